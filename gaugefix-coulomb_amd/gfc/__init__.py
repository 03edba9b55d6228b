"""Python binding of libgfc, the MI355X Coulomb gauge-fixing engine.

A thin ctypes layer over the C ABI declared in include/gfc.h.  Field arrays
are numpy complex128 arrays in the reference's memory image
(include/fields.h:33-43 of jesuel-marques/GaugeFix-Coulomb):

    U.shape == (V, 4, 3, 3)   site-major, mu = 0:T 1:Z 2:Y 3:X
    G.shape == (V, 3, 3)

with site = ((t*ns + z)*ns + y)*ns + x.  Every device computation runs in the
HIP library; importing this module fails loudly when libgfc.so is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.normpath(os.path.join(_HERE, "..", "lib"))
LIB_PATH = os.path.join(LIB_DIR, "libgfc.so")
CLI_PATH = os.path.join(LIB_DIR, "gauge_fix_coulomb")

OK = 0
ENOTCONVERGED = -1
EPARAMS = -2
EDEVICE = -3
EGEOMETRY = -4
ENOMEM = -5
EINVAL = -6
EPROJECT = -7

PROXY_E2 = 0
PROXY_THETA = 1
PREC_FP64 = 0
PREC_MIXED = 1


class GfcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gfc error {code}: {msg}")
        self.code = code


class _Params(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double),
        ("gfix_proxy", C.c_int),
        ("max_sweeps_to_fix", C.c_uint),
        ("estimate_sweeps_to_gf_progress", C.c_uint),
        ("sweeps_to_reunitarization", C.c_uint),
        ("error", C.c_int),
        ("omega_OR", C.c_float),
        ("hits", C.c_uint),
    ]


class _Geom(C.Structure):
    _fields_ = [("ns", C.c_int), ("nt", C.c_int)]


class _Options(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("mixed_switch", C.c_double),
        ("mixed_probe_every", C.c_int),
        ("mixed_defer", C.c_int),
        ("verbose", C.c_int),
        ("time_sweeps", C.c_int),
        ("retire_slices", C.c_int),
        ("slice_group", C.c_int),
        ("slice_lanes", C.c_int),
    ]


class _Report(C.Structure):
    _fields_ = [
        ("sweeps", C.c_int),
        ("sweeps_fp32", C.c_int),
        ("residue_evals", C.c_int),
        ("residue_initial", C.c_double),
        ("residue_final", C.c_double),
        ("seconds", C.c_double),
        ("seconds_sweeps", C.c_double),
        ("slice_sweeps", C.c_double),
    ]


class _Completion(C.Structure):
    _fields_ = [
        ("id", C.c_int64),
        ("tag", C.c_uint64),
        ("status", C.c_int),
        ("sweeps", C.c_int),
        ("sweeps_fp32", C.c_int),
        ("ns", C.c_int),
        ("nt", C.c_int),
        ("residue_initial", C.c_double),
        ("residue_final", C.c_double),
        ("F", C.c_double),
        ("t_submit", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
    ]


class _ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64),
        ("chunks", C.c_int64), ("graphs_built", C.c_int64), ("graph_replays", C.c_int64),
        ("residue_batches", C.c_int64), ("site_updates", C.c_int64),
        ("queued", C.c_int), ("active", C.c_int),
        ("bytes_used", C.c_int64), ("bytes_budget", C.c_int64),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libgfc.so not found at {LIB_PATH}: build it with "
            "`make -C gaugefix-coulomb_amd/csrc` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P, D, I, I64 = C.c_void_p, C.POINTER(C.c_double), C.c_int, C.c_int64
    sig = {
        "gfc_params_default": (None, [C.POINTER(_Params)]),
        "gfc_params_from_file": (I, [C.c_char_p, C.POINTER(_Params)]),
        "gfc_params_valid": (I, [C.POINTER(_Params)]),
        "gfc_params_print": (None, [C.POINTER(_Params)]),
        "gfc_next_check": (I, [C.c_uint, C.c_double, C.POINTER(_Params)]),
        "gfc_geom_check": (I, [C.POINTER(_Geom)]),
        "gfc_geom_volume": (I64, [C.POINTER(_Geom)]),
        "gfc_load_config": (I, [C.c_char_p, C.POINTER(_Geom), I, I, D]),
        "gfc_write_config": (I, [C.c_char_p, C.POINTER(_Geom), I, I, D]),
        "gfc_write_gauge_transf": (I, [C.c_char_p, C.POINTER(_Geom), D]),
        "gfc_load_gauge_transf": (I, [C.c_char_p, C.POINTER(_Geom), D]),
        "gfc_write_sweeps_record": (I, [C.c_char_p, C.POINTER(_Geom), C.c_char_p, I]),
        "gfc_generate_smooth": (I, [C.POINTER(_Geom), C.c_uint64, C.c_double, D]),
        "gfc_field_to_fp32": (None, [D, C.POINTER(C.c_float), I64]),
        "gfc_options_default": (None, [C.POINTER(_Options)]),
        "gfc_last_error": (C.c_char_p, []),
        "gfc_version": (C.c_char_p, []),
        "gfc_ctx_create": (I, [I, C.POINTER(P)]),
        "gfc_ctx_destroy": (None, [P]),
        "gfc_ctx_copy_probe": (I, [P, I64, I, D]),
        "gfc_ctx_read_probe": (I, [P, I64, I, D]),
        "gfc_job_footprint": (I64, [C.POINTER(_Geom), I]),
        "gfc_job_create": (I, [P, C.POINTER(_Geom), I, C.POINTER(P)]),
        "gfc_job_destroy": (None, [P]),
        "gfc_job_upload_config": (I, [P, P, I]),
        "gfc_job_reunitarize": (I, [P]),
        "gfc_job_observables": (I, [P, D, D, D]),
        "gfc_job_plaquette": (I, [P, D, D]),
        "gfc_job_fix": (I, [P, C.POINTER(_Params), C.POINTER(_Options), C.POINTER(_Report)]),
        "gfc_job_fix_trace": (I, [P, C.POINTER(_Params), C.POINTER(_Options),
                                  C.POINTER(_Report), D, I]),
        "gfc_job_download_config": (I, [P, D]),
        "gfc_job_download_gauge": (I, [P, D]),
        "gfc_job_get_link": (I, [P, I64, I, D]),
        "gfc_job_save_config": (I, [P]),
        "gfc_job_restore_config": (I, [P]),
        "gfc_fix_coulomb_or": (I, [P, C.POINTER(_Geom), C.POINTER(_Params),
                                   C.POINTER(_Options), D, D, C.POINTER(_Report)]),
        "gfc_reunitarize_field": (I, [P, D, I64]),
        "gfc_config_observables": (I, [P, C.POINTER(_Geom), D, D, D, D, D, D]),
        "gfc_config_observables_noalloc": (I, [P, C.POINTER(_Geom), C.c_void_p, D, D, D, D,
                                               D]),
        "gfc_job_generate_smooth": (I, [P, C.c_uint64, C.c_double]),
        "gfc_job_load_config": (I, [P, C.c_char_p, I, I]),
        "gfc_job_write_gauge_transf": (I, [P, C.c_char_p]),
        "gfc_job_write_config": (I, [P, C.c_char_p, I, I]),
        "gfc_server_create": (I, [P, I64, I, I, C.POINTER(P)]),
        "gfc_server_destroy": (None, [P]),
        "gfc_server_clock": (C.c_double, [P]),
        "gfc_server_submit": (I64, [P, C.POINTER(_Geom), C.c_uint64, C.c_double,
                                    C.POINTER(_Params), C.POINTER(_Options), C.c_uint64]),
        "gfc_server_submit_config": (I64, [P, C.POINTER(_Geom), C.c_void_p, I,
                                           C.POINTER(_Params), C.POINTER(_Options),
                                           C.c_uint64]),
        "gfc_server_poll": (I, [P, C.POINTER(_Completion), I]),
        "gfc_server_pending": (I, [P]),
        "gfc_server_get_stats": (I, [P, C.POINTER(_ServerStats)]),
        "gfc_server_result_gauge": (I, [P, I64, D]),
        "gfc_server_release": (I, [P, I64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib, sorted(sig)


lib, EXPORTED = _load()


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc: int) -> int:
    if rc < 0:
        raise GfcError(rc, lib.gfc_last_error().decode())
    return rc


@dataclass
class Geometry:
    ns: int
    nt: int

    @property
    def volume(self) -> int:
        return self.ns ** 3 * self.nt

    def _c(self) -> _Geom:
        return _Geom(self.ns, self.nt)


@dataclass
class Params:
    """Mirrors ORGaugeFixingParameters (include/gauge_fixing.h:31-54)."""
    tolerance: float = 1e-16
    gfix_proxy: int = PROXY_E2
    max_sweeps_to_fix: int = 100000
    estimate_sweeps_to_gf_progress: int = 1000
    sweeps_to_reunitarization: int = 250
    error: int = 0
    omega_OR: float = 1.95
    hits: int = 2

    def _c(self) -> _Params:
        return _Params(self.tolerance, self.gfix_proxy, self.max_sweeps_to_fix,
                       self.estimate_sweeps_to_gf_progress, self.sweeps_to_reunitarization,
                       self.error, self.omega_OR, self.hits)

    @staticmethod
    def _from_c(p: _Params) -> "Params":
        return Params(p.tolerance, p.gfix_proxy, p.max_sweeps_to_fix,
                      p.estimate_sweeps_to_gf_progress, p.sweeps_to_reunitarization, p.error,
                      p.omega_OR, p.hits)

    @staticmethod
    def default() -> "Params":
        p = _Params()
        lib.gfc_params_default(C.byref(p))
        return Params._from_c(p)

    @staticmethod
    def from_file(path: str) -> "Params":
        """initORGaugeFixingParameters: defaults overridden by the file; `error` set
        when the result is invalid (an unreadable file keeps the defaults)."""
        p = _Params()
        lib.gfc_params_from_file(path.encode(), C.byref(p))
        return Params._from_c(p)

    def valid(self) -> bool:
        p = self._c()
        return bool(lib.gfc_params_valid(C.byref(p)))

    def next_check(self, sweep: int, residue: float) -> int:
        p = self._c()
        return lib.gfc_next_check(sweep, residue, C.byref(p))


@dataclass
class Options:
    precision: int = PREC_FP64
    mixed_switch: float = 1e-9
    mixed_probe_every: int = 50
    mixed_defer: int = 1
    verbose: int = 0
    time_sweeps: int = 0
    retire_slices: int = 0
    slice_group: int = 0
    slice_lanes: int = 0

    def _c(self) -> _Options:
        return _Options(self.precision, self.mixed_switch, self.mixed_probe_every,
                        self.mixed_defer, self.verbose, self.time_sweeps, self.retire_slices,
                        self.slice_group, self.slice_lanes)


@dataclass
class Report:
    sweeps: int
    sweeps_fp32: int
    residue_evals: int
    residue_initial: float
    residue_final: float
    seconds: float
    seconds_sweeps: float
    slice_sweeps: float


def _report(r: _Report) -> Report:
    return Report(r.sweeps, r.sweeps_fp32, r.residue_evals, r.residue_initial,
                  r.residue_final, r.seconds, r.seconds_sweeps, r.slice_sweeps)


def new_config(geom: Geometry) -> np.ndarray:
    return np.zeros((geom.volume, 4, 3, 3), dtype=np.complex128)


def generate_smooth(geom: Geometry, seed: int, eps: float) -> np.ndarray:
    """Seeded synthetic configuration g(n) exp(i eps H_mu(n)) g(n+mu)^dagger."""
    U = new_config(geom)
    g = geom._c()
    _check(lib.gfc_generate_smooth(C.byref(g), seed, eps, _dptr(U)))
    return U


def to_fp32(U: np.ndarray) -> np.ndarray:
    out = np.empty(U.shape, dtype=np.complex64)
    lib.gfc_field_to_fp32(_dptr(U), out.ctypes.data_as(C.POINTER(C.c_float)),
                          U.size // 9)
    return out


def load_config(path: str, geom: Geometry, precision: int = 8, byte_swap: bool = True):
    """loadConfig: returns (rc, U) with the reference's return codes."""
    U = new_config(geom)
    g = geom._c()
    rc = lib.gfc_load_config(path.encode(), C.byref(g), precision, int(byte_swap), _dptr(U))
    return rc, U


def write_config(path: str, geom: Geometry, U: np.ndarray, precision: int = 8,
                 byte_swap: bool = False) -> int:
    g = geom._c()
    U = np.ascontiguousarray(U, dtype=np.complex128)
    return lib.gfc_write_config(path.encode(), C.byref(g), precision, int(byte_swap), _dptr(U))


def write_gauge_transf(path: str, geom: Geometry, G: np.ndarray) -> int:
    g = geom._c()
    G = np.ascontiguousarray(G, dtype=np.complex128)
    return lib.gfc_write_gauge_transf(path.encode(), C.byref(g), _dptr(G))


def load_gauge_transf(path: str, geom: Geometry):
    G = np.zeros((geom.volume, 3, 3), dtype=np.complex128)
    g = geom._c()
    rc = lib.gfc_load_gauge_transf(path.encode(), C.byref(g), _dptr(G))
    return rc, G


def write_sweeps_record(directory, geom: Geometry, identifier: str, sweeps: int) -> int:
    g = geom._c()
    d = directory.encode() if directory else None
    return lib.gfc_write_sweeps_record(d, C.byref(g), identifier.encode(), sweeps)


def footprint(geom: Geometry, precision: int = PREC_FP64) -> int:
    g = geom._c()
    return lib.gfc_job_footprint(C.byref(g), precision)


class Context:
    """One HIP device and stream."""

    def __init__(self, device: int = 0):
        self._p = C.c_void_p()
        # jobs and servers on this context: closed before it (their C objects
        # use the context's streams when they are destroyed)
        self._children = weakref.WeakSet()
        _check(lib.gfc_ctx_create(device, C.byref(self._p)))

    def copy_probe(self, nbytes: int, reps: int = 5) -> float:
        """Mean seconds of a device copy of nbytes (streaming 16-byte copy kernel)."""
        s = C.c_double()
        _check(lib.gfc_ctx_copy_probe(self._p, nbytes, reps, C.byref(s)))
        return s.value

    def read_probe(self, nbytes: int, reps: int = 20) -> float:
        """Seconds of one read pass over a device buffer of nbytes (Infinity-Cache
        resident up to 256 MiB, HBM above), without the launch's ramp-up and
        drain: many passes per launch, differenced over reps passes."""
        s = C.c_double()
        _check(lib.gfc_ctx_read_probe(self._p, nbytes, reps, C.byref(s)))
        return s.value

    def close(self):
        if self._p:
            for child in list(self._children):
                child.close()
            lib.gfc_ctx_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Job:
    """One lattice resident in HBM (links U, gauge transformation G)."""

    def __init__(self, ctx: Context, geom: Geometry, precision: int = PREC_FP64):
        self.ctx = ctx
        self.geom = geom
        self._p = C.c_void_p()
        g = geom._c()
        _check(lib.gfc_job_create(ctx._p, C.byref(g), precision, C.byref(self._p)))
        ctx._children.add(self)

    def close(self):
        if self._p:
            lib.gfc_job_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload_config(self, U: np.ndarray):
        if U.dtype == np.complex64:
            prec = 4
        else:
            U = np.ascontiguousarray(U, dtype=np.complex128)
            prec = 8
        assert U.shape == (self.geom.volume, 4, 3, 3)
        U = np.ascontiguousarray(U)
        _check(lib.gfc_job_upload_config(self._p, U.ctypes.data_as(C.c_void_p), prec))

    def reunitarize(self) -> int:
        return _check(lib.gfc_job_reunitarize(self._p))

    def observables(self):
        e2, th, f = C.c_double(), C.c_double(), C.c_double()
        _check(lib.gfc_job_observables(self._p, C.byref(e2), C.byref(th), C.byref(f)))
        return e2.value, th.value, f.value

    def plaquette(self):
        s, t = C.c_double(), C.c_double()
        _check(lib.gfc_job_plaquette(self._p, C.byref(s), C.byref(t)))
        return s.value, t.value

    def fix(self, params: Params, options: Options | None = None, trace_len: int = 0):
        """gaugefixOverrelaxation on the resident job.  Returns (rc, report[, trace]);
        rc is the sweep count or a negative code (-1 not converged, -2 bad
        parameters); device errors raise."""
        p = params._c()
        o = (options or Options())._c()
        r = _Report()
        if trace_len:
            tr = np.full(trace_len, np.nan)
            rc = lib.gfc_job_fix_trace(self._p, C.byref(p), C.byref(o), C.byref(r),
                                       _dptr(tr), trace_len)
        else:
            rc = lib.gfc_job_fix(self._p, C.byref(p), C.byref(o), C.byref(r))
        if rc < -2:
            _check(rc)
        if trace_len:
            return rc, _report(r), tr
        return rc, _report(r)

    def download_config(self) -> np.ndarray:
        U = new_config(self.geom)
        _check(lib.gfc_job_download_config(self._p, _dptr(U)))
        return U

    def download_gauge(self) -> np.ndarray:
        G = np.zeros((self.geom.volume, 3, 3), dtype=np.complex128)
        _check(lib.gfc_job_download_gauge(self._p, _dptr(G)))
        return G

    def generate_smooth(self, seed: int, eps: float):
        """Synthesise the smooth configuration directly in HBM (G <- 1)."""
        _check(lib.gfc_job_generate_smooth(self._p, seed, eps))

    def load_config(self, path: str, in_precision: int = 8, byte_swap: bool = True) -> int:
        """Stream a configuration file into the job (loadConfig rules); returns
        the loader's code: 0, -1 open, -2 short, -3 trailing bytes, -4 precision."""
        rc = lib.gfc_job_load_config(self._p, os.fsencode(path), in_precision, int(byte_swap))
        if rc < -4:
            _check(rc)
        return rc

    def write_gauge_transf(self, path: str) -> int:
        rc = lib.gfc_job_write_gauge_transf(self._p, os.fsencode(path))
        if rc < -4:
            _check(rc)
        return rc

    def write_config(self, path: str, out_precision: int = 8, byte_swap: bool = True) -> int:
        rc = lib.gfc_job_write_config(self._p, os.fsencode(path), out_precision, int(byte_swap))
        if rc < -4:
            _check(rc)
        return rc

    def save_config(self):
        _check(lib.gfc_job_save_config(self._p))

    def restore_config(self):
        _check(lib.gfc_job_restore_config(self._p))

    def get_link(self, site: int, mu: int) -> np.ndarray:
        out = np.zeros((3, 3), dtype=np.complex128)
        _check(lib.gfc_job_get_link(self._p, site, mu, _dptr(out)))
        return out


def fix_coulomb_or(ctx: Context, geom: Geometry, params: Params, U: np.ndarray,
                   G: np.ndarray, options: Options | None = None):
    """The drop-in gaugefixOverrelaxation(U, G, params): U, G updated in place."""
    assert U.dtype == np.complex128 and U.flags.c_contiguous
    assert G.dtype == np.complex128 and G.flags.c_contiguous
    g, p = geom._c(), params._c()
    o = (options or Options())._c()
    r = _Report()
    rc = lib.gfc_fix_coulomb_or(ctx._p, C.byref(g), C.byref(p), C.byref(o), _dptr(U),
                                _dptr(G), C.byref(r))
    if rc < -2:
        _check(rc)
    return rc, _report(r)


def reunitarize_field(ctx: Context, field: np.ndarray) -> int:
    assert field.dtype == np.complex128 and field.flags.c_contiguous
    return lib.gfc_reunitarize_field(ctx._p, _dptr(field), field.size // 9)


def config_observables(ctx: Context, geom: Geometry, U: np.ndarray):
    """(e2, theta, F, plaquette spatial, plaquette temporal) of a host config."""
    g = geom._c()
    vals = [C.c_double() for _ in range(5)]
    U = np.ascontiguousarray(U, dtype=np.complex128)
    _check(lib.gfc_config_observables(ctx._p, C.byref(g), _dptr(U),
                                      *[C.byref(v) for v in vals]))
    return tuple(v.value for v in vals)


def config_observables_noalloc(ctx: Context, geom: Geometry, U_dev: int):
    """config_observables of a configuration already in device memory (U_dev:
    a device address of 4V x 18 doubles); allocates nothing."""
    g = geom._c()
    vals = [C.c_double() for _ in range(5)]
    _check(lib.gfc_config_observables_noalloc(ctx._p, C.byref(g), C.c_void_p(U_dev),
                                              *[C.byref(v) for v in vals]))
    return tuple(v.value for v in vals)


@dataclass
class Completion:
    id: int
    tag: int
    status: int
    sweeps: int
    sweeps_fp32: int
    ns: int
    nt: int
    residue_initial: float
    residue_final: float
    F: float
    t_submit: float
    t_admit: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class Server:
    """Continuous-batching scheduler (gfc_server_*): many lattices resident in
    HBM, admitted by footprint, their half-sweeps packed into shared launches."""

    def __init__(self, ctx: Context, mem_budget: int = 0, max_jobs: int = 64,
                 keep_results: bool = False):
        self.ctx = ctx
        self._p = C.c_void_p()
        _check(lib.gfc_server_create(ctx._p, mem_budget, max_jobs, int(keep_results),
                                     C.byref(self._p)))
        ctx._children.add(self)
        self._buf = (_Completion * 64)()

    def close(self):
        if self._p:
            lib.gfc_server_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def clock(self) -> float:
        return lib.gfc_server_clock(self._p)

    def submit(self, geom: Geometry, seed: int, eps: float, params: Params,
               options: Options | None = None, tag: int = 0) -> int:
        g, p = geom._c(), params._c()
        o = (options or Options())._c()
        return _check(lib.gfc_server_submit(self._p, C.byref(g), seed, eps, C.byref(p),
                                            C.byref(o), tag))

    def submit_config(self, geom: Geometry, U: np.ndarray, params: Params,
                      options: Options | None = None, tag: int = 0) -> int:
        """Queues the caller's configuration U (complex128 or complex64, shape
        (V, 4, 3, 3)); the server keeps its own copy."""
        if U.dtype == np.complex64:
            prec = 4
        else:
            U = np.asarray(U, dtype=np.complex128)
            prec = 8
        assert U.shape == (geom.volume, 4, 3, 3)
        U = np.ascontiguousarray(U)
        g, p = geom._c(), params._c()
        o = (options or Options())._c()
        return _check(lib.gfc_server_submit_config(self._p, C.byref(g),
                                                   U.ctypes.data_as(C.c_void_p), prec,
                                                   C.byref(p), C.byref(o), tag))

    def poll(self) -> list:
        n = _check(lib.gfc_server_poll(self._p, self._buf, len(self._buf)))
        return [Completion(*(getattr(self._buf[i], f) for f, _ in _Completion._fields_))
                for i in range(n)]

    def pending(self) -> int:
        return lib.gfc_server_pending(self._p)

    def stats(self) -> dict:
        s = _ServerStats()
        _check(lib.gfc_server_get_stats(self._p, C.byref(s)))
        return {f: getattr(s, f) for f, _ in _ServerStats._fields_}

    def result_gauge(self, rid: int, geom: Geometry) -> np.ndarray:
        G = np.zeros((geom.volume, 3, 3), dtype=np.complex128)
        _check(lib.gfc_server_result_gauge(self._p, rid, _dptr(G)))
        return G

    def release(self, rid: int):
        _check(lib.gfc_server_release(self._p, rid))

    def drain(self, timeout: float = 600.0) -> list:
        """Polls until every submitted request has completed."""
        out = []
        t0 = self.clock()
        while self.pending() > 0:
            out += self.poll()
            if self.clock() - t0 > timeout:
                raise TimeoutError("server did not drain")
        return out
