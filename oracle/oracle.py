"""ctypes wrapper of liboracle.so, the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product path (gfc) never does.  Fields use the same numpy
layout as gfc: U (V, 4, 3, 3) complex128, G (V, 3, 3) complex128.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")


class _Geom(C.Structure):
    _fields_ = [("ns", C.c_int), ("nt", C.c_int)]


class _Params(C.Structure):
    _fields_ = [
        ("tolerance", C.c_double),
        ("proxy", C.c_int),
        ("max_sweeps", C.c_uint),
        ("estimate", C.c_uint),
        ("reunit_every", C.c_uint),
        ("omega", C.c_float),
        ("hits", C.c_uint),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    D, P = C.POINTER(C.c_double), C.c_void_p
    G, PR = C.POINTER(_Geom), C.POINTER(_Params)
    sig = {
        "orc_prod": (None, [D, D, D]),
        "orc_inverse": (None, [D, D]),  # complex return ignored
        "orc_cm_step": (None, [D, C.c_int]),
        "orc_project_su3": (C.c_int, [D]),
        "orc_power_binomial": (C.c_int, [D, C.c_double, D]),
        "orc_decompose": (None, [D, D]),
        "orc_la_update": (None, [D, C.c_uint, D]),
        "orc_half_sweep": (None, [D, D, G, C.c_int, C.c_double, C.c_uint]),
        "orc_e2": (C.c_double, [D, G]),
        "orc_theta": (C.c_double, [D, G]),
        "orc_functional": (C.c_double, [D, G]),
        "orc_plaquette": (C.c_double, [D, G, C.c_int]),
        "orc_reunitarize": (C.c_int, [D, C.c_int64]),
        "orc_set_identity": (None, [D, C.c_int64]),
        "orc_apply_gauge": (None, [D, D, G]),
        "orc_next_check": (C.c_int, [C.c_uint, C.c_double, PR]),
        "orc_params_valid": (C.c_int, [PR]),
        "orc_gaugefix": (C.c_int, [D, D, G, PR, D, C.c_int, C.c_int]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    # the complex-returning functions: read det through a (double, double) struct
    class _Cplx(C.Structure):
        _fields_ = [("re", C.c_double), ("im", C.c_double)]
    lib.orc_det.restype = _Cplx
    lib.orc_det.argtypes = [D]
    lib.orc_inverse.restype = _Cplx
    lib.orc_inverse.argtypes = [D, D]
    return lib


lib = _load()


def _d(a: np.ndarray):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def geom(ns: int, nt: int) -> _Geom:
    return _Geom(ns, nt)


def params(tolerance=1e-16, proxy=0, max_sweeps=100000, estimate=1000, reunit_every=250,
           omega=1.95, hits=2) -> _Params:
    return _Params(tolerance, proxy, max_sweeps, estimate, reunit_every, omega, hits)


def det(A):
    r = lib.orc_det(_d(np.ascontiguousarray(A, np.complex128)))
    return complex(r.re, r.im)


def inverse(A):
    A = np.ascontiguousarray(A, np.complex128)
    out = np.zeros((3, 3), np.complex128)
    lib.orc_inverse(_d(A), _d(out))
    return out


def prod(A, B):
    out = np.zeros((3, 3), np.complex128)
    lib.orc_prod(_d(np.ascontiguousarray(A, np.complex128)),
                 _d(np.ascontiguousarray(B, np.complex128)), _d(out))
    return out


def project_su3(A):
    X = np.array(A, dtype=np.complex128, order="C")
    st = lib.orc_project_su3(_d(X))
    return st, X


def power_binomial(A, omega):
    A = np.ascontiguousarray(A, np.complex128)
    out = np.zeros((3, 3), np.complex128)
    st = lib.orc_power_binomial(_d(A), omega, _d(out))
    return st, out


def cm_step(A, sub):
    X = np.array(A, dtype=np.complex128, order="C")
    lib.orc_cm_step(_d(X), sub)
    return X


def la_update(w, hits):
    W = np.array(w, dtype=np.complex128, order="C")
    g = np.zeros((3, 3), np.complex128)
    lib.orc_la_update(_d(W), hits, _d(g))
    return g


def decompose(A):
    A = np.ascontiguousarray(A, np.complex128)
    c = np.zeros(9)
    lib.orc_decompose(_d(A), c.ctypes.data_as(C.POINTER(C.c_double)))
    return c


def e2(U, g):
    return lib.orc_e2(_d(U), C.byref(g))


def theta(U, g):
    return lib.orc_theta(_d(U), C.byref(g))


def functional(U, g):
    return lib.orc_functional(_d(U), C.byref(g))


def plaquette(U, g, which):
    return lib.orc_plaquette(_d(U), C.byref(g), which)


def reunitarize(field):
    return lib.orc_reunitarize(_d(field), field.size // 9)


def apply_gauge(U, G, g):
    lib.orc_apply_gauge(_d(U), _d(G), C.byref(g))


def half_sweep(U, G, g, sweep, omega, hits):
    lib.orc_half_sweep(_d(U), _d(G), C.byref(g), sweep, omega, hits)


def next_check(sweep, res, p):
    return lib.orc_next_check(sweep, res, C.byref(p))


def gaugefix(U, G, g, p, trace_len=0, verbose=0):
    """orc_gaugefix: returns (rc, trace) with trace[k-1] = proxy after sweep k."""
    tr = np.full(max(trace_len, 1), np.nan)
    rc = lib.orc_gaugefix(_d(U), _d(G), C.byref(g), C.byref(p),
                          tr.ctypes.data_as(C.POINTER(C.c_double)) if trace_len else None,
                          trace_len, verbose)
    return rc, tr[:trace_len]
