"""Benchmark: Coulomb gauge fixing of a 48^3x96 lattice on one MI355X.

A step is one full gauge fix to e2 < 1e-16 of a synthetic smooth
configuration (eps 0.35, stored as FP32 SCIDAC input like the mixed-precision
workload, converted to FP64 and reunitarized on the device), under the
reference's own convergence-check schedule (estimate_sweeps_to_gf_progress
1000), omega 1.95, hits 2, reunitarization every 10 sweeps.  Inputs are
resident in HBM before the timed region (a device-side copy restores the
unfixed configuration at the start of every step).

value        site updates per second of the FP64 drop-in path
               (half-sweeps x V/2 / wall time of the step)
fp64_seconds wall time to e2 < 1e-16, FP64 path
mixed_*      the FP32 -> FP64 path on the same input and schedule
roofline     the kernel of the timed fix, k_sweep<double> in cache-resident
             timeslice waves: the factored update's algorithmic bytes (768 B
             per site update: the 6 spatial u0 links in five stored complex
             numbers, G(n) read and written, one neighbour G per update) per lattice
             half-sweep over the half-sweep's duration measured with HIP
             events on the job's stream (fork to join of the waves), and the
             copy-calibrated L2-miss traffic of the same launches from the
             committed rocprofv3 counters (profiles/sweep_fix_48x96_pmc.txt)
cpu_baseline the reference's OpenMP CPU path (compiled from its sources into
             oracle/_ref) timed on this host on a bounded sample of the same
             workload: a few half-sweeps of the same lattice
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gaugefix-coulomb_amd"))

# The serving leg's server runs seven streams (one per size lane, three more
# for the large lane's timeslice waves).  With HIP's default of four hardware
# queues some of them share a queue, and kernels on one queue run in order: a
# lone 48^3x96 request then takes 0.92 s instead of 0.72 s.  Eight queues
# (read when the HIP runtime starts, so set before any HIP call) give each
# stream its own; the single-lattice fixes use four streams either way.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

# bytes per site update credited to the reference's update: 6 spatial links
# and G(n), full 3x3 matrices read and written once (14 x 144 B); only used
# for reference-equivalent rates
REF_BYTES_PER_UPDATE_FP64 = 2016
REF_BYTES_PER_UPDATE_FP32 = 1008
# the factored update's algorithmic bytes in the engine's storage: the 6
# spatial u0 links (FP64: five complex numbers, 80 B; FP32: two rows, 48 B),
# G(n) read and written and one neighbour G (two rows: 96 B in FP64, 48 B in
# FP32; each of the 6 neighbour G read by an update is shared by 6 updates)
ALG_BYTES_PER_UPDATE_FP64 = 6 * 80 + 96 + 96 + 96
ALG_BYTES_PER_UPDATE_FP32 = 6 * 48 + 48 + 48 + 48
HBM_PEAK_GBS = 8000.0
# VALU issue peak: 256 CUs x 4 SIMDs, a wave64 FP64 (or unpacked FP32)
# instruction every 4 cycles at 2.4 GHz
VALU_WAVE_INSTR_PEAK = 256 * 4 * 2.4e9 / 4


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pmc_value(name: str, key: str):
    """A figure from a committed rocprofv3 PMC summary (profiles/<name>)."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    m = re.search(key + r":\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def pmc_traffic(ns: int, nt: int, tag: str):
    """L2-miss bytes per whole-lattice launch of a sweep kernel from the
    committed rocprofv3 PMC summary for this lattice (FETCH_SIZE + WRITE_SIZE,
    copy-calibrated)."""
    return pmc_value(f"sweep_{ns}x{nt}_pmc.txt", tag + "_hbm_bytes_per_launch")


class _CStdoutToStderr:
    """Sends file descriptor 1 to stderr while the CPU baseline runs: the
    reference library prints its "Sweeps:" log with printf, and stdout must
    carry only the JSON record."""

    def __enter__(self):
        sys.stdout.flush()
        self.libc = C.CDLL(None)
        self.saved = os.dup(1)
        os.dup2(2, 1)

    def __exit__(self, *exc):
        self.libc.fflush(None)
        sys.stdout.flush()
        os.dup2(self.saved, 1)
        os.close(self.saved)
        return False


def _host_has_avx512() -> bool:
    """The ISA extensions -march=skylake-avx512 code uses (F, CD, BW, DQ, VL)."""
    try:
        flags = set()
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                flags = set(line.split(":", 1)[1].split())
                break
        return {"avx512f", "avx512cd", "avx512bw", "avx512dq", "avx512vl"} <= flags
    except OSError:
        return False


def cpu_baseline(U: np.ndarray, ns: int, nt: int, seconds: float):
    """Times the reference's gaugefixOverrelaxation (oracle/_ref build) on n
    half-sweeps of the same lattice; falls back to the C oracle port.  The
    reference is the build with its own Makefile flags (-O4
    -march=skylake-avx512, Makefile:8) where the host has AVX-512, else the
    portable -O2 build."""
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    vh = ns ** 3 * nt // 2
    perf_so = os.path.join(ROOT, "oracle", "_ref", "libref_le64_perf.so")
    if os.path.exists(perf_so) and _host_has_avx512():
        ref_so, build = perf_so, "-std=c99 -O4 -march=skylake-avx512 -mtune=skylake-avx512 -fopenmp"
    else:
        ref_so, build = (os.path.join(ROOT, "oracle", "_ref", "libref_le64_mt.so"),
                         "-std=c99 -O2 -fopenmp (host without AVX-512)")
    Uw = U.copy()
    G = np.zeros((U.shape[0], 3, 3), np.complex128)
    dp = lambda a: a.ctypes.data_as(C.c_void_p)
    if os.path.exists(ref_so):
        lib = C.CDLL(ref_so)

        class Generic(C.Structure):
            _fields_ = [("tolerance", C.c_double), ("gfix_proxy", C.c_void_p),
                        ("max_sweeps_to_fix", C.c_uint), ("estimate", C.c_uint),
                        ("reunit", C.c_uint), ("error", C.c_bool)]

        class ORParams(C.Structure):
            _fields_ = [("generic_gf", Generic), ("omega_OR", C.c_float), ("hits", C.c_uint)]

        lib.initGeometry.argtypes = [C.c_short, C.c_short]
        lib.gaugefixOverrelaxation.argtypes = [C.c_void_p, C.c_void_p, ORParams]
        lib.gaugefixOverrelaxation.restype = C.c_int
        lib.reunitarizeField.argtypes = [C.c_void_p, C.c_uint]
        lib.initGeometry(ns, nt)
        lib.reunitarizeField(dp(Uw), 4 * ns ** 3 * nt)
        e2fn = C.cast(lib.calculate_e2, C.c_void_p).value

        def run(n):
            p = ORParams(Generic(1e-300, e2fn, n, 100000000, 10, False), 1.95, 2)
            t = time.perf_counter()
            lib.gaugefixOverrelaxation(dp(Uw), dp(G), p)
            return time.perf_counter() - t
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as orc
        g = orc.geom(ns, nt)
        orc.reunitarize(Uw)

        def run(n):
            p = orc.params(tolerance=1e-300, estimate=100000000, max_sweeps=n, reunit_every=10)
            t = time.perf_counter()
            orc.gaugefix(Uw, G, g, p)
            return time.perf_counter() - t
        kind = "port"
        build = "oracle/gfc_oracle.c, -std=c99 -O2 -ffp-contract=off -fopenmp"
    with _CStdoutToStderr():
        t1 = run(1)  # setup (G <- 1, initial e2) + one half-sweep
        per = max(t1 * 0.5, 1e-3)
        n = max(2, min(40, int(seconds / per)))
        tn = run(1 + n)
    per_sweep = max(tn - t1, 1e-9) / n
    return {"value": vh / per_sweep, "unit": "site-updates/s", "cores": threads,
            "kind": kind,
            "build": build,
            "sample": f"{n} half-sweeps of {ns}^3x{nt} (FP64, omega 1.95, hits 2) after one "
                      f"warm-up call; time of the (1+{n})-sweep call minus the 1-sweep call"}


def profile_kernels(ctx, job, n: int):
    """n FP64 half-sweeps, n FP32 half-sweeps (one whole-lattice launch each)
    and 1 GiB copies, for counters and per-launch durations."""
    import gfc
    fixed = gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=100000000,
                       max_sweeps_to_fix=n, sweeps_to_reunitarization=10 ** 6)
    job.restore_config()
    rc, rep = job.fix(fixed, gfc.Options(precision=gfc.PREC_FP64, slice_group=-1))
    log(f"profile: {rep.sweeps} FP64 half-sweeps")
    job.restore_config()
    rc, rep = job.fix(fixed, gfc.Options(precision=gfc.PREC_MIXED, mixed_switch=0.0,
                                         mixed_probe_every=10 ** 6, slice_group=-1))
    log(f"profile: {rep.sweeps_fp32} FP32 half-sweeps")
    sec = ctx.copy_probe(2 ** 30, reps=3)
    log(f"profile: copied 1 GiB x 4 with k_copy16 (calibration kernel), "
        f"{2 * 2 ** 30 / sec / 1e9:.1f} GB/s read+write")


def profile_fix(ctx, job, n: int, fp32: bool = False):
    """n FP64 (or, with fp32, FP32: the mixed path's first phase, no hand-off)
    half-sweeps run as the timed fix runs them (default grouping:
    cache-resident timeslice waves on several streams, reunitarization folded
    in every 10 sweeps), then 1 GiB copies: counters for the dominant kernel of
    the timed region and their calibration."""
    import gfc
    fixed = gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=100000000,
                       max_sweeps_to_fix=n, sweeps_to_reunitarization=10)
    job.restore_config()
    if fp32:
        opt = gfc.Options(precision=gfc.PREC_MIXED, mixed_switch=0.0, mixed_probe_every=10 ** 6)
    else:
        opt = gfc.Options(precision=gfc.PREC_FP64)
    rc, rep = job.fix(fixed, opt)
    log(f"profile: {rep.sweeps} {'FP32' if fp32 else 'FP64'} half-sweeps "
        f"({rep.sweeps_fp32} in FP32) in cache-resident timeslice waves")
    sec = ctx.copy_probe(2 ** 30, reps=3)
    log(f"profile: copied 1 GiB x 4 with k_copy16 (calibration kernel), "
        f"{2 * 2 ** 30 / sec / 1e9:.1f} GB/s read+write")


def io_leg(job, geom, U):
    """Load of the FP32 little-endian input file (hot page cache): the device
    stream (gfc_job_load_config: parallel pread into pinned blocks, DMA,
    conversion on the device) against the host loader (gfc_load_config) plus
    the upload of the FP64 image."""
    import tempfile
    import gfc
    d = tempfile.mkdtemp(dir=os.environ.get("TMPDIR", "/tmp"))
    path = os.path.join(d, "cfg_fp32_le.bin")
    try:
        if gfc.write_config(path, geom, U, 4, False) != 0:
            return None
        nbytes = os.path.getsize(path)
        t = time.perf_counter()
        rc = job.load_config(path, 4, False)
        t_dev = time.perf_counter() - t
        assert rc == 0, rc
        Ud = job.download_config()
        t = time.perf_counter()
        rc, Uh = gfc.load_config(path, geom, 4, False)
        job.upload_config(Uh)
        t_host = time.perf_counter() - t
        same = bool(np.array_equal(Ud, Uh))
    finally:
        if os.path.exists(path):
            os.remove(path)
        os.rmdir(d)
    out = {"file": "FP32 little-endian, hot page cache", "bytes": nbytes,
           "device_stream_seconds": round(t_dev, 4),
           "device_stream_GBs": round(nbytes / t_dev / 1e9, 2),
           "host_load_plus_upload_seconds": round(t_host, 4),
           "same_result": same}
    log(f"io: {out}")
    return out


def lattice_64x128(ctx, ns: int, nt: int, eps: float, seed: int):
    """64^3x128 (33.5 M sites): FP64, mixed, and mixed with per-timeslice
    retirement on the reference schedule, plus the GB/s of the FP64 and FP32
    sweep kernels: as the fixes run them (half-sweeps in the default grouping)
    and in whole-lattice launches (HIP events).  The input is
    synthesised in HBM by the device generator (the host generator's
    construction), so no 19 GB host image is needed."""
    import gfc
    geom = gfc.Geometry(ns, nt)
    vh = geom.volume // 2
    params = gfc.Params(tolerance=1e-16, estimate_sweeps_to_gf_progress=1000,
                        sweeps_to_reunitarization=10, omega_OR=1.95, hits=2)
    out = {"lattice": f"{ns}^3x{nt}", "sites": geom.volume}
    with gfc.Job(ctx, geom, gfc.PREC_MIXED) as job:
        t = time.perf_counter()
        job.generate_smooth(seed, eps)
        job.reunitarize()
        job.save_config()
        log(f"{ns}^3x{nt}: synthesised in HBM in {time.perf_counter() - t:.2f}s")
        plaq0 = job.plaquette()
        runs = {"fp64": gfc.Options(precision=gfc.PREC_FP64),
                "mixed": gfc.Options(precision=gfc.PREC_MIXED),
                "mixed_retire": gfc.Options(precision=gfc.PREC_MIXED, retire_slices=1)}
        F64 = None
        for name, opt in runs.items():
            job.restore_config()
            t = time.perf_counter()
            rc, rep = job.fix(params, opt)
            dt = time.perf_counter() - t
            e2, th, F = job.observables()
            plaq = job.plaquette()
            F64 = F if F64 is None else F64
            out[name] = {"seconds": round(dt, 4), "sweeps": rc, "sweeps_fp32": rep.sweeps_fp32,
                         "slice_sweeps": round(rep.slice_sweeps, 2), "final_e2": e2,
                         "F_rel_diff_vs_fp64": abs(F - F64) / abs(F64),
                         "plaquette_drift": max(abs(plaq[0] - plaq0[0]),
                                                abs(plaq[1] - plaq0[1])),
                         "site_updates_per_s": rep.slice_sweeps * vh / dt}
            log(f"{ns}^3x{nt} {name}: {dt:.3f}s sweeps={rc} fp32={rep.sweeps_fp32} "
                f"slice-sweeps={rep.slice_sweeps:.1f} e2={e2:.3e}")
        out["mixed_speedup_vs_fp64"] = round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3)
        out["mixed_retire_speedup_vs_fp64"] = round(
            out["fp64"]["seconds"] / out["mixed_retire"]["seconds"], 3)
        # the fixes' own sweeps (default grouping, folded reunitarization): HIP
        # events around the runs of blocked sweeps, 200 half-sweeps; traffic from
        # the committed counters of the same sweeps
        fixed200 = gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=10 ** 8,
                              max_sweeps_to_fix=200, sweeps_to_reunitarization=10)
        for name, opt, bpu, prof in (
                ("fp64", gfc.Options(precision=gfc.PREC_FP64, time_sweeps=1),
                 ALG_BYTES_PER_UPDATE_FP64, f"sweep_fix_{ns}x{nt}_pmc.txt"),
                ("fp32", gfc.Options(precision=gfc.PREC_MIXED, mixed_switch=0.0,
                                     mixed_probe_every=10 ** 6, time_sweeps=1),
                 ALG_BYTES_PER_UPDATE_FP32, f"sweep_fix_fp32_{ns}x{nt}_pmc.txt")):
            job.restore_config()
            rc, rep = job.fix(fixed200, opt)
            ms = rep.seconds_sweeps / rep.sweeps * 1e3
            t_upd = pmc_value(prof, f"fix_{name}_traffic_bytes_per_update")
            out[f"fix_{name}_half_sweep_ms"] = round(ms, 4)
            out[f"fix_{name}_GBs_algorithmic"] = round(bpu * vh / (ms * 1e-3) / 1e9, 1)
            out[f"fix_{name}_traffic_bytes_per_update"] = t_upd
            if t_upd:
                out[f"fix_{name}_traffic_GBs"] = round(t_upd * vh / (ms * 1e-3) / 1e9, 1)
            log(f"{ns}^3x{nt} {name} fix sweeps: {ms:.4f} ms per half-sweep")
        fixed = gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=10 ** 8,
                           max_sweeps_to_fix=40, sweeps_to_reunitarization=10 ** 6)
        for name, opt, bpu in (("fp64", gfc.PREC_FP64, ALG_BYTES_PER_UPDATE_FP64),
                               ("fp32", gfc.PREC_MIXED, ALG_BYTES_PER_UPDATE_FP32)):
            job.restore_config()
            rc, rep = job.fix(fixed, gfc.Options(precision=opt, mixed_switch=0.0,
                                                 mixed_probe_every=10 ** 6, time_sweeps=1,
                                                 slice_group=-1))
            ms = rep.seconds_sweeps / rep.sweeps * 1e3
            gbs = bpu * vh / (ms * 1e-3) / 1e9
            traffic = pmc_traffic(ns, nt, name)
            out[f"{name}_sweep_kernel_ms"] = round(ms, 4)
            out[f"{name}_sweep_GBs_algorithmic"] = round(gbs, 1)
            out[f"{name}_sweep_frac_of_hbm_peak"] = round(gbs / HBM_PEAK_GBS, 4)
            out[f"{name}_sweep_hbm_bytes_per_launch"] = traffic
            if traffic:
                out[f"{name}_sweep_traffic_GBs"] = round(traffic / (ms * 1e-3) / 1e9, 1)
            log(f"{ns}^3x{nt} {name} sweep kernel: {ms:.4f} ms/launch, {gbs:.1f} GB/s algorithmic")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ns", type=int, default=48)
    ap.add_argument("--nt", type=int, default=96)
    ap.add_argument("--eps", type=float, default=0.35)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-io", action="store_true", help="skip the file-load leg")
    ap.add_argument("--no-64x128", action="store_true",
                    help="skip the 64^3x128 mixed/retirement leg")
    ap.add_argument("--serve", type=int, default=256, metavar="N",
                    help="requests in the serving stream (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--profile-kernels", type=int, default=0, metavar="N",
                    help="only run N FP64 and N FP32 half-sweeps plus a 1 GiB device copy "
                         "(calibration for rocprofv3 PMC counters), then exit")
    ap.add_argument("--profile-fix", type=int, default=0, metavar="N",
                    help="only run N FP64 half-sweeps the way the timed fix runs them, plus "
                         "a 1 GiB device copy (rocprofv3 counters of the timed kernels), "
                         "on an input synthesised in HBM, then exit")
    ap.add_argument("--profile-fp32", action="store_true",
                    help="with --profile-fix: FP32 half-sweeps (mixed path, no hand-off)")
    ap.add_argument("--device-input", action="store_true",
                    help="with --profile-kernels: synthesise the input in HBM")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the engine fixes a lattice on one MI355X "
              "(splitting one over several GPUs by timeslices is a design, not built); "
              "run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import gfc
    geom = gfc.Geometry(a.ns, a.nt)
    V = geom.volume
    vh = V // 2
    if a.profile_fix:
        ctx = gfc.Context(0)
        with gfc.Job(ctx, geom, gfc.PREC_MIXED if a.profile_fp32 else gfc.PREC_FP64) as job:
            job.generate_smooth(a.seed, a.eps)
            job.reunitarize()
            job.save_config()
            log(f"profile: {a.ns}^3x{a.nt} input synthesised in HBM")
            profile_fix(ctx, job, a.profile_fix, a.profile_fp32)
        ctx.close()
        return
    if a.profile_kernels and a.device_input:
        ctx = gfc.Context(0)
        with gfc.Job(ctx, geom, gfc.PREC_MIXED) as job:
            job.generate_smooth(a.seed, a.eps)
            job.reunitarize()
            job.save_config()
            profile_kernels(ctx, job, a.profile_kernels)
        ctx.close()
        return
    log(f"generating {a.ns}^3x{a.nt} eps={a.eps} seed={a.seed}")
    t = time.time()
    U = gfc.generate_smooth(geom, a.seed, a.eps)
    U32 = gfc.to_fp32(U)
    log(f"  generated in {time.time() - t:.1f}s")
    params = gfc.Params(tolerance=1e-16, estimate_sweeps_to_gf_progress=1000,
                        sweeps_to_reunitarization=10, omega_OR=1.95, hits=2)
    ctx = gfc.Context(0)
    job = gfc.Job(ctx, geom, gfc.PREC_MIXED)
    job.upload_config(U32)
    job.reunitarize()
    job.save_config()
    copy_s = ctx.copy_probe(2 ** 30, reps=5)
    copy_gbs = 2 * 2 ** 30 / copy_s / 1e9
    log(f"device copy: {copy_gbs:.1f} GB/s (read + write)")
    # read bandwidth of a working set resident in the Infinity Cache (the
    # cache budget of the sweeps' waves) and of one in HBM: what the sweeps'
    # L2-miss traffic runs against
    cache_read_gbs = max((mib << 20) / ctx.read_probe(mib << 20, reps=50) / 1e9
                         for mib in (128, 192, 256))
    hbm_read_gbs = (2 << 30) / ctx.read_probe(2 << 30, reps=10) / 1e9
    log(f"device read: {cache_read_gbs:.1f} GB/s from a 128-256 MiB set, "
        f"{hbm_read_gbs:.1f} GB/s from 2 GiB")
    if a.profile_kernels:
        profile_kernels(ctx, job, a.profile_kernels)
        return
    plaq0 = job.plaquette()
    fp64 = gfc.Options(precision=gfc.PREC_FP64)
    mixed = gfc.Options(precision=gfc.PREC_MIXED)

    for i in range(a.warmup):
        job.restore_config()
        rc, rep = job.fix(params, fp64)
        log(f"warmup fp64: sweeps={rc} {rep.seconds:.3f}s")
    # the timed steps' sweep kernels: HIP events on the job's stream bracket
    # every timeslice-blocked run of half-sweeps (fork of the waves to their
    # join), so seconds_sweeps / sweeps is the duration of one lattice
    # half-sweep as the timed region runs it
    fp64_timed = gfc.Options(precision=gfc.PREC_FP64, time_sweeps=1)
    sweeps, sweep_s = [], 0.0
    t0 = time.perf_counter()
    for i in range(a.steps):
        job.restore_config()
        rc, rep = job.fix(params, fp64_timed)
        assert rc > 0, rc
        sweeps.append(rc)
        sweep_s += rep.seconds_sweeps
    t1 = time.perf_counter()
    step_s = (t1 - t0) / a.steps
    value = sum(sweeps) * vh / (t1 - t0)
    e2_64, th64, F64 = job.observables()
    plaq64 = job.plaquette()
    log(f"fp64: {a.steps} steps, sweeps={sweeps}, {step_s:.3f}s/step, {value:.4e} upd/s, "
        f"e2={e2_64:.3e}")
    half_ms = sweep_s / sum(sweeps) * 1e3
    achieved = ALG_BYTES_PER_UPDATE_FP64 * vh / (half_ms * 1e-3) / 1e9
    # the committed counters are of the 48^3x96 fix
    fixprof = f"sweep_fix_{a.ns}x{a.nt}_pmc.txt"
    traffic_upd = pmc_value(fixprof, "fix_fp64_traffic_bytes_per_update")
    valu_upd = pmc_value(fixprof, "fix_fp64_valu_instructions_per_update")
    log(f"fp64 fix sweeps: {half_ms:.4f} ms per half-sweep, {achieved:.1f} GB/s algorithmic")
    # one whole-lattice launch per half-sweep (streaming loads), for comparison
    job.restore_config()
    rc, rep = job.fix(params, gfc.Options(precision=gfc.PREC_FP64, time_sweeps=1,
                                          slice_group=-1))
    sweep_ms = rep.seconds_sweeps / rep.sweeps * 1e3
    log(f"fp64 whole-lattice sweep kernel: {sweep_ms:.4f} ms/launch")

    def time_fix(prm, opt, reps):
        ts, last = [], None
        for _ in range(reps):
            job.restore_config()
            tm = time.perf_counter()
            rc, last = job.fix(prm, opt)
            ts.append(time.perf_counter() - tm)
            assert rc > 0, rc
        return float(np.mean(ts)), last

    out_mixed = {}
    if not a.no_mixed:
        time_fix(params, mixed, 1)  # warm-up
        mixed_s, rep_m = time_fix(params, mixed, a.steps)
        e2m, thm, Fm = job.observables()
        plaqm = job.plaquette()
        # hand-off as soon as the FP32 phase is finished (no deferral)
        imm_s, rep_i = time_fix(params, gfc.Options(precision=gfc.PREC_MIXED, mixed_defer=0), 1)
        # FP32 sweep-kernel duration (HIP events), 50 FP32 half-sweeps
        job.restore_config()
        rc, rep32 = job.fix(gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=10 ** 8,
                                       max_sweeps_to_fix=50, sweeps_to_reunitarization=10),
                            gfc.Options(precision=gfc.PREC_MIXED, mixed_switch=0.0,
                                        mixed_probe_every=10 ** 6, time_sweeps=1,
                                        slice_group=-1))
        sweep32_ms = rep32.seconds_sweeps / rep32.sweeps * 1e3
        # FP32 half-sweeps as the mixed fix runs them (default grouping), HIP
        # events around each run of blocked sweeps, 200 half-sweeps
        job.restore_config()
        rc, rep32f = job.fix(gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=10 ** 8,
                                        max_sweeps_to_fix=200, sweeps_to_reunitarization=10),
                             gfc.Options(precision=gfc.PREC_MIXED, mixed_switch=0.0,
                                         mixed_probe_every=10 ** 6, time_sweeps=1))
        half32_ms = rep32f.seconds_sweeps / rep32f.sweeps * 1e3
        fixprof32 = f"sweep_fix_fp32_{a.ns}x{a.nt}_pmc.txt"
        t32_upd = pmc_value(fixprof32, "fix_fp32_traffic_bytes_per_update")
        v32_upd = pmc_value(fixprof32, "fix_fp32_valu_instructions_per_update")
        # the same comparison under the dense check policy (a check every 10 sweeps)
        dense = gfc.Params(tolerance=1e-16, estimate_sweeps_to_gf_progress=0,
                           sweeps_to_reunitarization=10, omega_OR=1.95, hits=2)
        d64_s, d64 = time_fix(dense, fp64, 1)
        dmx_s, dmx = time_fix(dense, mixed, 1)
        out_mixed = {
            "mixed_seconds": round(mixed_s, 4),
            "mixed_sweeps": rep_m.sweeps, "mixed_sweeps_fp32": rep_m.sweeps_fp32,
            "mixed_speedup_vs_fp64": round(step_s / mixed_s, 3),
            "mixed_final_e2": e2m, "mixed_F": Fm,
            "mixed_F_rel_diff_vs_fp64": abs(Fm - F64) / abs(F64),
            "mixed_plaquette_drift": max(abs(plaqm[0] - plaq0[0]), abs(plaqm[1] - plaq0[1])),
            "mixed_immediate_handoff": {"seconds": round(imm_s, 4), "sweeps": rep_i.sweeps,
                                        "sweeps_fp32": rep_i.sweeps_fp32,
                                        "speedup_vs_fp64": round(step_s / imm_s, 3)},
            "fp32_sweep_kernel_ms": round(sweep32_ms, 5),
            "fix_fp32_half_sweep_ms": round(half32_ms, 5),
            "fix_fp32_over_fp64_half_sweep": round(half32_ms / half_ms, 4),
            # the FP32 half-sweep's L2-miss traffic and VALU issue, from the
            # committed counters of the same FP32 sweeps
            "fix_fp32_traffic_bytes_per_update": t32_upd,
            "fix_fp32_traffic_GBs": round(t32_upd * vh / (half32_ms * 1e-3) / 1e9, 1)
            if t32_upd else None,
            "fix_fp32_valu_issue_frac": round(v32_upd * vh / 64 / (half32_ms * 1e-3) /
                                              VALU_WAVE_INSTR_PEAK, 4) if v32_upd else None,
            "fp32_sweep_GBs_algorithmic": round(ALG_BYTES_PER_UPDATE_FP32 * vh / sweep32_ms / 1e6,
                                                1),
            "dense_schedule": {"fp64_seconds": round(d64_s, 4), "fp64_sweeps": d64.sweeps,
                               "mixed_seconds": round(dmx_s, 4), "mixed_sweeps": dmx.sweeps,
                               "mixed_sweeps_fp32": dmx.sweeps_fp32,
                               "mixed_speedup_vs_fp64": round(d64_s / dmx_s, 3)},
        }
        log(f"mixed: {mixed_s:.3f}s sweeps={rep_m.sweeps} fp32={rep_m.sweeps_fp32} e2={e2m:.3e} "
            f"speedup={step_s / mixed_s:.3f}; immediate hand-off {imm_s:.3f}s "
            f"(fp32 {rep_i.sweeps_fp32}) speedup {step_s / imm_s:.3f}; fp32 sweep {sweep32_ms:.4f} ms; dense schedule: "
            f"fp64 {d64_s:.3f}s ({d64.sweeps}) mixed {dmx_s:.3f}s ({dmx.sweeps}, "
            f"fp32 {dmx.sweeps_fp32}) speedup {d64_s / dmx_s:.3f}")
    io = None
    if not a.no_io:
        try:
            io = io_leg(job, geom, U)
        except Exception as e:  # e.g. no room for the 3 GB file: report, go on
            io = {"error": f"{type(e).__name__}: {e}"}
            log(f"io leg skipped: {io['error']}")
    job.close()

    serve_out = None
    if a.serve > 0:
        from gfc import serve
        reqs = serve.stream_requests(n=a.serve, rate=a.serve_rate)
        log(f"serving {a.serve} requests, Poisson rate {a.serve_rate}/s, mixed precision")
        sopt = gfc.Options(precision=gfc.PREC_MIXED)
        with gfc.Server(ctx, max_jobs=64) as srv:
            serve_out, comps = serve.run_stream(srv, reqs, sopt, log=log)
        # offered load against the engine's lone-request service times
        lone = serve.lone_service_times(ctx, reqs, sopt)
        serve_out.update(serve.utilisation(reqs, comps, lone))
        serve_out["lone_service_ms"] = {f"{k[0]}^3x{k[1]} tol {k[2]:.0e}": round(v[0] * 1e3, 2)
                                        for k, v in sorted(lone.items())}
        serve_out = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in serve_out.items()}
        serve_out["by_size"] = {k: {kk: (round(vv, 2) if isinstance(vv, float) else vv)
                                    for kk, vv in v.items()}
                                for k, v in serve_out["by_size"].items()}
        serve_out["workload"] = ("serving mix: sizes {16^3x32: 40%, 24^3x48: 30%, 32^3x64: 20%, "
                                 "48^3x96: 10%}, eps U[0.2,0.4], tol {1e-14, 1e-16}, "
                                 f"Poisson {a.serve_rate}/s, inputs synthesised in HBM")
        log(f"serve: {serve_out}")
    c5 = None
    if not a.no_64x128:
        c5 = lattice_64x128(ctx, 64, 128, a.eps, a.seed)
    ctx.close()

    cpu = None
    if not a.no_cpu:
        log("timing the CPU reference path ...")
        cpu = cpu_baseline(U, a.ns, a.nt, a.cpu_seconds)
        log(f"cpu: {cpu}")

    rec = {
        "metric": "site-updates/s & time to e2<1e-16, FP64 vs mixed, 48^3x96; serve p50/p99 ms",
        "value": value,
        "unit": "site-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded smooth SU(3) configuration, FP32 input)",
        "config": {"workload": f"Coulomb gauge fix {a.ns}^3x{a.nt} eps {a.eps} to e2<1e-16, "
                               "FP64 drop-in path, reference check schedule (est 1000), "
                               "omega 1.95, hits 2, reunitarize every 10",
                   "lattice": f"{a.ns}^3x{a.nt}", "sites": V, "parallelism": "single GPU"},
        "fp64_seconds": round(step_s, 4),
        "fp64_sweeps": sweeps[0],
        "fp64_final_e2": e2_64,
        "fp64_F": F64,
        "fp64_plaquette_drift": max(abs(plaq64[0] - plaq0[0]), abs(plaq64[1] - plaq0[1])),
        "fix_half_sweep_ms": round(half_ms, 5),
        "whole_lattice_sweep_kernel_ms": round(sweep_ms, 5),
        "whole_lattice_sweep_GBs_algorithmic": round(
            ALG_BYTES_PER_UPDATE_FP64 * vh / (sweep_ms * 1e-3) / 1e9, 1),
        # the fix's rate in the bytes the reference's own update moves (2016 B
        # per site update): a reference-equivalent rate, not a memory rate
        "fix_reference_equivalent_GBs": round(value * REF_BYTES_PER_UPDATE_FP64 / 1e9, 1),
        **out_mixed,
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(traffic_upd * vh) if traffic_upd else None,
            "kernel": "k_sweep<double> of the timed fix in its default grouping "
                      "(48^3x96: 96 single-timeslice launches of 55296 site updates per "
                      "half-sweep, cache-resident waves on 4 streams)",
            "per": f"lattice half-sweep ({vh} site updates)",
            "algorithmic_bytes_per_update": ALG_BYTES_PER_UPDATE_FP64,
            "traffic_bytes_per_update": traffic_upd,
            "traffic_GBs": round(traffic_upd * vh / (half_ms * 1e-3) / 1e9, 1)
            if traffic_upd else None,
            "valu_issue_frac": round(valu_upd * vh / 64 / (half_ms * 1e-3) / VALU_WAVE_INSTR_PEAK,
                                     4) if valu_upd else None,
            # the measured read rate of an Infinity-Cache-resident set (best of
            # 128, 192 and 256 MiB) and of HBM, and the traffic's rate as a
            # fraction of the former
            "cache_read_GBs": round(cache_read_gbs, 1),
            "hbm_read_GBs": round(hbm_read_gbs, 1),
            "traffic_frac_of_cache_read": round(
                traffic_upd * vh / (half_ms * 1e-3) / 1e9 / cache_read_gbs, 4)
            if traffic_upd else None},
        "copy_kernel_GBs": round(copy_gbs, 1),
        "serve": serve_out,
        "lattice_64x128": c5,
        "io": io,
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
