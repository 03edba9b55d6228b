"""Parity of the HIP engine with the CPU oracle away from the default
parameters: overrelaxation omega (stored as float, as in the reference's
struct), the number of Cabibbo-Marinari hits, the reunitarization interval R
and the check estimate (gaugefixOverrelaxation, src/gauge_fixing.c:713-826;
parameters as parsed in src/gauge_fixing.c:164-244).

The reference fixtures cover omega 1.95, 2 hits, R 10 (and omega 1.5 / 1.0 on
the degenerate lattice); here the oracle, which matches those fixtures, gives
the expected values for other settings on the same seeded inputs:
  * e2 after every one of 61 half-sweeps of one run (whole-lattice sweeps)
    within 1e-12 relative plus 1e-14 sqrt(e2), and G after them within 1e-10.
    Some of these runs converge within the 61 sweeps; once e2 is small, the
    divergence it sums carries rounding of about 1e-15 (the size of the G
    difference), which moves e2 by about 2e-15 sqrt(e2): relative to e2 that
    exceeds 1e-12 below e2 ~ 4e-6;
  * a run to tolerance on the default (timeslice-blocked) path: the same
    sweep count, G within 1e-10, e2 at or below the tolerance.
"""
import numpy as np
import pytest

import gfc
import oracle as orc

pytestmark = pytest.mark.gpu

E2_REL = 1e-12
E2_SQRT = 1e-14  # absolute part, times sqrt(e2)
G_ABS = 1e-10

# (omega, hits, R, estimate): omega 1 is the plain Los Alamos update, 1.3 is not
# a float exactly, R 1 reunitarizes after every sweep, R 250 never inside 61
# sweeps, R 3 / 7 / 13 fall on both parities and off the probe grid
VARIANTS = [
    (1.7, 1, 7, 1000),
    (1.0, 3, 3, 0),
    (1.95, 4, 250, 500),
    (1.5, 2, 1, 1000),
    (1.3, 2, 13, 200),
]
LATTICES = [(4, 8, 5, 0.3), (6, 4, 9, 0.35)]


def _inputs(ns, nt, seed, eps):
    geom = gfc.Geometry(ns, nt)
    U = gfc.generate_smooth(geom, seed, eps)
    U1 = U.copy()
    assert orc.reunitarize(U1) == 0
    return geom, U, U1


def _job(ctx, geom, U):
    job = gfc.Job(ctx, geom)
    job.upload_config(U)
    assert job.reunitarize() == 0
    return job


@pytest.mark.parametrize("ns,nt,seed,eps", LATTICES)
@pytest.mark.parametrize("omega,hits,R,est", VARIANTS)
def test_trace_and_state_after_61_sweeps(ctx, ns, nt, seed, eps, omega, hits, R, est):
    geom, U, U1 = _inputs(ns, nt, seed, eps)
    k = 61
    G1 = np.zeros((geom.volume, 3, 3), np.complex128)
    rc_o, tr_o = orc.gaugefix(U1, G1, orc.geom(ns, nt),
                              orc.params(tolerance=1e-300, max_sweeps=k, estimate=100000000,
                                         reunit_every=R, omega=omega, hits=hits), trace_len=k)
    job = _job(ctx, geom, U)
    p = gfc.Params(tolerance=1e-300, max_sweeps_to_fix=k, estimate_sweeps_to_gf_progress=100000000,
                   sweeps_to_reunitarization=R, omega_OR=omega, hits=hits)
    rc, rep, tr = job.fix(p, trace_len=k)
    assert rc == rc_o == gfc.ENOTCONVERGED and rep.sweeps == k
    diff = np.abs(tr - tr_o)
    bound = E2_REL * tr_o + E2_SQRT * np.sqrt(tr_o)
    dG = np.max(np.abs(job.download_gauge() - G1))
    print(f"{ns}^3x{nt} omega {omega} hits {hits} R {R}: e2 {tr_o[0]:.1e} -> {tr_o[-1]:.1e}, "
          f"max rel e2 {np.max(diff / tr_o):.2e}, max diff/bound {np.max(diff / bound):.2f}, "
          f"max|dG| {dG:.2e}")
    assert np.all(diff <= bound)
    assert dG < G_ABS
    job.close()


@pytest.mark.parametrize("ns,nt,seed,eps", LATTICES)
@pytest.mark.parametrize("omega,hits,R,est", VARIANTS)
def test_run_to_tolerance(ctx, ns, nt, seed, eps, omega, hits, R, est):
    geom, U, U1 = _inputs(ns, nt, seed, eps)
    tol = 1e-14
    G1 = np.zeros((geom.volume, 3, 3), np.complex128)
    rc_o, _ = orc.gaugefix(U1, G1, orc.geom(ns, nt),
                           orc.params(tolerance=tol, estimate=est, reunit_every=R, omega=omega,
                                      hits=hits))
    job = _job(ctx, geom, U)
    p = gfc.Params(tolerance=tol, estimate_sweeps_to_gf_progress=est,
                   sweeps_to_reunitarization=R, omega_OR=omega, hits=hits)
    rc, rep = job.fix(p)
    e2 = job.observables()[0]
    dG = np.max(np.abs(job.download_gauge() - G1))
    print(f"{ns}^3x{nt} omega {omega} hits {hits} R {R} est {est}: sweeps {rc} "
          f"(oracle {rc_o}), e2 {e2:.2e}, max|dG| {dG:.2e}")
    assert rc_o > 0 and rc == rc_o
    assert e2 <= tol
    assert dG < G_ABS
    job.close()


@pytest.mark.parametrize("omega,hits,R,est", VARIANTS)
def test_mixed_precision_variants(ctx, omega, hits, R, est):
    """The FP32 -> FP64 path under the same settings, on 8^3x16: it stops at a
    scheduled check no earlier than the FP64 run's, with e2 at or below the
    tolerance and the gauge functional F equal to the FP64 one to 1e-10
    relative (the same gauge copy)."""
    geom = gfc.Geometry(8, 16)
    p = gfc.Params(tolerance=1e-14, estimate_sweeps_to_gf_progress=est,
                   sweeps_to_reunitarization=R, omega_OR=omega, hits=hits)
    with gfc.Job(ctx, geom, gfc.PREC_MIXED) as job:
        job.generate_smooth(21, 0.3)
        job.reunitarize()
        job.save_config()
        rc64, _ = job.fix(p, gfc.Options(precision=gfc.PREC_FP64))
        F64 = job.observables()[2]
        job.restore_config()
        rc, rep = job.fix(p, gfc.Options(precision=gfc.PREC_MIXED))
        e2, _, F = job.observables()
        print(f"omega {omega} hits {hits} R {R} est {est}: FP64 {rc64}, mixed {rc} "
              f"(fp32 {rep.sweeps_fp32}), e2 {e2:.2e}, F rel {abs(F - F64) / F64:.2e}")
        assert rc64 > 0 and rc >= rc64
        assert e2 <= 1e-14
        assert abs(F - F64) / abs(F64) < 1e-10
