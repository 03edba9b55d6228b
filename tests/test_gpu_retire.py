"""Per-timeslice convergence tracking (gfc_options.retire_slices).

Timeslices are independent for w, e2 and theta, so a slice whose own e2 is at
or below the tolerance stops being swept.  The global e2 must still reach the
tolerance, the plaquettes must not move, the result must be the same
gauge (functional equal to the plain run's), and fewer slice half-sweeps are
spent.
"""
import numpy as np
import pytest

import gfc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [gfc.PREC_FP64, gfc.PREC_MIXED])
@pytest.mark.parametrize("est", [1000, 0])
def test_retirement_reaches_tolerance_with_less_work(ctx, precision, est):
    geom = gfc.Geometry(16, 32)
    U = gfc.generate_smooth(geom, 21, 0.35)
    p = gfc.Params(sweeps_to_reunitarization=10, estimate_sweeps_to_gf_progress=est)
    with gfc.Job(ctx, geom, precision) as job:
        job.upload_config(U)
        job.reunitarize()
        job.save_config()
        p0 = job.plaquette()
        rc0, rep0 = job.fix(p, gfc.Options(precision=precision))
        F0 = job.observables()[2]
        job.restore_config()
        rc, rep = job.fix(p, gfc.Options(precision=precision, retire_slices=1,
                                         mixed_probe_every=20))
        e2, th, F = job.observables()
        p1 = job.plaquette()
    print(f"est={est} prec={precision}: sweeps {rc} vs {rc0}, slice-sweeps {rep.slice_sweeps:.1f}, "
          f"e2 {e2:.3e}")
    assert rc > 0 and e2 <= 1e-16
    # under the reference schedule the first check comes long after most
    # slices converged; with a check every 10 sweeps there is less to save
    assert rep.slice_sweeps < rep.sweeps if est else rep.slice_sweeps <= rep.sweeps
    assert abs(F - F0) / abs(F0) < 1e-10
    assert abs(p1[0] - p0[0]) < 1e-14 and abs(p1[1] - p0[1]) < 1e-14


def test_retirement_needs_ns_multiple_of_8(ctx):
    geom = gfc.Geometry(6, 8)
    with gfc.Job(ctx, geom) as job:
        job.generate_smooth(1, 0.3)
        with pytest.raises(gfc.GfcError) as e:
            job.fix(gfc.Params(), gfc.Options(retire_slices=1))
        assert e.value.code == gfc.EINVAL


@pytest.mark.parametrize("precision", [gfc.PREC_FP64, gfc.PREC_MIXED])
def test_retirement_against_reference(ctx, precision):
    """Slice retirement on 8^3x16 (eps 0.25, seed 2: the reference's fixture run to
    1e-16 under its default schedule): the same sweep count, e2 <= 1e-16 and
    the reference's gauge functional F to 1e-10 relative (retired slices stop
    moving below the tolerance, so G is not compared)."""
    import hashlib
    from conftest import golden
    f = golden("lat_8x16.npz")
    geom = gfc.Geometry(int(f["ns"]), int(f["nt"]))
    U = gfc.generate_smooth(geom, int(f["seed"]), float(f["eps"]))
    assert hashlib.sha256(U.tobytes()).hexdigest() == str(f["config_sha256"])
    p = gfc.Params(tolerance=1e-16, sweeps_to_reunitarization=10,
                   estimate_sweeps_to_gf_progress=1000)
    with gfc.Job(ctx, geom, precision) as job:
        job.upload_config(U)
        assert job.reunitarize() == 0
        rc, rep = job.fix(p, gfc.Options(precision=precision, retire_slices=1,
                                         mixed_probe_every=20))
        e2, th, F = job.observables()
    Fref = float(f["full_est1000_obs"][2])
    print(f"precision {precision}: sweeps {rc} (reference {int(f['full_est1000_rc'])}), "
          f"slice-sweeps {rep.slice_sweeps:.1f}, e2 {e2:.2e}, F rel {abs(F - Fref) / Fref:.2e}")
    assert rc == int(f["full_est1000_rc"])
    assert e2 <= 1e-16 and rep.slice_sweeps < rep.sweeps
    assert abs(F - Fref) / abs(Fref) < 1e-10
