"""Timeslice-blocked sweeps (gfc_options.slice_group).

Between two host decisions the sweeps run group of timeslices by group of
timeslices on several streams.  A site update reads only its own timeslice,
so the gauge transformation, the sweep count and every residue must equal
those of whole-lattice sweeps bit for bit, for any grouping (ragged last
group, one group, one slice per group) and in both precisions.  The sweeps
read u0 and G(n) through the cache when what the concurrent launches sweep
fits in 192 MiB and with non-temporal loads otherwise; both load policies
give the same bits.
"""
import numpy as np
import pytest

import gfc

pytestmark = pytest.mark.gpu


def _run(job, params, options):
    job.restore_config()
    rc, rep = job.fix(params, options)
    return rc, rep, job.download_gauge(), job.download_config()


@pytest.mark.parametrize("precision", [gfc.PREC_FP64, gfc.PREC_MIXED])
def test_blocked_equals_whole_lattice(ctx, precision):
    geom = gfc.Geometry(8, 16)
    p = gfc.Params(estimate_sweeps_to_gf_progress=400)
    with gfc.Job(ctx, geom, precision) as job:
        job.generate_smooth(5, 0.3)
        job.reunitarize()
        job.save_config()
        o = dict(precision=precision, mixed_probe_every=30)
        rc0, rep0, G0, U0 = _run(job, p, gfc.Options(slice_group=-1, **o))
        assert rc0 > 0
        for grp in (1, 3, 7, 16, 0):
            rc, rep, G, U = _run(job, p, gfc.Options(slice_group=grp, **o))
            print(f"precision {precision} group {grp}: {rc} sweeps, e2 {rep.residue_final:.3e}")
            assert rc == rc0
            assert rep.sweeps_fp32 == rep0.sweeps_fp32
            assert rep.residue_final == rep0.residue_final
            assert np.array_equal(G, G0)
            assert np.array_equal(U, U0)


def test_blocked_stop_at_max_sweeps(ctx):
    # max sweeps (not a multiple of the reunitarization period) ends a block
    geom = gfc.Geometry(8, 16)
    p = gfc.Params(tolerance=1e-300, max_sweeps_to_fix=137, sweeps_to_reunitarization=10)
    with gfc.Job(ctx, geom) as job:
        job.generate_smooth(6, 0.3)
        job.reunitarize()
        job.save_config()
        out = []
        for grp in (-1, 5):
            job.restore_config()
            rc, rep = job.fix(p, gfc.Options(slice_group=grp))
            assert rc == gfc.ENOTCONVERGED and rep.sweeps == 137
            out.append(job.download_gauge())
        assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("nt", [16, 6])
@pytest.mark.parametrize("precision", [gfc.PREC_FP64, gfc.PREC_MIXED])
def test_cached_and_streamed_loads_agree(ctx, precision, nt):
    # 48^3 x nt: one timeslice holds 40.5 MiB of u0 and G in FP64 (20.25 MiB
    # in FP32), so whole-lattice sweeps and groups of 3 timeslices on four
    # streams use non-temporal loads, groups of 1 and the default grouping
    # cached ones (with nt = 6 the FP64 waves end with a ragged one of 2
    # groups, and in FP32 the whole lattice, 122 MiB, is swept from the cache)
    geom = gfc.Geometry(48, nt)
    p = gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=100000000,
                   max_sweeps_to_fix=64, sweeps_to_reunitarization=10)
    with gfc.Job(ctx, geom, precision) as job:
        job.generate_smooth(8, 0.35)
        job.reunitarize()
        job.save_config()
        o = dict(precision=precision, mixed_probe_every=30)
        rc0, rep0, G0, _ = _run(job, p, gfc.Options(slice_group=-1, **o))
        assert rc0 == gfc.ENOTCONVERGED and rep0.sweeps == 64
        for grp in (1, 3, 0):
            rc, rep, G, _ = _run(job, p, gfc.Options(slice_group=grp, **o))
            assert rc == rc0 and rep.sweeps_fp32 == rep0.sweeps_fp32
            assert np.array_equal(G, G0)


@pytest.mark.parametrize("ns,nt", [(10, 20), (14, 28)])
@pytest.mark.parametrize("precision", [gfc.PREC_FP64, gfc.PREC_MIXED])
def test_blocked_tiles_shared_across_timeslices(ctx, precision, ns, nt):
    """ns = 2 mod 4: a timeslice (ns^3/2 parity sites) is not a whole number of
    32-site tiles, so groups of timeslices on different streams write
    different elements of the same G tiles and cache lines concurrently; the
    result must still equal whole-lattice sweeps bit for bit."""
    geom = gfc.Geometry(ns, nt)
    assert (ns ** 3 // 2) % 32 != 0
    p = gfc.Params(estimate_sweeps_to_gf_progress=400, sweeps_to_reunitarization=10)
    with gfc.Job(ctx, geom, precision) as job:
        job.generate_smooth(ns + 40, 0.3)
        job.reunitarize()
        job.save_config()
        o = dict(precision=precision, mixed_probe_every=30)
        rc0, rep0, G0, U0 = _run(job, p, gfc.Options(slice_group=-1, **o))
        assert rc0 > 0
        for grp in (1, 3, 0):
            rc, rep, G, U = _run(job, p, gfc.Options(slice_group=grp, **o))
            print(f"{ns}^3x{nt} precision {precision} group {grp}: {rc} sweeps")
            assert rc == rc0 and rep.sweeps_fp32 == rep0.sweeps_fp32
            assert rep.residue_final == rep0.residue_final
            assert np.array_equal(G, G0)
            assert np.array_equal(U, U0)
