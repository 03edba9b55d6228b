"""Full fixes on lattice shapes that take each branch of the default sweep
grouping (whole lattice from the cache; cache-sized groups of several
timeslices with cached loads; single timeslices with streaming loads), with
spatial extents that are not multiples of 8.

For each shape, under the reference schedule:
* the default grouping gives G bit for bit equal to whole-lattice sweeps
  (slice_group = -1), with the same sweep count;
* the fix reaches e2 <= 1e-16 and leaves the plaquettes unchanged to 1e-14;
* the mixed-precision fix reaches the tolerance too, and its gauge functional
  agrees with the FP64 one to 1e-10 relative.
"""
import numpy as np
import pytest

import gfc

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

# (ns, nt): 12^3 x 24 and 20^3 x 40 fit in the cache whole; 40^3 x 80 runs
# cached groups of 2 timeslices (23.4 MiB each in FP64); 56^3 x 112 runs single
# timeslices (64.3 MiB in FP64, streamed; 32.2 MiB in FP32, cached)
SHAPES = [(12, 24), (20, 40), (40, 80), (56, 112)]


@pytest.mark.parametrize("ns,nt", SHAPES)
def test_shape(ctx, ns, nt):
    geom = gfc.Geometry(ns, nt)
    p = gfc.Params(sweeps_to_reunitarization=10)
    with gfc.Job(ctx, geom, gfc.PREC_MIXED) as job:
        job.generate_smooth(ns + nt, 0.3)
        assert job.reunitarize() == 0
        job.save_config()
        plaq0 = job.plaquette()
        G = {}
        for grp in (-1, 0):
            job.restore_config()
            rc, rep = job.fix(p, gfc.Options(precision=gfc.PREC_FP64, slice_group=grp))
            assert rc > 0 and rep.residue_final <= p.tolerance
            G[grp] = (rc, job.download_gauge())
        assert G[0][0] == G[-1][0]
        assert np.array_equal(G[0][1], G[-1][1])
        plaq = job.plaquette()
        assert abs(plaq[0] - plaq0[0]) < 1e-14 and abs(plaq[1] - plaq0[1]) < 1e-14
        F64 = job.observables()[2]
        job.restore_config()
        rc, rep = job.fix(p, gfc.Options(precision=gfc.PREC_MIXED))
        print(f"{ns}^3x{nt}: {rc} sweeps, fp32 {rep.sweeps_fp32}, e2 {rep.residue_final:.2e}")
        assert rc > 0 and rep.residue_final <= p.tolerance
        assert abs(job.observables()[2] - F64) / F64 < 1e-10
