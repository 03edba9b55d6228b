"""The largest lattice one job accepts, and the first one it refuses.

The kernels address one parity's field of one direction with 32-bit byte
offsets, which caps a job at 89 478 464 sites.  80^3 x 174 (89 088 000 sites)
puts those fields at 99.6% of the 32-bit range.  At that size:

* timeslices are independent in Coulomb gauge, so after a fixed number of
  half-sweeps G on timeslices 0 and 1 equals, bit for bit, G of an 80^3 x 2
  lattice with the same spatial links (the device generator is a function of
  the site index, and temporal links do not enter the update);
* a full FP64 fix stops at the first scheduled check with e2 <= 1e-16, leaves
  the plaquettes unchanged to 1e-14, and U = G U0 G(n+mu)^dagger holds on
  links sampled over the whole lattice, the last timeslice and last site
  included (1e-12), with G in SU(3) (1e-13).

82^3 x 164 (90 424 352 sites) is refused with GFC_EGEOMETRY before anything
is allocated.
"""
import numpy as np
import pytest

import gfc

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

NS, NT = 80, 174
SEED, EPS = 5, 0.35


def _neighbour(site, mu, ns, nt):
    x = site % ns
    y = (site // ns) % ns
    z = (site // ns ** 2) % ns
    t = site // ns ** 3
    c = [t, z, y, x]
    c[mu] = (c[mu] + 1) % (nt if mu == 0 else ns)
    return ((c[0] * ns + c[1]) * ns + c[2]) * ns + c[3]


def test_too_large_is_refused(ctx):
    with pytest.raises(gfc.GfcError) as e:
        gfc.Job(ctx, gfc.Geometry(82, 164))
    assert e.value.code == gfc.EGEOMETRY


def test_largest_lattice(ctx):
    geom = gfc.Geometry(NS, NT)
    assert geom.volume == 89_088_000
    S = NS ** 3
    fixed = gfc.Params(tolerance=1e-300, estimate_sweeps_to_gf_progress=100000000,
                       max_sweeps_to_fix=21, sweeps_to_reunitarization=10)
    with gfc.Job(ctx, gfc.Geometry(NS, 2)) as small:
        small.generate_smooth(SEED, EPS)
        assert small.reunitarize() == 0
        rc, _ = small.fix(fixed)
        assert rc == gfc.ENOTCONVERGED
        G_small = small.download_gauge()
    with gfc.Job(ctx, geom) as job:
        job.generate_smooth(SEED, EPS)
        assert job.reunitarize() == 0
        rc, _ = job.fix(fixed)
        assert rc == gfc.ENOTCONVERGED
        G = job.download_gauge()
        assert np.array_equal(G[:2 * S], G_small)
        print(f"{NS}^3x{NT}: G on timeslices 0-1 after 21 half-sweeps equals {NS}^3x2 bit for bit",
              flush=True)
        del G, G_small

        # full fix from the unfixed input
        job.generate_smooth(SEED, EPS)
        assert job.reunitarize() == 0
        rng = np.random.default_rng(11)
        sites = [int(s) for s in rng.integers(0, geom.volume, 48)]
        sites += [geom.volume - 1, geom.volume - S, (NT - 1) * S + S // 2]
        samples = [(s, mu, job.get_link(s, mu)) for s in sites for mu in (0, 3)]
        plaq0 = job.plaquette()
        F0 = job.observables()[2]
        p = gfc.Params(sweeps_to_reunitarization=10)
        rc, rep = job.fix(p)
        print(f"{NS}^3x{NT} FP64: {rc} sweeps, e2 {rep.residue_final:.3e}, {rep.seconds:.3f} s",
              flush=True)
        assert rc == p.next_check(0, rep.residue_initial)
        e2, _, F = job.observables()
        assert e2 <= p.tolerance and F > F0
        plaq = job.plaquette()
        assert abs(plaq[0] - plaq0[0]) < 1e-14 and abs(plaq[1] - plaq0[1]) < 1e-14
        G = job.download_gauge()
        eye = np.eye(3)
        for site, mu, u0 in samples:
            g, gn = G[site], G[_neighbour(site, mu, NS, NT)]
            assert np.abs(g @ g.conj().T - eye).max() < 1e-13
            assert abs(np.linalg.det(g) - 1) < 1e-13
            assert np.abs(job.get_link(site, mu) - g @ u0 @ gn.conj().T).max() < 1e-12
