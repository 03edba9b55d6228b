"""Full-size runs (48^3 x 96 from FP32 input; 64^3 x 128 with per-timeslice
retirement).  The 48^3 x 96 run is compared with the compiled reference's
own run on the same input (tests/golden/ref_48x96_fix.json, written by
tests/golden/make_ref_fullsize.py): same sweep count, F to 1e-12 (FP64) and
1e-10 (mixed) relative, plaquettes to 1e-14.  Both runs are also checked
through properties that hold at any size:

* the fix stops at the first scheduled check, sweep gfc_next_check(0, e2(0)),
  with e2 <= 1e-16 there;
* plaquettes are gauge invariant (1e-14);
* every sampled link satisfies U_mu(n) = G(n) U0_mu(n) G(n+mu)^dagger, with U0
  the reunitarized input (1e-12), and G(n) is in SU(3);
* the gauge functional F rises, and the mixed-precision result reaches the
  same gauge as the FP64 one (F to 1e-10 relative).
"""
import json
import os

import numpy as np
import pytest

import gfc

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

RNG = np.random.default_rng(7)
REF_48 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_48x96_fix.json")


def _neighbour(site, mu, ns, nt):
    x = site % ns
    y = (site // ns) % ns
    z = (site // ns ** 2) % ns
    t = site // ns ** 3
    c = [t, z, y, x]
    c[mu] = (c[mu] + 1) % (nt if mu == 0 else ns)
    return ((c[0] * ns + c[1]) * ns + c[2]) * ns + c[3]


def _samples(job, geom, n=64):
    sites = RNG.integers(0, geom.volume, n)
    mus = RNG.integers(0, 4, n)
    return [(int(s), int(m), job.get_link(int(s), int(m))) for s, m in zip(sites, mus)]


def _check_fixed(job, geom, samples, plaq0, params, res0):
    e2, theta, F = job.observables()
    assert e2 <= params.tolerance
    assert abs(theta - e2 / 6) <= 1e-9 * e2
    plaq = job.plaquette()
    assert abs(plaq[0] - plaq0[0]) < 1e-14 and abs(plaq[1] - plaq0[1]) < 1e-14
    G = job.download_gauge()
    eye = np.eye(3)
    for site, mu, u0 in samples:
        g, gn = G[site], G[_neighbour(site, mu, geom.ns, geom.nt)]
        assert np.abs(g @ g.conj().T - eye).max() < 1e-13
        assert abs(np.linalg.det(g) - 1) < 1e-13
        u = job.get_link(site, mu)
        assert np.abs(u - g @ u0 @ gn.conj().T).max() < 1e-12
    return F


def test_48x96_fp32_input_fp64_and_mixed(ctx):
    geom = gfc.Geometry(48, 96)
    U32 = gfc.to_fp32(gfc.generate_smooth(geom, 3, 0.35))
    p = gfc.Params(sweeps_to_reunitarization=10)
    with gfc.Job(ctx, geom, gfc.PREC_MIXED) as job:
        job.upload_config(U32)
        del U32
        job.reunitarize()
        job.save_config()
        samples = _samples(job, geom)
        plaq0 = job.plaquette()
        e2_0, _, F0 = job.observables()
        ref = json.load(open(REF_48))
        assert abs(plaq0[0] - ref["plaq0"][0]) < 1e-14 and abs(plaq0[1] - ref["plaq0"][1]) < 1e-14
        assert abs(e2_0 - ref["e2_0"]) / ref["e2_0"] < 1e-12
        assert abs(F0 - ref["F0"]) < 1e-15
        Fs = {}
        for prec in (gfc.PREC_FP64, gfc.PREC_MIXED):
            job.restore_config()
            rc, rep = job.fix(p, gfc.Options(precision=prec))
            first = p.next_check(0, rep.residue_initial)
            print(f"48^3x96 precision {prec}: {rc} sweeps (first check {first}), "
                  f"fp32 {rep.sweeps_fp32}, e2 {rep.residue_final:.3e}, {rep.seconds:.3f} s")
            assert rc == first == ref["sweeps"]
            Fs[prec] = _check_fixed(job, geom, samples, plaq0, p, rep.residue_initial)
            plaq = job.plaquette()
            rel = abs(Fs[prec] - ref["F"]) / ref["F"]
            print(f"  vs reference: F rel diff {rel:.2e}, plaquette diffs "
                  f"{abs(plaq[0] - ref['plaq'][0]):.1e} {abs(plaq[1] - ref['plaq'][1]):.1e}")
            assert rel < (1e-12 if prec == gfc.PREC_FP64 else 1e-10)
            assert abs(plaq[0] - ref["plaq"][0]) < 1e-14 and abs(plaq[1] - ref["plaq"][1]) < 1e-14
        assert Fs[gfc.PREC_FP64] > F0
        assert abs(Fs[gfc.PREC_MIXED] - Fs[gfc.PREC_FP64]) / Fs[gfc.PREC_FP64] < 1e-10


def test_64x128_mixed_with_slice_retirement(ctx):
    geom = gfc.Geometry(64, 128)
    p = gfc.Params(sweeps_to_reunitarization=10)
    with gfc.Job(ctx, geom, gfc.PREC_MIXED) as job:
        job.generate_smooth(3, 0.35)
        job.reunitarize()
        samples = _samples(job, geom)
        plaq0 = job.plaquette()
        F0 = job.observables()[2]
        rc, rep = job.fix(p, gfc.Options(precision=gfc.PREC_MIXED, retire_slices=1))
        print(f"64^3x128 mixed + retirement: {rc} sweeps, fp32 {rep.sweeps_fp32}, "
              f"slice-sweeps {rep.slice_sweeps:.1f}, e2 {rep.residue_final:.3e}, "
              f"{rep.seconds:.3f} s")
        assert rc == p.next_check(0, rep.residue_initial)
        assert rep.slice_sweeps <= rep.sweeps
        F = _check_fixed(job, geom, samples, plaq0, p, rep.residue_initial)
        assert F > F0
