"""Serving: continuous batching of many lattices.

Each request gets the same result alone (gfc_job_fix on its own lattice) and
in the mix (packed with others by the server): the per-site arithmetic and
the reduction order are shared, so G is compared bit for bit and the sweep
counts must agree exactly.
"""
import numpy as np
import pytest

import gfc
from gfc import serve

pytestmark = pytest.mark.gpu


def test_device_generator_matches_host(ctx):
    geom = gfc.Geometry(4, 8)
    U = gfc.generate_smooth(geom, 77, 0.3)
    with gfc.Job(ctx, geom) as job:
        job.generate_smooth(77, 0.3)
        Ud = job.download_config()
    assert np.max(np.abs(Ud - U)) < 1e-12


def _alone(ctx, geom, seed, eps, params, opts):
    with gfc.Job(ctx, geom, opts.precision) as job:
        job.generate_smooth(seed, eps)
        job.reunitarize()
        rc, rep = job.fix(params, opts)
        return rc, rep, job.download_gauge()


@pytest.mark.parametrize("precision", [gfc.PREC_FP64, gfc.PREC_MIXED])
def test_alone_equals_in_the_mix(ctx, precision):
    reqs = [(gfc.Geometry(4, 8), 11, 0.3, 1e-16), (gfc.Geometry(8, 8), 12, 0.25, 1e-14),
            (gfc.Geometry(6, 12), 13, 0.35, 1e-16), (gfc.Geometry(8, 16), 14, 0.2, 1e-16),
            (gfc.Geometry(4, 8), 15, 0.4, 1e-14)]
    opts = gfc.Options(precision=precision, mixed_probe_every=20)
    with gfc.Server(ctx, max_jobs=3, keep_results=True) as srv:
        ids = {}
        for k, (geom, seed, eps, tol) in enumerate(reqs):
            p = gfc.Params(tolerance=tol, sweeps_to_reunitarization=10)
            ids[srv.submit(geom, seed, eps, p, opts, tag=k)] = k
        comps = srv.drain(timeout=300)
        assert len(comps) == len(reqs)
        st = srv.stats()
        assert st["graph_replays"] > 0  # the packed chunks were replayed from a graph
        for c in comps:
            geom, seed, eps, tol = reqs[c.tag]
            p = gfc.Params(tolerance=tol, sweeps_to_reunitarization=10)
            rc, rep, G_alone = _alone(ctx, geom, seed, eps, p, opts)
            assert c.status == rc and c.sweeps == rep.sweeps
            assert c.sweeps_fp32 == rep.sweeps_fp32
            assert c.residue_final == rep.residue_final
            G_mix = srv.result_gauge(c.id, geom)
            assert np.array_equal(G_mix, G_alone), np.max(np.abs(G_mix - G_alone))
            srv.release(c.id)


def test_admission_by_footprint(ctx):
    """With a budget for two 8^3x16 lattices, the third waits in the queue."""
    geom = gfc.Geometry(8, 16)
    budget = 2 * gfc.footprint(geom, gfc.PREC_FP64) + 1
    p = gfc.Params(sweeps_to_reunitarization=10)
    with gfc.Server(ctx, mem_budget=budget, max_jobs=8) as srv:
        for k in range(3):
            srv.submit(geom, 100 + k, 0.3, p, gfc.Options(), tag=k)
        srv.poll()
        st = srv.stats()
        assert st["active"] == 2 and st["queued"] == 1
        assert st["bytes_used"] <= budget
        comps = srv.drain(timeout=300)
        assert sorted(c.tag for c in comps) == [0, 1, 2]
        assert all(c.status > 0 for c in comps)
        # the third was admitted only after one of the first two completed
        first_done = min(c.t_done for c in comps if c.tag < 2)
        third = [c for c in comps if c.tag == 2][0]
        assert third.t_admit >= first_done - 1e-6
    with gfc.Server(ctx, mem_budget=budget, max_jobs=8) as srv:
        with pytest.raises(gfc.GfcError):
            srv.submit(gfc.Geometry(16, 32), 1, 0.3, p)


def test_short_stream(ctx):
    """A short Poisson stream of small lattices: every request converges."""
    reqs = serve.stream_requests(n=12, rate=20.0, sizes=((8, 16, 0.5), (12, 24, 0.5)))
    with gfc.Server(ctx, max_jobs=16) as srv:
        summary, comps = serve.run_stream(srv, reqs)
    assert summary["all_converged"]
    assert summary["p99_ms"] >= summary["p50_ms"] > 0


def test_every_lane_alone_equals_in_the_mix(ctx):
    """One lattice per size lane (each lane has its own stream and chunk
    plans), two of the small one: each result equals its lone run bit for
    bit, in mixed precision with the deferred hand-off."""
    reqs = [(gfc.Geometry(8, 16), 21, 0.3), (gfc.Geometry(24, 24), 22, 0.3),
            (gfc.Geometry(32, 32), 23, 0.3), (gfc.Geometry(48, 48), 24, 0.3),
            (gfc.Geometry(8, 16), 25, 0.35)]
    opts = gfc.Options(precision=gfc.PREC_MIXED)
    p = gfc.Params(tolerance=1e-16, sweeps_to_reunitarization=10)
    with gfc.Server(ctx, max_jobs=8, keep_results=True) as srv:
        for k, (geom, seed, eps) in enumerate(reqs):
            srv.submit(geom, seed, eps, p, opts, tag=k)
        comps = srv.drain(timeout=300)
        assert len(comps) == len(reqs)
        # the small lattices finish long before the large one
        t = {c.tag: c.t_done for c in comps}
        assert t[0] < t[3] and t[4] < t[3]
        for c in comps:
            geom, seed, eps = reqs[c.tag]
            rc, rep, G_alone = _alone(ctx, geom, seed, eps, p, opts)
            print(f"{geom.ns}^3x{geom.nt}: sweeps {c.sweeps} fp32 {c.sweeps_fp32} "
                  f"latency {(c.t_done - c.t_submit) * 1e3:.1f} ms")
            assert c.status == rc and c.sweeps == rep.sweeps
            assert c.sweeps_fp32 == rep.sweeps_fp32
            assert c.residue_final == rep.residue_final
            G_mix = srv.result_gauge(c.id, geom)
            assert np.array_equal(G_mix, G_alone), np.max(np.abs(G_mix - G_alone))
            srv.release(c.id)


def _fixture(name):
    import hashlib
    from conftest import golden
    f = golden(name)
    geom = gfc.Geometry(int(f["ns"]), int(f["nt"]))
    U = gfc.generate_smooth(geom, int(f["seed"]), float(f["eps"]))
    assert hashlib.sha256(U.tobytes()).hexdigest() == str(f["config_sha256"])
    return f, geom, U


@pytest.mark.parametrize("precision", [gfc.PREC_FP64, gfc.PREC_MIXED])
def test_served_results_equal_reference(ctx, precision):
    """Every served request runs gaugefixOverrelaxation
    (src/gauge_fixing.c:713-826): the reference's fixture lattices, submitted
    as configurations to the server together, end with the reference's sweep
    count under its default schedule (estimate 1000, tolerance 1e-16, R 10),
    its functional F (1e-12 relative in FP64, 1e-10 mixed) and, in FP64, its
    G (1e-10).  In mixed precision the hand-off is deferred by an estimate of
    the FP64 sweeps the check needs; where that falls short (the 2^3x2
    lattice) the request converges at a later scheduled check, never
    earlier."""
    names = ["lat_4x8.npz", "lat_6x4.npz", "lat_8x16.npz", "lat_2x2.npz"]
    p = gfc.Params(tolerance=1e-16, estimate_sweeps_to_gf_progress=1000,
                   sweeps_to_reunitarization=10)
    opts = gfc.Options(precision=precision)
    with gfc.Server(ctx, max_jobs=8, keep_results=True) as srv:
        fx = {}
        for k, name in enumerate(names):
            f, geom, U = _fixture(name)
            fx[k] = (name, f, geom)
            srv.submit_config(geom, U, p, opts, tag=k)
        comps = srv.drain(timeout=300)
        assert len(comps) == len(names)
        for c in comps:
            name, f, geom = fx[c.tag]
            ref_rc = int(f["full_est1000_rc"])
            Fref = float(f["full_est1000_obs"][2])
            G = srv.result_gauge(c.id, geom)
            Gref = f["full_est1000_G"]
            sel = f["G_sites"] if len(Gref) < geom.volume else slice(None)
            dG = np.max(np.abs(G[sel] - Gref))
            print(f"{name} precision {precision}: sweeps {c.status} (reference {ref_rc}), "
                  f"F rel {abs(c.F - Fref) / Fref:.2e}, max|dG| {dG:.2e}")
            assert c.residue_final <= 1e-16
            if precision == gfc.PREC_FP64:
                assert c.status == ref_rc
                assert abs(c.F - Fref) / abs(Fref) < 1e-12
                assert dG < 1e-10
            else:
                assert c.status >= ref_rc
                assert abs(c.F - Fref) / abs(Fref) < 1e-10
            srv.release(c.id)


def test_served_null_row_fails(ctx):
    """An input with a null row cannot be reunitarized: the request ends with
    GFC_EPROJECT and no sweep (the CLI exits with failure there,
    gauge_fix_coulomb.c:126-131); a good request beside it is unaffected."""
    geom = gfc.Geometry(4, 8)
    good = gfc.generate_smooth(geom, 51, 0.3)
    bad = good.copy()
    bad[17, 2, 1] = 0
    p = gfc.Params(sweeps_to_reunitarization=10)
    with gfc.Server(ctx, max_jobs=4, keep_results=True) as srv:
        srv.submit_config(geom, bad, p, tag=0)
        srv.submit_config(geom, good, p, tag=1)
        srv.submit_config(geom, bad.astype(np.complex64), p, tag=2)
        comps = {c.tag: c for c in srv.drain(timeout=120)}
        assert comps[0].status == gfc.EPROJECT and comps[0].sweeps == 0
        assert comps[2].status == gfc.EPROJECT and comps[2].sweeps == 0
        assert comps[1].status > 0 and comps[1].residue_final <= 1e-16
        with gfc.Job(ctx, geom) as job:
            job.upload_config(good)
            job.reunitarize()
            rc, rep = job.fix(p)
            assert comps[1].status == rc
            assert np.array_equal(srv.result_gauge(comps[1].id, geom), job.download_gauge())


def test_serving_mix_stream_alone_equals_in_the_mix(ctx):
    """A stream shaped like the serving workload: 8 requests with the serving workload's sizes
    (16^3x32 ... 48^3x96, one of each at least), eps in [0.2, 0.4], tolerances
    1e-14 and 1e-16, Poisson arrivals at 4/s, mixed precision.  Each served G
    equals its lone gfc_job_fix bit for bit, with the same sweep counts."""
    reqs = serve.stream_requests(n=8, rate=4.0, seed=7)
    kinds = [(16, 32), (24, 48), (32, 64), (48, 96), (16, 32), (24, 48), (16, 32), (32, 64)]
    for r, (ns, nt) in zip(reqs, kinds):
        r.geom = gfc.Geometry(ns, nt)
    for r, tol in zip(reqs, (1e-14, 1e-16) * 4):
        r.tolerance = tol
    opts = gfc.Options(precision=gfc.PREC_MIXED)
    with gfc.Server(ctx, max_jobs=16, keep_results=True) as srv:
        summary, comps = serve.run_stream(srv, reqs, opts)
        print({k: summary[k] for k in ("p50_ms", "p99_ms", "throughput_req_per_s")})
        assert summary["all_converged"]
        for c in comps:
            r = reqs[c.tag]
            rc, rep, G_alone = _alone(ctx, r.geom, r.seed, r.eps, serve.request_params(r), opts)
            print(f"{r.geom.ns}^3x{r.geom.nt} tol {r.tolerance:.0e}: sweeps {c.sweeps} "
                  f"(fp32 {c.sweeps_fp32}) latency {(c.t_done - c.t_submit) * 1e3:.0f} ms")
            assert c.status == rc and c.sweeps == rep.sweeps
            assert c.sweeps_fp32 == rep.sweeps_fp32
            assert c.residue_final == rep.residue_final
            G_mix = srv.result_gauge(c.id, r.geom)
            assert np.array_equal(G_mix, G_alone), np.max(np.abs(G_mix - G_alone))
            del G_mix, G_alone
            srv.release(c.id)
