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
