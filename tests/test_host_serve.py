"""Host logic of the serving workload (gfc.serve): the request stream of the serving mix,
the latency summary and the offered load rho = lambda E[S], on synthetic
completions (no GPU)."""
from types import SimpleNamespace

import numpy as np
import pytest

serve = pytest.importorskip("gfc.serve")


def test_stream_requests_follow_the_serving_mix():
    reqs = serve.stream_requests(n=4000, rate=4.0)
    t = np.array([r.t_arrival for r in reqs])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    # Poisson arrivals at 4/s: mean gap 0.25 s
    assert abs(np.diff(t).mean() - 0.25) < 0.02
    sizes = [(r.geom.ns, r.geom.nt) for r in reqs]
    for ns, nt, w in serve.SERVING_MIX_SIZES:
        assert abs(sizes.count((ns, nt)) / len(reqs) - w) < 0.03
    assert all(0.2 <= r.eps <= 0.4 for r in reqs)
    assert {r.tolerance for r in reqs} == {1e-14, 1e-16}
    # the same seed gives the same stream
    again = serve.stream_requests(n=4000, rate=4.0)
    assert [(r.t_arrival, r.seed, r.eps) for r in again] == [(r.t_arrival, r.seed, r.eps)
                                                             for r in reqs]


def _comp(tag, ns, nt, t_done, sweeps=1085, status=1085, res=1e-17):
    return SimpleNamespace(tag=tag, ns=ns, nt=nt, t_done=t_done, sweeps=sweeps,
                           status=status, residue_final=res)


def test_summary_latency_and_by_size():
    reqs = [SimpleNamespace(t_arrival=a, tolerance=1e-16) for a in (0.0, 1.0, 2.0, 3.0)]
    t0 = 10.0
    comps = [_comp(0, 16, 32, 10.1), _comp(1, 48, 96, 11.5),
             _comp(2, 16, 32, 12.3), _comp(3, 48, 96, 14.0)]
    st = {"chunks": 7, "graph_replays": 5, "graphs_built": 2}
    s = serve.summarise(reqs, comps, t0, st)
    lat = np.array([0.1, 0.5, 0.3, 1.0])
    assert s["p50_ms"] == pytest.approx(np.percentile(lat, 50) * 1e3)
    assert s["p99_ms"] == pytest.approx(np.percentile(lat, 99) * 1e3)
    assert s["max_ms"] == pytest.approx(1000.0)
    assert s["makespan_s"] == pytest.approx(4.0)
    assert s["by_size"]["16^3x32"]["requests"] == 2
    assert s["by_size"]["48^3x96"]["p50_ms"] == pytest.approx(750.0)
    assert s["all_converged"]
    comps[1].residue_final = 1e-15  # above its tolerance
    assert not serve.summarise(reqs, comps, t0, st)["all_converged"]


def test_utilisation_rho():
    geoms = [SimpleNamespace(ns=16, nt=32), SimpleNamespace(ns=48, nt=96)]
    reqs = [SimpleNamespace(t_arrival=float(i), geom=geoms[i % 2], tolerance=1e-16)
            for i in range(5)]
    comps = [_comp(i, reqs[i].geom.ns, reqs[i].geom.nt, 0.0, sweeps=1085 if i % 2 else 790)
             for i in range(5)]
    lone = {(16, 32, 1e-16): (0.010, 790), (48, 96, 1e-16): (0.500, 1085)}
    u = serve.utilisation(reqs, comps, lone)
    # 4 gaps over 4 s: lambda = 1/s; E[S] = (3 x 0.010 + 2 x 0.500) / 5
    assert u["arrival_rate_per_s"] == pytest.approx(1.0)
    assert u["mean_lone_service_ms"] == pytest.approx((3 * 10 + 2 * 500) / 5)
    assert u["rho"] == pytest.approx(u["mean_lone_service_ms"] / 1e3)
    work = 3 * 790 * 16 ** 3 * 32 // 2 + 2 * 1085 * 48 ** 3 * 96 // 2
    assert u["capacity_site_updates_per_s"] == pytest.approx(work / (3 * 0.010 + 2 * 0.500))
