"""Request stream for the continuous-batching server (gfc.Server).

`stream_requests` draws the serving workload: Poisson arrivals, lattice sizes
from a weighted set, eps uniform in a range, tolerance from a set; every
request is an independent seeded smooth configuration.  `run_stream` replays
the arrival times in real time against a server, polling it between arrivals,
and measures each request's latency from its scheduled arrival to the moment
its gauge transformation is materialised in HBM.  `lone_service_times`
measures what each request class costs alone on the idle GPU, from which
`utilisation` gives the offered load rho = lambda E[S].
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np

from . import Geometry, Job, Options, Params, PREC_MIXED, Server

SERVING_MIX_SIZES = ((16, 32, 0.4), (24, 48, 0.3), (32, 64, 0.2), (48, 96, 0.1))


@dataclass
class Request:
    idx: int
    t_arrival: float
    geom: Geometry
    seed: int
    eps: float
    tolerance: float


def stream_requests(n=256, rate=4.0, sizes=SERVING_MIX_SIZES, eps_range=(0.2, 0.4),
                    tolerances=(1e-14, 1e-16), seed=4, base_seed=1000):
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n)
    t = np.cumsum(gaps) - gaps[0]
    probs = np.array([s[2] for s in sizes])
    kinds = rng.choice(len(sizes), size=n, p=probs / probs.sum())
    eps = rng.uniform(*eps_range, size=n)
    tols = rng.choice(tolerances, size=n)
    return [Request(i, float(t[i]), Geometry(sizes[kinds[i]][0], sizes[kinds[i]][1]),
                    base_seed + i, float(eps[i]), float(tols[i])) for i in range(n)]


def request_params(r: Request) -> Params:
    return Params(tolerance=r.tolerance, sweeps_to_reunitarization=10, omega_OR=1.95, hits=2,
                  estimate_sweeps_to_gf_progress=1000)


def run_stream(server: Server, reqs, options: Options | None = None, log=None):
    """Submits reqs at their arrival times (relative to now) and polls the server
    until all have completed.  Returns a summary dict and the completions."""
    options = options or Options(precision=PREC_MIXED)
    t0 = server.clock()
    by_tag = {}
    i = 0
    n = len(reqs)
    last_log = 0.0
    while len(by_tag) < n:
        now = server.clock() - t0
        while i < n and reqs[i].t_arrival <= now:
            r = reqs[i]
            server.submit(r.geom, r.seed, r.eps, request_params(r), options, tag=r.idx)
            i += 1
        if server.pending() == 0:
            if i < n:
                time.sleep(min(0.005, max(0.0, reqs[i].t_arrival - now)))
            continue
        for c in server.poll():
            by_tag[c.tag] = c
        if log and now - last_log > 10.0:
            last_log = now
            log(f"serve: t={now:6.1f}s submitted={i} done={len(by_tag)} "
                f"pending={server.pending()}")
    comps = [by_tag[k] for k in range(n)]
    return summarise(reqs, comps, t0, server.stats()), comps


def summarise(reqs, comps, t0: float, stats: dict) -> dict:
    """Latency summary of a served stream: comps[k] is the completion of reqs[k]
    (times on the server clock, the stream started at t0).  Latency is the
    completion time minus the scheduled arrival; by_size splits it by lattice."""
    n = len(reqs)
    lat = np.array([c.t_done - (t0 + reqs[c.tag].t_arrival) for c in comps])
    by_size = {}
    for key in sorted({(c.ns, c.nt) for c in comps}):
        l = np.array([lat[i] for i, c in enumerate(comps) if (c.ns, c.nt) == key])
        by_size[f"{key[0]}^3x{key[1]}"] = {"requests": int(l.size),
                                          "p50_ms": float(np.percentile(l, 50) * 1e3),
                                          "p99_ms": float(np.percentile(l, 99) * 1e3)}
    makespan = max(c.t_done for c in comps) - t0
    upd = sum(c.sweeps * (c.ns ** 3 * c.nt // 2) for c in comps)
    return {
        "requests": n,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "mean_ms": float(lat.mean() * 1e3),
        "max_ms": float(lat.max() * 1e3),
        "makespan_s": makespan,
        "throughput_req_per_s": n / makespan,
        "offered_req_per_s": n / max(reqs[-1].t_arrival, 1e-9),
        "site_updates_per_s": upd / makespan,
        "all_converged": all(c.status >= 0 and c.residue_final <= reqs[c.tag].tolerance
                             for c in comps),
        "chunks": stats["chunks"], "graph_replays": stats["graph_replays"],
        "graphs_built": stats["graphs_built"],
        "by_size": by_size,
    }


def lone_service_times(ctx, reqs, options: Options | None = None, reps: int = 2):
    """Service time of each (lattice size, tolerance) class of reqs when its fix
    runs alone on the idle GPU: gfc_job_fix of the class's first request
    (after one warm-up fix), with the server's per-request options.  Returns
    {(ns, nt, tolerance): (seconds, sweeps)}."""
    options = options or Options(precision=PREC_MIXED)
    out = {}
    for r in reqs:
        key = (r.geom.ns, r.geom.nt, r.tolerance)
        if key in out:
            continue
        with Job(ctx, r.geom, options.precision) as job:
            job.generate_smooth(r.seed, r.eps)
            job.reunitarize()
            job.save_config()
            best = None
            for _ in range(reps):
                job.restore_config()
                t = time.perf_counter()
                rc, rep = job.fix(request_params(r), options)
                dt = time.perf_counter() - t
                best = dt if best is None else min(best, dt)
            out[key] = (best, rep.sweeps)
    return out


def utilisation(reqs, comps, lone) -> dict:
    """Offered load of the stream against the engine's lone-request capacity.

    rho = lambda E[S]: arrivals per second over the stream times the mean of
    each request's lone service time S (its class's lone fix time, scaled by
    its own sweep count).  rho < 1 means a server that ran one request at a
    time at the lone speed would keep up; the scheduler also overlaps
    requests, so it can sustain somewhat more.  capacity_site_updates_per_s is
    the lone engine's rate on this mix: the site updates the requests needed
    over the sum of their lone service times."""
    span = max(reqs[-1].t_arrival - reqs[0].t_arrival, 1e-9)
    lam = (len(reqs) - 1) / span
    S, work = [], 0.0
    for c in comps:
        r = reqs[c.tag]
        sec, sweeps = lone[(r.geom.ns, r.geom.nt, r.tolerance)]
        S.append(sec * c.sweeps / max(sweeps, 1))
        work += c.sweeps * (c.ns ** 3 * c.nt // 2)
    ES = float(np.mean(S))
    return {"rho": lam * ES, "arrival_rate_per_s": lam, "mean_lone_service_ms": ES * 1e3,
            "offered_site_updates_per_s": work / span,
            "capacity_site_updates_per_s": work / float(np.sum(S))}
