// server.hip -- continuous-batching gauge-fix server (gfc_server_* in gfc.h).
//
// Requests (lattice size, parameters, precision, and either a configuration
// supplied by the caller or a seed and eps to synthesise one) queue on the
// host.  A request is admitted when a job slot and its HBM footprint are free:
// its configuration is copied (or synthesised) into HBM and reunitarized --
// an input with a null row fails with GFC_EPROJECT, as the CLI does -- and its
// fixing state is the one gfc_job_fix keeps (sweep count, check schedule,
// FP32/FP64 phase).
//
// Lattices are served in four lanes by size (fewer than 2^18, 2^20, 2^22
// sites, and larger), one stream each, so that a small lattice is never
// stepped in lockstep with a large one: its half-sweeps take microseconds
// while a large lattice's take a millisecond, and the streams share the GPU.
// The large lane runs each lattice's half-sweeps as gfc_job_fix does, in
// cache-resident timeslice waves over its own set of four streams.
// A lane's poll:
//   1. admits the queued requests of its size class that fit;
//   2. one batched residue evaluation for every lattice at a check, a mixed-
//      precision probe or its start (one synchronisation of the lane);
//   3. the decisions of gaugefixOverrelaxation (src/gauge_fixing.c:771-820)
//      per lattice: stall, convergence, next check, max sweeps,
//      reunitarization; finished lattices are materialised and retired;
//   4. a chunk of half-sweeps: each lattice runs up to its next event (next
//      multiple of sweeps_to_reunitarization, next check, next probe, hand-off,
//      max), and every half-sweep step is ONE launch per precision that packs
//      the 256-site chunks of the lane's lattices (k_sweep_batch).  The chunk's
//      launch sequence is captured in a HIP graph and replayed while the plan
//      repeats.  The large lane instead runs its oldest lattice up to its
//      next check, probe or hand-off (at most kWaveChunk half-sweeps) in
//      timeslice waves, with its reunitarizations folded in, and its other
//      lattices wait their turn.  At most two chunks of a lane are in flight.
// gfc_server_poll serves, in turn, every lane that can proceed without
// waiting for the GPU, so a busy lane never holds up the others.
// Per-site arithmetic and reduction order are those of the single-lattice
// path, so a lattice gets the same result alone or in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <vector>

#include "internal.hpp"

using gfc::BatchJob;
using gfc::cplx;
using gfc::kBlock;
using gfcx::fail;
using gfcx::nblocks;

namespace {

enum State { QUEUED, INIT, RUN, DONE };

constexpr int kLanes = gfc_ctx::kLanes;

// A lane whose runners' u0 and G (192 B per site in FP32, 384 B in FP64) fit
// in this share of the 256 MiB Infinity Cache sweeps with cached loads, so
// that its data stays in the cache from one half-sweep to the next; larger
// lanes stream with non-temporal loads.
constexpr int64_t kLaneCacheBytes = 96ll << 20;

// The lane of the largest lattices, which sweeps in timeslice waves, and the
// most half-sweeps of one lattice in one of its chunks (so that a request
// queued behind it is admitted within about a tenth of a second).
constexpr int kWaveLane = gfc_ctx::kLanes - 1;
constexpr int kWaveChunk = 100;

// Size class of a lattice: its lane.
int lane_of(const gfc_geom &g) {
    const int64_t V = gfc_geom_volume(&g);
    return V < (1 << 18) ? 0 : V < (1 << 20) ? 1 : V < (1 << 22) ? 2 : 3;
}

struct Req {
    int64_t id = 0;
    uint64_t tag = 0;
    gfc_geom geom{0, 0};
    uint64_t seed = 0;
    double eps = 0;
    // a caller-supplied configuration (4V x 18 reals of in_prec bytes), kept
    // until its copy to the device has completed; empty: synthesise from seed
    std::vector<char> host_cfg;
    int in_prec = 0;
    gfc_params p{};
    gfc_options o{};
    double t_submit = 0, t_admit = 0, t_done = 0;
    gfc_job *job = nullptr;
    int64_t footprint = 0;
    int prec = GFC_PREC_FP64;
    int lane = 0;
    State state = QUEUED;
    // fixing state, as in run_fix (engine.hip)
    int sweep = 0, due = 0, evals = 0, sweeps_fp32 = 0, rc = 0;
    double res = 0, res0 = 0, F = 0;
    bool fp32 = false;
    gfcx::Handoff ho;
    int slot = -1;
    double obs[3] = {0, 0, 0};
};

template <typename T>
T *dev_alloc(size_t n) {
    T *p = nullptr;
    if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) return nullptr;
    return p;
}

// One size class: its stream, resident lattices and device-side tables.
struct Lane {
    hipStream_t st = nullptr;
    std::vector<std::unique_ptr<Req>> active;
    BatchJob *d_jobs = nullptr;
    BatchJob *h_jobs = nullptr;  // pinned, max_jobs descriptors by slot
    // sweep items and graphs of the chunk plans, double-buffered: a new plan
    // is written into the buffer the chunk in flight does not use
    int4 *d_items[2] = {nullptr, nullptr}, *h_items[2] = {nullptr, nullptr};
    size_t items_cap[2] = {0, 0};
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    int cur = 0;  // buffer of the current plan
    int4 *d_oitems = nullptr, *h_oitems = nullptr;  // residue-evaluation items
    size_t oitems_cap = 0;
    int *d_sel = nullptr, *h_sel = nullptr;
    double *d_part = nullptr;
    size_t part_cap = 0;
    double *d_out = nullptr, *h_out = nullptr;
    // reunitarization failure flag of each slot's input (set by k_reunit_aos at
    // admission, read back with the slot's first residue)
    int *d_flags = nullptr, *h_flags = nullptr;
    std::vector<int64_t> plan_sig;
    hipEvent_t chunk_ev[2] = {nullptr, nullptr};
    int64_t chunks = 0;
    // timeslice-wave lane: st and three streams of its own, default priority
    bool waves = false;
    gfcx::StreamSet wss;
};

}  // namespace

struct gfc_server {
    gfc_ctx *ctx = nullptr;
    int64_t budget = 0, used = 0;
    int max_jobs = 0, keep = 0;
    std::chrono::steady_clock::time_point t0;
    int64_t next_id = 1;
    std::deque<std::unique_ptr<Req>> queue;
    Lane lanes[kLanes];
    std::map<int64_t, std::unique_ptr<Req>> kept;
    std::multimap<std::tuple<int, int, int>, gfc_job *> cache;  // (ns, nt, prec) -> free job
    std::vector<gfc_completion> done;
    hipEvent_t created = nullptr;  // job creation on ctx->stream, awaited by the lanes
    gfc_server_stats stats{};
    double clock() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    int n_active() const {
        int n = 0;
        for (const Lane &L : lanes) n += (int)L.active.size();
        return n;
    }
};

namespace {

int64_t footprint_of(const Req &r) { return gfc_job_footprint(&r.geom, r.prec); }

int grow_table(Lane &L, int4 *&d, int4 *&h, size_t &cap, size_t n) {
    if (n <= cap) return GFC_OK;
    const size_t c = std::max(n, 2 * cap);
    HIP_TRY(hipStreamSynchronize(L.st));
    hipFree(d);
    hipHostFree(h);
    d = nullptr;
    h = nullptr;
    cap = 0;
    HIP_TRY(hipMalloc(&d, c * sizeof(int4)));
    HIP_TRY(hipHostMalloc(&h, c * sizeof(int4), hipHostMallocDefault));
    cap = c;
    return GFC_OK;
}

int grow_part(Lane &L, size_t blocks) {
    if (blocks <= L.part_cap) return GFC_OK;
    HIP_TRY(hipStreamSynchronize(L.st));
    hipFree(L.d_part);
    L.d_part = nullptr;
    L.part_cap = 0;
    HIP_TRY(hipMalloc(&L.d_part, blocks * 3 * sizeof(double)));
    L.part_cap = blocks;
    return GFC_OK;
}

void fill_desc(Lane &L, Req &r) {
    BatchJob &b = L.h_jobs[r.slot];
    gfc_job *j = r.job;
    b.u0 = r.fp32 ? (const void *)j->u0f : (const void *)j->u0;
    b.g = r.fp32 ? (void *)j->gf : (void *)j->g;
    b.u0d = j->u0;
    b.gd = j->g;
    b.gf = j->gf;
    b.omega = (double)r.p.omega_OR;
    b.ns = j->geom.ns;
    b.Vh = j->Vh;
    b.hits = (int)r.p.hits;
}

int upload_desc(gfc_server *s, Lane &L) {
    HIP_TRY(hipMemcpyAsync(L.d_jobs, L.h_jobs, sizeof(BatchJob) * s->max_jobs,
                           hipMemcpyHostToDevice, L.st));
    return GFC_OK;
}

// Takes a free slot of lane L and a job (from the cache or newly allocated)
// for r.
int admit_one(gfc_server *s, Lane &L, Req &r, bool &admitted) {
    admitted = false;
    if (s->n_active() >= s->max_jobs) return GFC_OK;
    auto key = std::make_tuple(r.geom.ns, r.geom.nt, r.prec);
    auto it = s->cache.find(key);
    gfc_job *job = nullptr;
    if (it != s->cache.end()) {
        job = it->second;
        s->cache.erase(it);
    } else {
        // evict cached jobs of other shapes until the footprint fits
        while (s->used + r.footprint > s->budget && !s->cache.empty()) {
            auto e = s->cache.begin();
            s->used -= gfc_job_footprint(&e->second->geom, e->second->precision);
            gfc_job_destroy(e->second);
            s->cache.erase(e);
        }
        if (s->used + r.footprint > s->budget) return GFC_OK;  // wait for memory
        RC_TRY(gfc_job_create(s->ctx, &r.geom, r.prec, &job));
        s->used += r.footprint;
        // the job's initial G was set on ctx->stream
        HIP_TRY(hipEventRecord(s->created, s->ctx->stream));
        HIP_TRY(hipStreamWaitEvent(L.st, s->created, 0));
    }
    r.job = job;
    // slot: lowest index not used by another request of this lane
    std::vector<char> busy(s->max_jobs, 0);
    for (auto &a : L.active) busy[a->slot] = 1;
    int slot = 0;
    while (busy[slot]) slot++;
    r.slot = slot;
    // input: the caller's configuration or one synthesised in HBM,
    // reunitarized like the CLI does after loading (gauge_fix_coulomb.c:126)
    if (!r.host_cfg.empty()) {
        const int64_t nreal = job->V * 4 * 18;
        if (r.in_prec == 8) {
            HIP_TRY(hipMemcpyAsync(job->cfg, r.host_cfg.data(), nreal * sizeof(double),
                                   hipMemcpyHostToDevice, L.st));
        } else {
            // FP32 staging in the u0 planes (rebuilt below), as gfc_job_upload_config
            float *stage = reinterpret_cast<float *>(job->u0);
            HIP_TRY(hipMemcpyAsync(stage, r.host_cfg.data(), nreal * sizeof(float),
                                   hipMemcpyHostToDevice, L.st));
            hipLaunchKernelGGL(gfc::k_f32_to_f64, dim3(nblocks(nreal)), dim3(kBlock), 0, L.st,
                               stage, reinterpret_cast<double *>(job->cfg), nreal);
            LAUNCH_CHECK();
        }
    } else {
        RC_TRY(gfcx::generate_smooth_device(job, L.st, r.seed, r.eps));
    }
    HIP_TRY(hipMemsetAsync(L.d_flags + slot, 0, sizeof(int), L.st));
    hipLaunchKernelGGL(gfc::k_reunit_aos, dim3(nblocks(job->V * 4)), dim3(kBlock), 0, L.st,
                       job->cfg, job->V * 4, L.d_flags + slot);
    LAUNCH_CHECK();
    RC_TRY(gfcx::build_u0(job, L.st));
    RC_TRY(gfcx::set_identity(job, L.st));
    r.fp32 = r.o.precision == GFC_PREC_MIXED;
    r.state = INIT;
    r.t_admit = s->clock();
    fill_desc(L, r);
    admitted = true;
    s->stats.admitted++;
    return GFC_OK;
}

// Residues of every request in `which` (one batched launch + one sync of L).
int evaluate(gfc_server *s, Lane &L, std::vector<Req *> &which) {
    if (which.empty()) return GFC_OK;
    size_t nblk = 0;
    for (Req *r : which) nblk += nblocks(2 * (int64_t)r->job->Vh);
    RC_TRY(grow_part(L, nblk));
    RC_TRY(grow_table(L, L.d_oitems, L.h_oitems, L.oitems_cap, nblk));
    HIP_TRY(hipStreamSynchronize(L.st));  // pinned staging is reused
    size_t k = 0;
    for (size_t q = 0; q < which.size(); q++) {
        Req *r = which[q];
        L.h_jobs[r->slot].part_off = (int)k;
        const int nb = (int)nblocks(2 * (int64_t)r->job->Vh);
        // the initial residue reads the FP64 G, like run_fix (engine.hip)
        const int g32 = r->state == RUN && r->fp32 ? 1 : 0;
        for (int b = 0; b < nb; b++) L.h_oitems[k++] = make_int4(r->slot, b, g32, 0);
        L.h_sel[q] = r->slot;
    }
    hipStream_t st = L.st;
    RC_TRY(upload_desc(s, L));
    HIP_TRY(hipMemcpyAsync(L.d_oitems, L.h_oitems, k * sizeof(int4), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(L.d_sel, L.h_sel, which.size() * sizeof(int), hipMemcpyHostToDevice,
                           st));
    hipLaunchKernelGGL(gfc::k_obs_batch, dim3((unsigned)k), dim3(kBlock), 0, st, L.d_jobs,
                       L.d_oitems, L.d_part);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(gfc::k_reduce_batch, dim3((unsigned)which.size()), dim3(kBlock), 0, st,
                       L.d_jobs, L.d_sel, L.d_part, L.d_out);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(L.h_out, L.d_out, 3 * which.size() * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(L.h_flags, L.d_flags, s->max_jobs * sizeof(int),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (size_t q = 0; q < which.size(); q++) {
        Req *r = which[q];
        const double V = (double)r->job->V;
        r->obs[0] = L.h_out[3 * q] / V;
        r->obs[1] = L.h_out[3 * q + 1] / (3.0 * V);
        r->obs[2] = L.h_out[3 * q + 2] / (9.0 * V);
        r->evals++;
    }
    s->stats.residue_batches++;
    return GFC_OK;
}

int promote(Lane &L, Req &r) {
    RC_TRY(gfcx::promote_g_only(r.job, L.st));
    r.fp32 = false;
    fill_desc(L, r);
    L.plan_sig.clear();
    return GFC_OK;
}

// End of gaugefixOverrelaxation for r: hand-off, final reunitarization (unless
// it stopped on max sweeps or a stalled residue), U = G U0 G^dagger.
int finish(Lane &L, Req &r, int rc) {
    if (r.fp32) RC_TRY(promote(L, r));
    if (rc >= 0) RC_TRY(gfcx::reunit_g(r.job, L.st, false, 0, 2));
    RC_TRY(gfcx::materialize(r.job, L.st, rc >= 0));
    r.rc = rc >= 0 ? r.sweep : rc;
    r.state = DONE;
    return GFC_OK;
}

// A request that cannot be fixed (its input could not be reunitarized): no
// sweep, no result, the error code as its status.
void fail_request(Req &r, int code) {
    r.rc = code;
    r.state = DONE;
}

int next_multiple(int s, int m) { return (s / m + 1) * m; }

// A chunk of the timeslice-wave lane: its oldest running lattice sweeps up to
// its next check, probe, hand-off or max sweeps (at most kWaveChunk
// half-sweeps) in timeslice waves on the lane's streams.  The lane serves its
// lattices first come, first served: a lattice admitted behind another waits
// until that one is done, rather than sharing the lane with it chunk by chunk,
// which made both late (the 256-request serving stream: p99 1.035 -> 0.991 s,
// mean latency 173 -> 166 ms).
int run_waves(gfc_server *s, Lane &L) {
    Req *run = nullptr;
    for (auto &a : L.active)  // in order of admission
        if (a->state == RUN) {
            run = a.get();
            break;
        }
    if (run == nullptr) return GFC_OK;
    Req &r = *run;
    int64_t stop = (int64_t)r.sweep + kWaveChunk;
    if (r.due > r.sweep) stop = std::min<int64_t>(stop, r.due);
    if ((int64_t)r.p.max_sweeps_to_fix > r.sweep)
        stop = std::min<int64_t>(stop, (int64_t)r.p.max_sweeps_to_fix);
    if (r.fp32) {
        if (!r.ho.at_floor) stop = std::min(stop, (int64_t)next_multiple(r.sweep, r.o.mixed_probe_every));
        if (r.ho.switch_at > r.sweep) stop = std::min<int64_t>(stop, r.ho.switch_at);
    }
    hipEvent_t &ev = L.chunk_ev[L.chunks % 2];
    HIP_TRY(hipEventSynchronize(ev));
    int group = gfcx::group_size(r.job, r.o, r.fp32);
    if (group <= 0) group = r.job->geom.nt;  // fits the cache whole: one group
    RC_TRY(gfcx::sweep_blocked(r.job, L.wss, r.fp32, r.sweep, (int)stop, group,
                               gfcx::lane_count(r.job, r.o, r.fp32), (double)r.p.omega_OR,
                               (int)r.p.hits, r.p.sweeps_to_reunitarization));
    HIP_TRY(hipEventRecord(ev, L.st));
    L.chunks++;
    s->stats.chunks++;
    const int n = (int)stop - r.sweep;
    r.sweep = (int)stop;
    if (r.fp32) r.sweeps_fp32 += n;
    s->stats.site_updates += (int64_t)n * r.job->Vh;
    return GFC_OK;
}

int plan_and_run(gfc_server *s, Lane &L) {
    if (L.waves) return run_waves(s, L);
    // stop points: next reunitarization point, next check, next probe,
    // hand-off, max
    std::vector<Req *> run;
    std::vector<int> steps;
    int K = 0;
    for (auto &a : L.active) {
        if (a->state != RUN) continue;
        Req &r = *a;
        const int R = (int)r.p.sweeps_to_reunitarization;
        int stop = next_multiple(r.sweep, R);
        if (r.due > r.sweep) stop = std::min(stop, r.due);
        if ((int64_t)r.p.max_sweeps_to_fix > r.sweep)
            stop = (int)std::min<int64_t>(stop, (int64_t)r.p.max_sweeps_to_fix);
        if (r.fp32) {
            if (!r.ho.at_floor) stop = std::min(stop, next_multiple(r.sweep, r.o.mixed_probe_every));
            if (r.ho.switch_at > r.sweep) stop = std::min(stop, r.ho.switch_at);
        }
        run.push_back(&r);
        steps.push_back(stop - r.sweep);
        K = std::max(K, stop - r.sweep);
    }
    if (run.empty()) return GFC_OK;
    int64_t bytes = 0;
    for (Req *r : run) bytes += r->job->V * (r->fp32 ? 192 : 384);
    const bool cached = bytes <= kLaneCacheBytes;
    // plan signature: load policy, then (slot, job, fp32, steps, parity) of
    // every runner
    std::vector<int64_t> sig{cached};
    for (size_t q = 0; q < run.size(); q++) {
        sig.push_back(run[q]->slot);
        sig.push_back((int64_t)(intptr_t)run[q]->job);
        sig.push_back(run[q]->fp32);
        sig.push_back(steps[q]);
        sig.push_back(run[q]->sweep & 1);
    }
    hipStream_t st = L.st;
    if (sig != L.plan_sig || L.graph[L.cur] == nullptr) {
        // the other buffer: its plan's chunks have completed (the lane is
        // served only once its older chunk is done, and the newer one runs
        // the current plan)
        const int b = L.cur ^ 1;
        // item table: for each step k, the 64-bit then the 32-bit runners
        std::vector<std::pair<size_t, size_t>> seg;  // per (step, precision): offset, count
        size_t total = 0;
        for (int k = 0; k < K; k++)
            for (int prec = 0; prec < 2; prec++)
                for (size_t q = 0; q < run.size(); q++)
                    if (steps[q] > k && (int)run[q]->fp32 == prec)
                        total += nblocks(run[q]->job->Vh);
        RC_TRY(grow_table(L, L.d_items[b], L.h_items[b], L.items_cap[b], total));
        int4 *h_items = L.h_items[b], *d_items = L.d_items[b];
        size_t n = 0;
        for (int k = 0; k < K; k++)
            for (int prec = 0; prec < 2; prec++) {
                const size_t off = n;
                for (size_t q = 0; q < run.size(); q++) {
                    Req &r = *run[q];
                    if (steps[q] <= k || (int)r.fp32 != prec) continue;
                    const int par = (r.sweep + k) & 1;
                    for (int h0 = 0; h0 < r.job->Vh; h0 += kBlock)
                        h_items[n++] = make_int4(r.slot, h0, par, 0);
                }
                seg.emplace_back(off, n - off);
            }
        RC_TRY(upload_desc(s, L));
        HIP_TRY(hipMemcpyAsync(d_items, h_items, n * sizeof(int4), hipMemcpyHostToDevice, st));
        if (L.graph[b]) {
            hipGraphExecDestroy(L.graph[b]);
            L.graph[b] = nullptr;
        }
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (size_t i = 0; i < seg.size(); i++) {
            const size_t off = seg[i].first, cnt = seg[i].second;
            if (cnt == 0) continue;
            const dim3 grid((unsigned)cnt);
            if (i % 2 == 0 && cached)
                hipLaunchKernelGGL((gfc::k_sweep_batch<double, false>), grid, dim3(kBlock), 0,
                                   st, L.d_jobs, d_items + off, (int)cnt);
            else if (i % 2 == 0)
                hipLaunchKernelGGL((gfc::k_sweep_batch<double, true>), grid, dim3(kBlock), 0,
                                   st, L.d_jobs, d_items + off, (int)cnt);
            else if (cached)
                hipLaunchKernelGGL((gfc::k_sweep_batch<float, false>), grid, dim3(kBlock), 0, st,
                                   L.d_jobs, d_items + off, (int)cnt);
            else
                hipLaunchKernelGGL((gfc::k_sweep_batch<float, true>), grid, dim3(kBlock), 0, st,
                                   L.d_jobs, d_items + off, (int)cnt);
        }
        hipError_t ce = hipStreamEndCapture(st, &g);
        if (ce != hipSuccess) return fail(GFC_EDEVICE, std::string("capture: ") + hipGetErrorString(ce));
        ce = hipGraphInstantiate(&L.graph[b], g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ce != hipSuccess) {
            L.graph[b] = nullptr;
            return fail(GFC_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(ce));
        }
        L.cur = b;
        L.plan_sig = sig;
        s->stats.graphs_built++;
    } else {
        s->stats.graph_replays++;
    }
    // at most two chunks of the lane in flight
    hipEvent_t &ev = L.chunk_ev[L.chunks % 2];
    HIP_TRY(hipEventSynchronize(ev));
    HIP_TRY(hipGraphLaunch(L.graph[L.cur], st));
    HIP_TRY(hipEventRecord(ev, st));
    L.chunks++;
    s->stats.chunks++;
    for (size_t q = 0; q < run.size(); q++) {
        run[q]->sweep += steps[q];
        if (run[q]->fp32) run[q]->sweeps_fp32 += steps[q];
        s->stats.site_updates += (int64_t)steps[q] * run[q]->job->Vh;
    }
    return GFC_OK;
}

int complete(gfc_server *s, std::unique_ptr<Req> &r) {
    gfc_completion c;
    std::memset(&c, 0, sizeof c);
    c.id = r->id;
    c.tag = r->tag;
    c.status = r->rc;
    c.sweeps = r->sweep;
    c.sweeps_fp32 = r->sweeps_fp32;
    c.residue_initial = r->res0;
    c.residue_final = r->res;
    c.F = r->F;
    c.t_submit = r->t_submit;
    c.t_admit = r->t_admit;
    c.t_done = r->t_done;
    c.ns = r->geom.ns;
    c.nt = r->geom.nt;
    s->done.push_back(c);
    s->stats.completed++;
    if (s->keep) {
        s->kept[r->id] = std::move(r);
    } else {
        s->cache.emplace(std::make_tuple(r->geom.ns, r->geom.nt, r->prec), r->job);
        r->job = nullptr;
    }
    return GFC_OK;
}

// Whether a lattice needs its residue at its current sweep.
bool needs_residue(const Req &r) {
    if (r.state == INIT) return true;
    if (r.state != RUN) return false;
    const bool check = r.sweep == r.due;
    return check || (r.fp32 && !r.ho.at_floor && r.sweep % r.o.mixed_probe_every == 0);
}

// Whether its decisions at the current sweep need the lane idle: a residue,
// or a change of its descriptor (hand-off, end on max sweeps), whose pinned
// copy an upload still queued behind the chunk in flight may read.
bool needs_idle(const Req &r) {
    if (needs_residue(r)) return true;
    return r.state == RUN && ((r.fp32 && r.sweep >= r.ho.switch_at) ||
                              (int64_t)r.sweep >= (int64_t)r.p.max_sweeps_to_fix);
}

bool has_queued(const gfc_server *s, int lane) {
    for (auto &q : s->queue)
        if (q->lane == lane) return true;
    return false;
}

// One poll of lane li (steps 1-4 of the header).
int poll_lane(gfc_server *s, int li) {
    Lane &L = s->lanes[li];
    // 1. admission (FIFO within the size class, by footprint and slot)
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        if ((*it)->lane != li) {
            ++it;
            continue;
        }
        bool ok = false;
        RC_TRY(admit_one(s, L, **it, ok));
        if (!ok) break;
        L.active.push_back(std::move(*it));
        it = s->queue.erase(it);
    }
    // 2. residues for starts, checks and probes
    std::vector<Req *> ev;
    for (auto &a : L.active)
        if (needs_residue(*a)) ev.push_back(a.get());
    RC_TRY(evaluate(s, L, ev));
    // 3. decisions (the order of gaugefixOverrelaxation's loop body)
    for (auto &a : L.active) {
        Req &r = *a;
        const int which = r.p.gfix_proxy == GFC_PROXY_THETA ? 1 : 0;
        if (r.state == INIT) {
            // the input's copy to the device has completed (evaluate synchronised
            // the lane)
            std::vector<char>().swap(r.host_cfg);
            if (L.h_flags[r.slot]) {
                fail_request(r, GFC_EPROJECT);
                continue;
            }
            r.res = r.res0 = r.obs[which];
            r.F = r.obs[2];
            r.due = gfc_next_check(0, r.res, &r.p);
            r.ho.start(r.res);
            r.state = RUN;
            if (!(r.res > r.p.tolerance)) RC_TRY(finish(L, r, 0));
            continue;
        }
        if (r.state != RUN || r.sweep == 0) continue;
        const int s_ = r.sweep;
        const int R = (int)r.p.sweeps_to_reunitarization;
        const bool at_stop = s_ % R == 0 || s_ == r.due ||
                             (int64_t)s_ >= (int64_t)r.p.max_sweeps_to_fix ||
                             (r.fp32 && ((!r.ho.at_floor && s_ % r.o.mixed_probe_every == 0) ||
                                         s_ >= r.ho.switch_at));
        if (!at_stop) continue;
        if (r.fp32 && !r.ho.at_floor && s_ != r.due && s_ % r.o.mixed_probe_every == 0) {
            if (r.ho.observe(r.o, r.p.tolerance, s_, r.obs[which], r.due)) RC_TRY(promote(L, r));
        }
        if (s_ == r.due) {
            const double last = r.res;
            r.res = r.obs[which];
            r.F = r.obs[2];
            if (std::fabs(r.res - last) / last < 10E-16) {
                RC_TRY(finish(L, r, GFC_ENOTCONVERGED));
                continue;
            } else if (r.res <= r.p.tolerance) {
                RC_TRY(finish(L, r, 0));
                continue;
            } else {
                r.due = gfc_next_check((unsigned)s_, r.res, &r.p);
                if (r.fp32 && r.ho.observe(r.o, r.p.tolerance, s_, r.res, r.due))
                    RC_TRY(promote(L, r));
            }
        }
        if (r.fp32 && s_ >= r.ho.switch_at) RC_TRY(promote(L, r));
        if ((unsigned)s_ >= r.p.max_sweeps_to_fix) {
            RC_TRY(finish(L, r, GFC_ENOTCONVERGED));
            continue;
        }
        if (!((unsigned)s_ % r.p.sweeps_to_reunitarization))
            RC_TRY(gfcx::reunit_g(r.job, L.st, r.fp32, 0, 2));
    }
    // retire finished requests (their results are materialised in HBM)
    bool any_done = false;
    for (auto &a : L.active) any_done |= a->state == DONE;
    if (any_done) {
        HIP_TRY(hipStreamSynchronize(L.st));
        const double t = s->clock();
        for (auto &a : L.active)
            if (a->state == DONE) {
                a->t_done = t;
                RC_TRY(complete(s, a));
            }
        L.active.erase(std::remove_if(L.active.begin(), L.active.end(),
                                      [](const std::unique_ptr<Req> &a) {
                                          return !a || a->state == DONE || a->job == nullptr;
                                      }),
                       L.active.end());
        L.plan_sig.clear();
    }
    // 4. one chunk of packed half-sweeps
    return plan_and_run(s, L);
}

// Serves every lane that can go on without waiting: one with queued requests
// or a lattice that needs the lane idle waits for all its chunks, any other
// one for its older chunk.  Sleeps a moment when none can.
int poll_once(gfc_server *s) {
    bool served = false;
    for (int li = 0; li < kLanes; li++) {
        Lane &L = s->lanes[li];
        const bool queued = has_queued(s, li);
        if (L.active.empty() && !queued) continue;
        bool sync = queued;
        for (auto &a : L.active) sync |= needs_idle(*a);
        // chunk_ev[chunks % 2] marks the older chunk in flight, the other the newer
        hipEvent_t e = L.chunk_ev[(L.chunks + (sync ? 1 : 0)) % 2];
        const hipError_t q = hipEventQuery(e);
        if (q == hipErrorNotReady) continue;
        if (q != hipSuccess) return fail(GFC_EDEVICE, hipGetErrorString(q));
        RC_TRY(poll_lane(s, li));
        served = true;
    }
    if (!served) std::this_thread::sleep_for(std::chrono::microseconds(20));
    return GFC_OK;
}

void free_lane(Lane &L) {
    if (L.st) hipStreamSynchronize(L.st);
    for (int b = 0; b < 2; b++) {
        if (L.graph[b]) hipGraphExecDestroy(L.graph[b]);
        hipFree(L.d_items[b]);
        hipHostFree(L.h_items[b]);
    }
    for (hipEvent_t e : L.chunk_ev)
        if (e) hipEventDestroy(e);
    // the timeslice-wave lane's own streams (lane 0 of the set is L.st)
    for (int l = 1; l < gfc_ctx::kLanes; l++)
        if (L.wss.lane[l]) {
            hipStreamSynchronize(L.wss.lane[l]);
            hipStreamDestroy(L.wss.lane[l]);
        }
    if (L.wss.fork) hipEventDestroy(L.wss.fork);
    for (int l = 0; l < gfc_ctx::kLanes; l++)
        if (L.wss.join[l]) hipEventDestroy(L.wss.join[l]);
    for (auto &a : L.active) gfc_job_destroy(a->job);
    hipFree(L.d_jobs);
    hipFree(L.d_oitems);
    hipFree(L.d_sel);
    hipFree(L.d_part);
    hipFree(L.d_out);
    hipFree(L.d_flags);
    hipHostFree(L.h_flags);
    hipHostFree(L.h_jobs);
    hipHostFree(L.h_oitems);
    hipHostFree(L.h_sel);
    hipHostFree(L.h_out);
}

}  // namespace

extern "C" {

int gfc_server_create(gfc_ctx *ctx, int64_t mem_budget, int max_jobs, int keep_results,
                      gfc_server **out) {
    if (ctx == nullptr || out == nullptr || max_jobs <= 0)
        return fail(GFC_EINVAL, "gfc_server_create: bad argument");
    *out = nullptr;
    HIP_TRY(hipSetDevice(ctx->device));
    std::unique_ptr<gfc_server, void (*)(gfc_server *)> s(new gfc_server(), gfc_server_destroy);
    s->ctx = ctx;
    s->max_jobs = max_jobs;
    s->keep = keep_results;
    s->t0 = std::chrono::steady_clock::now();
    if (mem_budget <= 0) {
        size_t fr = 0, tot = 0;
        HIP_TRY(hipMemGetInfo(&fr, &tot));
        mem_budget = (int64_t)(0.85 * (double)fr);
    }
    s->budget = mem_budget;
    HIP_TRY(hipEventCreateWithFlags(&s->created, hipEventDisableTiming));
    for (int li = 0; li < kLanes; li++) {
        Lane &L = s->lanes[li];
        L.st = ctx->lane[li];
        L.d_jobs = dev_alloc<BatchJob>(max_jobs);
        L.d_sel = dev_alloc<int>(max_jobs);
        L.d_out = dev_alloc<double>(3 * (size_t)max_jobs);
        L.d_flags = dev_alloc<int>(max_jobs);
        if (!L.d_jobs || !L.d_sel || !L.d_out || !L.d_flags)
            return fail(GFC_ENOMEM, "gfc_server_create");
        HIP_TRY(hipHostMalloc(&L.h_flags, sizeof(int) * max_jobs, hipHostMallocDefault));
        std::memset(L.h_flags, 0, sizeof(int) * max_jobs);
        HIP_TRY(hipHostMalloc(&L.h_jobs, sizeof(BatchJob) * max_jobs, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(&L.h_sel, sizeof(int) * max_jobs, hipHostMallocDefault));
        HIP_TRY(hipHostMalloc(&L.h_out, sizeof(double) * 3 * max_jobs, hipHostMallocDefault));
        std::memset(L.h_jobs, 0, sizeof(BatchJob) * max_jobs);
        for (int k = 0; k < 2; k++) {
            HIP_TRY(hipEventCreateWithFlags(&L.chunk_ev[k], hipEventDisableTiming));
            HIP_TRY(hipEventRecord(L.chunk_ev[k], L.st));
        }
        if (li == kWaveLane) {
            L.waves = true;
            L.wss.lane[0] = L.st;
            for (int l = 1; l < gfc_ctx::kLanes; l++)
                HIP_TRY(hipStreamCreateWithFlags(&L.wss.lane[l], hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&L.wss.fork, hipEventDisableTiming));
            for (int l = 0; l < gfc_ctx::kLanes; l++)
                HIP_TRY(hipEventCreateWithFlags(&L.wss.join[l], hipEventDisableTiming));
        }
    }
    *out = s.release();
    return GFC_OK;
}

void gfc_server_destroy(gfc_server *s) {
    if (s == nullptr) return;
    hipSetDevice(s->ctx->device);
    for (Lane &L : s->lanes) free_lane(L);
    hipStreamSynchronize(s->ctx->stream);
    for (auto &kv : s->kept) gfc_job_destroy(kv.second->job);
    for (auto &kv : s->cache) gfc_job_destroy(kv.second);
    if (s->created) hipEventDestroy(s->created);
    delete s;
}

double gfc_server_clock(gfc_server *s) { return s ? s->clock() : 0.0; }

}  // extern "C"

namespace {

// The checks and queueing shared by both submit calls; fill sets the input.
template <typename Fill>
int64_t submit_req(gfc_server *s, const gfc_geom *g, const gfc_params *p, const gfc_options *o,
                   uint64_t tag, const char *who, Fill fill) {
    if (s == nullptr || g == nullptr || p == nullptr)
        return fail(GFC_EINVAL, std::string(who) + ": NULL argument");
    if (gfc_geom_check(g)) return fail(GFC_EGEOMETRY, std::string(who) + ": bad geometry");
    if (!gfc_params_valid(p)) return fail(GFC_EPARAMS, std::string(who) + ": invalid parameters");
    auto r = std::make_unique<Req>();
    r->geom = *g;
    r->p = *p;
    if (o)
        r->o = *o;
    else
        gfc_options_default(&r->o);
    if (r->o.mixed_probe_every <= 0) r->o.mixed_probe_every = 50;
    r->prec = r->o.precision == GFC_PREC_MIXED ? GFC_PREC_MIXED : GFC_PREC_FP64;
    r->lane = lane_of(r->geom);
    r->footprint = footprint_of(*r);
    if (r->footprint > s->budget)
        return fail(GFC_ENOMEM, std::string(who) + ": lattice larger than the memory budget");
    RC_TRY(fill(*r));
    r->id = s->next_id++;
    r->tag = tag;
    r->t_submit = s->clock();
    const int64_t id = r->id;
    s->queue.push_back(std::move(r));
    s->stats.submitted++;
    return id;
}

}  // namespace

extern "C" {

int64_t gfc_server_submit(gfc_server *s, const gfc_geom *g, uint64_t seed, double eps,
                          const gfc_params *p, const gfc_options *o, uint64_t tag) {
    return submit_req(s, g, p, o, tag, "gfc_server_submit", [&](Req &r) {
        r.seed = seed;
        r.eps = eps;
        return GFC_OK;
    });
}

int64_t gfc_server_submit_config(gfc_server *s, const gfc_geom *g, const void *U,
                                 int in_precision, const gfc_params *p, const gfc_options *o,
                                 uint64_t tag) {
    return submit_req(s, g, p, o, tag, "gfc_server_submit_config", [&](Req &r) {
        if (U == nullptr) return fail(GFC_EINVAL, "gfc_server_submit_config: NULL U");
        if (in_precision != 4 && in_precision != 8)
            return fail(GFC_EINVAL, "gfc_server_submit_config: precision must be 4 or 8");
        const size_t bytes = (size_t)gfc_geom_volume(g) * 4 * 18 * (size_t)in_precision;
        try {
            r.host_cfg.assign(static_cast<const char *>(U), static_cast<const char *>(U) + bytes);
        } catch (const std::bad_alloc &) {
            return fail(GFC_ENOMEM, "gfc_server_submit_config: host copy of the input");
        }
        r.in_prec = in_precision;
        return GFC_OK;
    });
}

int gfc_server_poll(gfc_server *s, gfc_completion *out, int max_out) {
    if (s == nullptr || (out == nullptr && max_out > 0))
        return fail(GFC_EINVAL, "gfc_server_poll: bad argument");
    HIP_TRY(hipSetDevice(s->ctx->device));
    if (s->done.empty()) RC_TRY(poll_once(s));
    const int n = std::min<int>(max_out, (int)s->done.size());
    for (int i = 0; i < n; i++) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + n);
    return n;
}

int gfc_server_pending(gfc_server *s) {
    return s ? (int)(s->queue.size() + s->n_active() + s->done.size()) : 0;
}

int gfc_server_get_stats(gfc_server *s, gfc_server_stats *st) {
    if (s == nullptr || st == nullptr) return fail(GFC_EINVAL, "gfc_server_get_stats: NULL");
    *st = s->stats;
    st->queued = (int)s->queue.size();
    st->active = s->n_active();
    st->bytes_used = s->used;
    st->bytes_budget = s->budget;
    return GFC_OK;
}

int gfc_server_result_gauge(gfc_server *s, int64_t id, double *G) {
    if (s == nullptr || G == nullptr) return fail(GFC_EINVAL, "gfc_server_result_gauge: NULL");
    auto it = s->kept.find(id);
    if (it == s->kept.end()) return fail(GFC_EINVAL, "gfc_server_result_gauge: unknown id");
    return gfc_job_download_gauge(it->second->job, G);
}

int gfc_server_release(gfc_server *s, int64_t id) {
    if (s == nullptr) return fail(GFC_EINVAL, "gfc_server_release: NULL");
    auto it = s->kept.find(id);
    if (it == s->kept.end()) return fail(GFC_EINVAL, "gfc_server_release: unknown id");
    Req &r = *it->second;
    s->cache.emplace(std::make_tuple(r.geom.ns, r.geom.nt, r.prec), r.job);
    s->kept.erase(it);
    return GFC_OK;
}

}  // extern "C"
