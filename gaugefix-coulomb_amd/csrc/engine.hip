// engine.hip -- C ABI of the Coulomb gauge-fixing engine (see include/gfc.h).
//
// One gfc_ctx owns four HIP streams on one device; one gfc_job owns the
// device fields of one lattice.  The driver loop of gfc_job_fix restates
// gaugefixOverrelaxation (src/gauge_fixing.c:713-826 of the reference):
// G <- 1, residue, check schedule, half-sweeps with alternating parity,
// stall / convergence / max-sweeps tests, reunitarization every
// sweeps_to_reunitarization sweeps and once at the end.  Between two host
// decisions the half-sweeps run timeslice group by timeslice group on the
// four streams (sweep_blocked), bit-identical to whole-lattice sweeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

using gfc::cplx;
using gfc::kBlock;
using gfcx::fail;
using gfcx::nblocks;

thread_local std::string gfcx::last_error;

using namespace gfcx;

// ---------------------------------------------------------------------------
// Scratch of gfc_config_observables_noalloc: per-block partial sums of the
// observable kernels (3 doubles per 256-site block, at most kNoallocBlocks
// blocks) followed by the six reduced sums.  Each GPU holds its own copy for
// the life of the process; its device address is looked up once per device
// when a context is created and kept in noalloc_scratch[device].  Every
// kernel writes every partial slot the following reduction reads, so nothing
// has to be cleared between calls.
namespace {
constexpr int kNoallocBlocks = 43680;  // V <= 11 182 080 sites (48^3 x 96 fits)
constexpr int kNoallocOut = 3 * kNoallocBlocks;
constexpr int kNoallocDoubles = kNoallocOut + 8;
static __device__ __align__(256) double scratch[kNoallocDoubles];

constexpr int kMaxDevices = 64;
std::mutex noalloc_mu;
double *noalloc_scratch[kMaxDevices] = {};

// Device address of `scratch` on the current device (which is `device`).
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(GFC_EDEVICE, "gfc_ctx_create: device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_scratch[device] != nullptr) return GFC_OK;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
        return fail(GFC_EDEVICE, std::string("gfc_ctx_create: scratch address: ") +
                                     hipGetErrorString(e));
    noalloc_scratch[device] = static_cast<double *>(ptr);
    return GFC_OK;
}
}  // namespace

extern "C" {

const char *gfc_last_error(void) { return gfcx::last_error.c_str(); }

const char *gfc_version(void) { return "gfc-mi355x 0.1 (gfx950)"; }

void gfc_options_default(gfc_options *o) {
    o->precision = GFC_PREC_FP64;
    o->mixed_switch = 1e-9;
    o->mixed_probe_every = 50;
    o->mixed_defer = 1;
    o->verbose = 0;
    o->time_sweeps = 0;
    o->retire_slices = 0;
    o->slice_group = 0;
    o->slice_lanes = 0;
}

int gfc_ctx_create(int device, gfc_ctx **out) {
    if (out == nullptr) return fail(GFC_EINVAL, "gfc_ctx_create: NULL output");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(GFC_EDEVICE, "gfc_ctx_create: no HIP device available");
    if (device < 0 || device >= n) return fail(GFC_EINVAL, "gfc_ctx_create: bad device index");
    HIP_TRY(hipSetDevice(device));
    gfc_ctx *c = new gfc_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->res_dev, 4 * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc(&c->res_host, 8 * sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&c->flag_dev, sizeof(int));
    c->lane[0] = c->stream;
    for (int k = 1; k < gfc_ctx::kLanes && e == hipSuccess; k++)
        e = hipStreamCreateWithFlags(&c->lane[k], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
    for (int k = 0; k < gfc_ctx::kLanes && e == hipSuccess; k++)
        e = hipEventCreateWithFlags(&c->join[k], hipEventDisableTiming);
    if (e != hipSuccess) {
        gfc_ctx_destroy(c);
        return fail(GFC_EDEVICE, std::string("gfc_ctx_create: ") + hipGetErrorString(e));
    }
    const int rc = noalloc_setup(device);
    if (rc != GFC_OK) {
        gfc_ctx_destroy(c);
        return rc;
    }
    *out = c;
    return GFC_OK;
}

void gfc_ctx_destroy(gfc_ctx *c) {
    if (c == nullptr) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    hipFree(c->part);
    hipFree(c->res_dev);
    hipHostFree(c->res_host);
    hipFree(c->flag_dev);
    for (int k = 1; k < gfc_ctx::kLanes; k++)
        if (c->lane[k]) {
            hipStreamSynchronize(c->lane[k]);
            hipStreamDestroy(c->lane[k]);
        }
    for (int k = 0; k < 2; k++) {
        if (c->pin[k]) hipHostFree(c->pin[k]);
        if (c->pin_ev[k]) hipEventDestroy(c->pin_ev[k]);
    }
    if (c->fork) hipEventDestroy(c->fork);
    for (int k = 0; k < gfc_ctx::kLanes; k++)
        if (c->join[k]) hipEventDestroy(c->join[k]);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int gfc_ctx_copy_probe(gfc_ctx *c, int64_t bytes, int reps, double *seconds) {
    if (c == nullptr || bytes < 16 || reps < 1 || seconds == nullptr)
        return fail(GFC_EINVAL, "gfc_ctx_copy_probe: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    const int64_t n = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GFC_OK;
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n * 16);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * 16, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        const unsigned grid = nblocks(n);
        hipLaunchKernelGGL(gfc::k_copy16, dim3(grid), dim3(kBlock), 0, c->stream, a, b, n);
        e = hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps && e == hipSuccess; r++) {
            hipLaunchKernelGGL(gfc::k_copy16, dim3(grid), dim3(kBlock), 0, c->stream, a, b, n);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        *seconds = ms * 1e-3 / reps;
    }
    if (e != hipSuccess) rc = fail(GFC_EDEVICE, std::string("gfc_ctx_copy_probe: ") + hipGetErrorString(e));
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return rc;
}

int gfc_ctx_read_probe(gfc_ctx *c, int64_t bytes, int reps, double *seconds) {
    // reps + 2 passes run in one launch
    if (c == nullptr || bytes < 16 || reps < 1 || reps > (1 << 20) || seconds == nullptr)
        return fail(GFC_EINVAL, "gfc_ctx_read_probe: bad argument");
    HIP_TRY(hipSetDevice(c->device));
    const int64_t n = bytes / 16;
    // one chunk per workgroup, at least one wave's loads wide
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, n / kBlock));
    const int64_t chunk = n / grid;
    double2 *a = nullptr, *sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = GFC_OK;
    hipError_t e = hipMalloc(&a, n * 16);
    if (e == hipSuccess) e = hipMalloc(&sink, (size_t)grid * kBlock * 16);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * 16, c->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    // best of three launches of `passes` passes, in ms
    auto timed = [&](int passes) -> float {
        float best = 0;
        for (int k = 0; k < 3 && e == hipSuccess; k++) {
            e = hipEventRecord(e0, c->stream);
            if (e == hipSuccess) {
                hipLaunchKernelGGL(gfc::k_read16, dim3(grid), dim3(kBlock), 0, c->stream, a,
                                   chunk, passes, sink);
                e = hipGetLastError();
            }
            if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (k == 0 || ms < best) best = ms;
        }
        return best;
    };
    if (e == hipSuccess) {
        // the first launch brings a cache-sized buffer into the Infinity Cache;
        // the difference of a (2 + reps)-pass and a 2-pass launch is reps
        // passes without the launch's ramp-up and drain
        timed(2);
        const float t1 = timed(2), t2 = timed(2 + reps);
        *seconds = (t2 > t1 ? (t2 - t1) / reps : t2 / (2 + reps)) * 1e-3;
    }
    if (e != hipSuccess) rc = fail(GFC_EDEVICE, std::string("gfc_ctx_read_probe: ") + hipGetErrorString(e));
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(a);
    hipFree(sink);
    return rc;
}

int64_t gfc_job_footprint(const gfc_geom *g, int precision) {
    if (gfc_geom_check(g)) return -1;
    const int64_t V = gfc_geom_volume(g);
    // G keeps two rows (6 complex numbers) per SU(3) matrix, in fields of
    // whole 32-site tiles per parity; the u0 fields are allocated at the same
    // size per direction (FP64 u0 uses five numbers of them; the allocation
    // also stages FP32 input)
    const int64_t fs = gfc::field_size(V / 2);
    int64_t b = V * 4 * 144       // cfg
                + 6 * fs * 16     // u0
                + 2 * fs * 16;    // g
    if (precision == GFC_PREC_MIXED) b += 6 * fs * 8 + 2 * fs * 8;
    return b;
}

void gfc_job_destroy(gfc_job *j) {
    if (j == nullptr) return;
    hipSetDevice(j->ctx->device);
    hipStreamSynchronize(j->ctx->stream);
    hipFree(j->cfg);
    hipFree(j->u0);
    hipFree(j->g);
    hipFree(j->u0f);
    hipFree(j->gf);
    hipFree(j->saved);
    delete j;
}

}  // extern "C"

namespace gfcx {

int ensure_partials(gfc_ctx *c, size_t blocks) {
    if (blocks <= c->part_cap) return GFC_OK;
    hipStreamSynchronize(c->stream);
    hipFree(c->part);
    c->part = nullptr;
    c->part_cap = 0;
    HIP_TRY(hipMalloc(&c->part, blocks * 3 * sizeof(double)));
    c->part_cap = blocks;
    return GFC_OK;
}

int set_identity(gfc_job *j) { return set_identity(j, j->ctx->stream); }

int set_identity(gfc_job *j, hipStream_t s) {
    const unsigned b = nblocks(2 * (int64_t)j->Vh);
    hipLaunchKernelGGL(gfc::k_g_identity<double>, dim3(b), dim3(kBlock), 0, s, j->g, j->Vh);
    LAUNCH_CHECK();
    if (j->gf) {
        hipLaunchKernelGGL(gfc::k_g_identity<float>, dim3(b), dim3(kBlock), 0, s, j->gf, j->Vh);
        LAUNCH_CHECK();
    }
    return GFC_OK;
}

}  // namespace

extern "C" int gfc_job_create(gfc_ctx *ctx, const gfc_geom *g, int precision, gfc_job **out) {
    if (ctx == nullptr || out == nullptr) return fail(GFC_EINVAL, "gfc_job_create: NULL argument");
    *out = nullptr;
    if (gfc_geom_check(g))
        return fail(GFC_EGEOMETRY, "gfc_job_create: ns and nt must be positive and even");
    if (precision != GFC_PREC_FP64 && precision != GFC_PREC_MIXED)
        return fail(GFC_EINVAL, "gfc_job_create: bad precision");
    const int64_t V = gfc_geom_volume(g);
    // the kernels address one parity's field of one direction with 32-bit
    // byte offsets
    if (gfc::field_size(V / 2) * (int64_t)sizeof(cplx<double>) > (int64_t)0xffffffffll)
        return fail(GFC_EGEOMETRY, "gfc_job_create: lattice too large (more than 89478464 sites)");
    HIP_TRY(hipSetDevice(ctx->device));
    gfc_job *j = new gfc_job();
    j->ctx = ctx;
    j->geom = *g;
    j->V = V;
    j->Vh = (int)(V / 2);
    j->precision = precision;
    hipError_t e = hipMalloc(&j->cfg, (size_t)V * 4 * 9 * sizeof(cplx<double>));
    const size_t fs = (size_t)gfc::field_size(j->Vh);  // per parity and direction
    if (e == hipSuccess) e = hipMalloc(&j->u0, 6 * fs * sizeof(cplx<double>));
    if (e == hipSuccess) e = hipMalloc(&j->g, 2 * fs * sizeof(cplx<double>));
    if (e == hipSuccess && precision == GFC_PREC_MIXED) {
        e = hipMalloc(&j->u0f, 6 * fs * sizeof(cplx<float>));
        if (e == hipSuccess) e = hipMalloc(&j->gf, 2 * fs * sizeof(cplx<float>));
    }
    if (e != hipSuccess) {
        gfc_job_destroy(j);
        return fail(GFC_ENOMEM, std::string("gfc_job_create: ") + hipGetErrorString(e));
    }
    int rc = ensure_partials(ctx, nblocks(V));
    if (rc == GFC_OK) rc = set_identity(j);
    if (rc != GFC_OK) {
        gfc_job_destroy(j);
        return rc;
    }
    *out = j;
    return GFC_OK;
}

extern "C" int gfc_job_upload_config(gfc_job *j, const void *U, int in_precision) {
    if (j == nullptr || U == nullptr) return fail(GFC_EINVAL, "gfc_job_upload_config: NULL");
    if (in_precision != 4 && in_precision != 8)
        return fail(GFC_EINVAL, "gfc_job_upload_config: precision must be 4 or 8");
    HIP_TRY(hipSetDevice(j->ctx->device));
    hipStream_t s = j->ctx->stream;
    const int64_t nreal = j->V * 4 * 18;
    if (in_precision == 8) {
        HIP_TRY(hipMemcpyAsync(j->cfg, U, nreal * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        // FP32 staging lands in the u0 planes: 18V complex doubles = 288V bytes,
        // exactly the 4V x 18 floats of an FP32 configuration
        float *stage = reinterpret_cast<float *>(j->u0);
        HIP_TRY(hipMemcpyAsync(stage, U, nreal * sizeof(float), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(gfc::k_f32_to_f64, dim3(nblocks(nreal)), dim3(kBlock), 0, s, stage,
                           reinterpret_cast<double *>(j->cfg), nreal);
        LAUNCH_CHECK();
    }
    int rc = set_identity(j);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    return GFC_OK;
}

int gfcx::generate_smooth_device(gfc_job *j, uint64_t seed, double eps) {
    return generate_smooth_device(j, j->ctx->stream, seed, eps);
}

int gfcx::generate_smooth_device(gfc_job *j, hipStream_t s, uint64_t seed, double eps) {
    // base state of the draws, as in host_gen.c: mix64(seed + golden gamma)
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    const uint64_t base = z ^ (z >> 31);
    // g(n) staging in the u0 planes (288V bytes >= 144V bytes needed)
    cplx<double> *tmp = j->u0;
    hipLaunchKernelGGL(gfc::k_gen_haar, dim3(nblocks(j->V)), dim3(kBlock), 0, s, base, j->V, tmp);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(gfc::k_gen_links, dim3(nblocks(j->V)), dim3(kBlock), 0, s, base, eps,
                       j->geom.ns, j->geom.nt, tmp, j->cfg);
    LAUNCH_CHECK();
    return set_identity(j, s);
}

extern "C" int gfc_job_generate_smooth(gfc_job *j, uint64_t seed, double eps) {
    if (j == nullptr) return fail(GFC_EINVAL, "gfc_job_generate_smooth: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    RC_TRY(generate_smooth_device(j, seed, eps));
    HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    return GFC_OK;
}

extern "C" int gfc_job_reunitarize(gfc_job *j) {
    if (j == nullptr) return fail(GFC_EINVAL, "gfc_job_reunitarize: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    hipStream_t s = j->ctx->stream;
    HIP_TRY(hipMemsetAsync(j->ctx->flag_dev, 0, sizeof(int), s));
    const int64_t n = j->V * 4;
    hipLaunchKernelGGL(gfc::k_reunit_aos, dim3(nblocks(n)), dim3(kBlock), 0, s, j->cfg, n,
                       j->ctx->flag_dev);
    LAUNCH_CHECK();
    int flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, j->ctx->flag_dev, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (flag) return fail(GFC_EPROJECT, "gfc_job_reunitarize: a link has a null row");
    return GFC_OK;
}

namespace gfcx {

// Reduces the per-block partials of the last observable kernel and returns the
// three sums on the host.
int finish_reduction(gfc_ctx *c, unsigned blocks, double out[3]) {
    hipLaunchKernelGGL(gfc::k_reduce_final, dim3(1), dim3(kBlock), 0, c->stream, c->part,
                       (int)blocks, c->res_dev);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(c->res_host, c->res_dev, 3 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out[0] = c->res_host[0];
    out[1] = c->res_host[1];
    out[2] = c->res_host[2];
    return GFC_OK;
}

// e2, theta, F of the factored state; fp32 selects the FP32 G (MIXED phase 1),
// evaluated against the FP64 u0.
int obs_factored(gfc_job *j, bool fp32, double out[3]) {
    gfc_ctx *c = j->ctx;
    const unsigned b = nblocks(2 * (int64_t)j->Vh);
    if (fp32)
        hipLaunchKernelGGL((gfc::k_obs_factored<double, float>), dim3(b), dim3(kBlock), 0,
                           c->stream, j->u0, j->gf, j->geom.ns, j->Vh, c->part);
    else
        hipLaunchKernelGGL((gfc::k_obs_factored<double, double>), dim3(b), dim3(kBlock), 0,
                           c->stream, j->u0, j->g, j->geom.ns, j->Vh, c->part);
    LAUNCH_CHECK();
    double s[3];
    int rc = finish_reduction(c, b, s);
    if (rc) return rc;
    const double V = (double)j->V;
    out[0] = s[0] / V;
    out[1] = s[1] / (3.0 * V);
    out[2] = s[2] / (9.0 * V);
    return GFC_OK;
}

int obs_aos(gfc_job *j, double out[3]) {
    gfc_ctx *c = j->ctx;
    const unsigned b = nblocks(j->V);
    hipLaunchKernelGGL(gfc::k_obs_aos, dim3(b), dim3(kBlock), 0, c->stream, j->cfg, j->geom.ns,
                       j->V, c->part);
    LAUNCH_CHECK();
    double s[3];
    int rc = finish_reduction(c, b, s);
    if (rc) return rc;
    const double V = (double)j->V;
    out[0] = s[0] / V;
    out[1] = s[1] / (3.0 * V);
    out[2] = s[2] / (9.0 * V);
    return GFC_OK;
}

int build_u0(gfc_job *j) { return build_u0(j, j->ctx->stream); }

int build_u0(gfc_job *j, hipStream_t s) {
    const unsigned b = nblocks(2 * (int64_t)j->Vh);
    hipLaunchKernelGGL(gfc::k_build_u0<double>, dim3(b), dim3(kBlock), 0, s, j->cfg, j->u0,
                       j->geom.ns, j->Vh);
    LAUNCH_CHECK();
    if (j->u0f) {
        hipLaunchKernelGGL(gfc::k_build_u0<float>, dim3(b), dim3(kBlock), 0, s, j->cfg, j->u0f,
                           j->geom.ns, j->Vh);
        LAUNCH_CHECK();
    }
    return GFC_OK;
}

// u0 from cfg and G <- 1 at the start of a fix; the FP32 copies only for a
// mixed-precision fix.
int prepare_fix(gfc_job *j, bool mixed) {
    hipStream_t s = j->ctx->stream;
    const unsigned b = nblocks(2 * (int64_t)j->Vh);
    hipLaunchKernelGGL(gfc::k_build_u0<double>, dim3(b), dim3(kBlock), 0, s, j->cfg, j->u0,
                       j->geom.ns, j->Vh);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(gfc::k_g_identity<double>, dim3(b), dim3(kBlock), 0, s, j->g, j->Vh);
    LAUNCH_CHECK();
    if (mixed) {
        hipLaunchKernelGGL(gfc::k_build_u0<float>, dim3(b), dim3(kBlock), 0, s, j->cfg, j->u0f,
                           j->geom.ns, j->Vh);
        LAUNCH_CHECK();
        hipLaunchKernelGGL(gfc::k_g_identity<float>, dim3(b), dim3(kBlock), 0, s, j->gf, j->Vh);
        LAUNCH_CHECK();
    }
    return GFC_OK;
}

int materialize(gfc_job *j, bool project) { return materialize(j, j->ctx->stream, project); }

int materialize(gfc_job *j, hipStream_t s, bool project) {
    // one thread per link
    if (project)
        hipLaunchKernelGGL(gfc::k_materialize<true>, dim3(nblocks(4 * j->V)), dim3(kBlock), 0, s,
                           j->cfg, j->g, j->geom.ns, j->geom.nt, j->Vh);
    else
        hipLaunchKernelGGL(gfc::k_materialize<false>, dim3(nblocks(4 * j->V)), dim3(kBlock), 0,
                           s, j->cfg, j->g, j->geom.ns, j->geom.nt, j->Vh);
    LAUNCH_CHECK();
    return GFC_OK;
}

// Host-side sweep state: which G array is live.
struct SweepState {
    bool fp32 = false;
};

// Bytes of one timeslice as the cache-budget policy counts them: u0 (3 links)
// and G at six complex numbers each (an FP64 u0 link reads five; the policy
// below was measured with this count).
int64_t slice_bytes(const gfc_job *j, bool fp32) {
    const int64_t S = (int64_t)j->geom.ns * j->geom.ns * j->geom.ns;
    return S * 4 * 6 * (fp32 ? 8 : 16);
}

// Part of the 256 MiB Infinity Cache that the data of concurrent sweep
// launches may fill and still be read from it again by the next half-sweep.
constexpr int64_t kCacheBytes = 192ll << 20;

// Whether sweeps over groups of `group` timeslices (0: the whole lattice)
// read u0 and G(n) with non-temporal loads: unless what the concurrent
// launches sweep fits in kCacheBytes.  Measured at 48^3 x 96 with groups of
// 1 (FP64) and 2 (FP32) timeslices on four streams: cached loads make a
// half-sweep 1% (FP64) and 9% (FP32) faster than the best streamed grouping;
// on sets larger than the cache they are 4-5% slower.
bool stream_loads(const gfc_job *j, bool fp32, int group, int lanes) {
    const int64_t ts = slice_bytes(j, fp32);
    const int nt = j->geom.nt;
    if (group <= 0) return (int64_t)nt * ts > kCacheBytes;
    const int ngroups = (nt + group - 1) / group;
    lanes = std::min(lanes, ngroups);
    return (int64_t)lanes * group * ts > kCacheBytes;
}

// Lattices of at least this many parity sites that fit in kCacheBytes whole
// are swept as three concurrent groups rather than one whole-lattice launch.
constexpr int64_t kSplitMinSites = 1 << 17;

// Streams over which the timeslice groups of precision fp32 run at once
// (gfc_options.slice_lanes, or by default, measured on full fixes):
// * three for a lattice of at least kSplitMinSites parity sites that fits the
//   cache whole (three groups of nt/3 timeslices: 20^3 x 40 FP64 and FP32 23%
//   and 18% faster than one whole-lattice launch, 24^3 x 48 FP32 20%);
// * two for FP64 single timeslices too large for a cache-resident wave
//   (64^3 x 128: 6% faster than four);
// * four otherwise.
int lane_count(const gfc_job *j, const gfc_options &o, bool fp32) {
    if (o.slice_lanes >= 1 && o.slice_lanes <= gfc_ctx::kLanes) return o.slice_lanes;
    if (o.slice_group != 0) return gfc_ctx::kLanes;
    const int64_t ts = slice_bytes(j, fp32);
    if ((int64_t)j->geom.nt * ts <= kCacheBytes) return j->Vh >= kSplitMinSites ? 3 : 1;
    if (!fp32 && gfc_ctx::kLanes * ts > kCacheBytes) return 2;
    return gfc_ctx::kLanes;
}

template <typename T, bool POUT, bool STREAM>
void launch_sweep_pin(const cplx<T> *u0, cplx<T> *g, hipStream_t s, int ns, int Vh, int p,
                      T omega, int hits, int h0, int count, bool pin) {
    const unsigned b = nblocks(count);
    if (pin)
        hipLaunchKernelGGL((gfc::k_sweep<T, POUT, STREAM, true>), dim3(b), dim3(kBlock), 0, s,
                           u0, g, ns, Vh, p, omega, hits, h0, count);
    else
        hipLaunchKernelGGL((gfc::k_sweep<T, POUT, STREAM, false>), dim3(b), dim3(kBlock), 0, s,
                           u0, g, ns, Vh, p, omega, hits, h0, count);
}

// A half-sweep launch: pout projects the swept parity's new G on output, pin
// its G on input (see sweep_site), stream selects non-temporal loads.
template <typename T>
int launch_sweep(const cplx<T> *u0, cplx<T> *g, hipStream_t s, int ns, int Vh, int p, T omega,
                 int hits, int h0, int count, bool pout, bool stream, bool pin = false) {
    if (pout && stream)
        launch_sweep_pin<T, true, true>(u0, g, s, ns, Vh, p, omega, hits, h0, count, pin);
    else if (pout)
        launch_sweep_pin<T, true, false>(u0, g, s, ns, Vh, p, omega, hits, h0, count, pin);
    else if (stream)
        launch_sweep_pin<T, false, true>(u0, g, s, ns, Vh, p, omega, hits, h0, count, pin);
    else
        launch_sweep_pin<T, false, false>(u0, g, s, ns, Vh, p, omega, hits, h0, count, pin);
    LAUNCH_CHECK();
    return GFC_OK;
}

int half_sweep(gfc_job *j, SweepState &st, int parity, double omega, int hits) {
    hipStream_t s = j->ctx->stream;
    const int ns = j->geom.ns, Vh = j->Vh;
    const bool stream = stream_loads(j, st.fp32, 0, 1);
    if (st.fp32)
        return launch_sweep(j->u0f, j->gf, s, ns, Vh, parity, (float)omega, hits, 0, Vh, false,
                            stream);
    return launch_sweep(j->u0, j->g, s, ns, Vh, parity, omega, hits, 0, Vh, false, stream);
}

// Projects G on parities [p0, p0 + np) of the live array.
int reunit_g(gfc_job *j, bool fp32, int p0, int np) {
    return reunit_g(j, j->ctx->stream, fp32, p0, np);
}

int reunit_g(gfc_job *j, hipStream_t s, bool fp32, int p0, int np) {
    const unsigned b = nblocks((int64_t)np * j->Vh);
    if (fp32)
        hipLaunchKernelGGL(gfc::k_g_reunit<float>, dim3(b), dim3(kBlock), 0, s, j->gf, j->Vh, p0,
                           np, 0, j->Vh);
    else
        hipLaunchKernelGGL(gfc::k_g_reunit<double>, dim3(b), dim3(kBlock), 0, s, j->g, j->Vh, p0,
                           np, 0, j->Vh);
    LAUNCH_CHECK();
    return GFC_OK;
}

// Timeslices per group, or 0 for whole-lattice sweeps, for the sweeps in
// precision fp32.  The groups are spread over the context's streams, a wave
// of one group per stream at a time, so that one group's kernel tail overlaps
// the others' work.  By default (slice_group = 0), with L = lane_count:
// * a lattice whose u0 and G fit in kCacheBytes: nt / L timeslices per group,
//   or the whole lattice in one launch when L = 1; from the cache;
// * otherwise, if a wave of L single timeslices fits, the largest groups whose
//   wave fits, evened out so that every wave fills all streams (64^3 x 128 in
//   FP64: single timeslices on two streams, from the cache);
// * otherwise single timeslices with non-temporal loads (a timeslice is then
//   large enough that a launch still fills the GPU; single timeslices on four
//   streams were 8-10% faster than groups of 8 timeslices at 64^3 x 128).
int group_size(const gfc_job *j, const gfc_options &o, bool fp32) {
    if (o.slice_group < 0) return 0;
    const int nt = j->geom.nt;
    if (o.slice_group > 0) return o.slice_group < nt ? o.slice_group : nt;
    const int64_t ts = slice_bytes(j, fp32);
    const int64_t L = lane_count(j, o, fp32);
    if ((int64_t)nt * ts <= kCacheBytes)
        return L > 1 && nt >= L ? (int)((nt + L - 1) / L) : 0;
    const int64_t nmax = kCacheBytes / (L * ts);
    if (nmax < 1) return 1;
    const int64_t waves = (nt + L * nmax - 1) / (L * nmax);
    return (int)((nt + L * waves - 1) / (L * waves));
}

StreamSet ctx_streams(gfc_ctx *c) {
    StreamSet ss;
    for (int l = 0; l < gfc_ctx::kLanes; l++) {
        ss.lane[l] = c->lane[l];
        ss.join[l] = c->join[l];
    }
    ss.fork = c->fork;
    return ss;
}

// Half-sweeps [from, to) (parity = sweep number & 1), timeslice group by
// timeslice group, with the reunitarization of G after every sweep number
// s + 1 < to that is a multiple of R folded into the two sweeps around it:
// sweep s projects its own parity's G on output, sweep s + 1 the other
// parity's on input.  Every site update reads only sites of its own
// timeslice, so each group runs through the sweeps on its own and the result
// equals that of whole-lattice sweeps bit for bit.
int sweep_blocked(gfc_job *j, const StreamSet &ss, bool fp32, int from, int to, int group,
                  int max_lanes, double omega, int hits, unsigned R,
                  const std::vector<char> *active) {
    const int ns = j->geom.ns, nt = j->geom.nt, Vh = j->Vh;
    const int S = ns * ns * ns / 2;  // parity sites per timeslice
    const int ngroups = (nt + group - 1) / group;
    // the groups to sweep: all, or (one timeslice per group) the active slices
    std::vector<int> groups;
    for (int gi = 0; gi < ngroups; gi++)
        if (active == nullptr || (*active)[gi]) groups.push_back(gi);
    const int ng = (int)groups.size();
    if (ng == 0) return GFC_OK;
    const int lanes = std::min(std::min(max_lanes, gfc_ctx::kLanes), ng);
    const bool stream = stream_loads(j, fp32, group, max_lanes);
    if (lanes > 1) {
        HIP_TRY(hipEventRecord(ss.fork, ss.lane[0]));
        for (int l = 1; l < lanes; l++) HIP_TRY(hipStreamWaitEvent(ss.lane[l], ss.fork, 0));
    }
    // `lanes` groups at a time, one per stream, launched in step so that the
    // streams stay busy together
    auto launch_all = [&]() -> int {
        for (int w0 = 0; w0 < ng; w0 += lanes) {
            const int w1 = std::min(ng, w0 + lanes);
            for (int k = from; k < to; k++) {
                // a reunitarization after sweep number k + 1 < to: sweep k projects
                // its own parity on output, sweep k + 1 the other parity on input
                const bool pout = k + 1 < to && (unsigned)(k + 1) % R == 0;
                const bool pin = k > from && (unsigned)k % R == 0;
                for (int wi = w0; wi < w1; wi++) {
                    hipStream_t s = ss.lane[wi - w0];
                    const int t0 = groups[wi] * group;
                    const int h0 = t0 * S;
                    const int count = (t0 + group <= nt ? group : nt - t0) * S;
                    if (fp32)
                        RC_TRY(launch_sweep(j->u0f, j->gf, s, ns, Vh, k & 1, (float)omega, hits,
                                            h0, count, pout, stream, pin));
                    else
                        RC_TRY(launch_sweep(j->u0, j->g, s, ns, Vh, k & 1, omega, hits, h0,
                                            count, pout, stream, pin));
                }
            }
        }
        return GFC_OK;
    };
    const int rc = launch_all();
    // the main stream waits for the other streams also after a failed launch,
    // so that no work of this call is left behind it
    if (lanes > 1) {
        for (int l = 1; l < lanes; l++) {
            HIP_TRY(hipEventRecord(ss.join[l], ss.lane[l]));
            HIP_TRY(hipStreamWaitEvent(ss.lane[0], ss.join[l], 0));
        }
    }
    return rc;
}

int promote_g_only(gfc_job *j, hipStream_t s) {
    const unsigned b = nblocks(2 * (int64_t)j->Vh);
    hipLaunchKernelGGL(gfc::k_g_promote, dim3(b), dim3(kBlock), 0, s, j->gf, j->g, j->Vh);
    LAUNCH_CHECK();
    return GFC_OK;
}

int promote_g(gfc_job *j, SweepState &st) {
    RC_TRY(promote_g_only(j, j->ctx->stream));
    st.fp32 = false;
    return GFC_OK;
}

int promote_g_only(gfc_job *j) { return promote_g_only(j, j->ctx->stream); }

struct SweepTimer {
    bool on = false;
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    hipStream_t s = nullptr;
    ~SweepTimer() {
        for (auto e : ev) hipEventDestroy(e);
    }
    void before() {
        if (!on) return;
        if (used + 2 > ev.size()) {
            for (int k = 0; k < 64; k++) {
                hipEvent_t e;
                hipEventCreate(&e);
                ev.push_back(e);
            }
        }
        hipEventRecord(ev[used], s);
    }
    void after() {
        if (!on) return;
        hipEventRecord(ev[used + 1], s);
        used += 2;
    }
    double seconds() {
        if (!on) return 0;
        hipStreamSynchronize(s);
        double tot = 0;
        for (size_t k = 0; k + 1 < used; k += 2) {
            float ms = 0;
            hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
            tot += ms;
        }
        return tot * 1e-3;
    }
};

// Per-timeslice convergence tracking (gfc_options.retire_slices).  w, e2 and
// theta couple only sites of one timeslice, so a slice whose own proxy is at or
// below the tolerance can stop being swept: its proxy stays frozen while the
// others keep converging, and the global proxy (the mean over slices) is
// still what the check schedule tests.  Active slices are swept with the
// batched sweep kernel over (h0, parity) items of those slices only.
struct SliceTracker {
    bool on = false;
    int nt = 0, S = 0;  // S = ns^3 / 2 parity sites per timeslice
    std::vector<char> active;
    int n_active = 0;
    double slice_sweeps = 0;
    gfc::BatchJob *d_job = nullptr;
    int4 *d_items[2] = {nullptr, nullptr};
    int n_items[2] = {0, 0};
    bool dirty = true;
    bool desc_fp32 = false;
    std::vector<double> part;
    ~SliceTracker() {
        hipFree(d_job);
        hipFree(d_items[0]);
        hipFree(d_items[1]);
    }
    int init(gfc_job *j) {
        on = true;
        nt = j->geom.nt;
        S = j->geom.ns * j->geom.ns * j->geom.ns / 2;
        active.assign(nt, 1);
        n_active = nt;
        HIP_TRY(hipMalloc(&d_job, sizeof(gfc::BatchJob)));
        for (int p = 0; p < 2; p++)
            HIP_TRY(hipMalloc(&d_items[p], sizeof(int4) * (size_t)(j->Vh / kBlock + 1)));
        return GFC_OK;
    }
    // Sweeps the active slices of parity p (nothing when all have retired).
    int sweep(gfc_job *j, bool fp32, int p, double omega, int hits) {
        hipStream_t s = j->ctx->stream;
        slice_sweeps += (double)n_active / nt;
        if (n_active == 0) return GFC_OK;
        if (dirty || desc_fp32 != fp32) {
            gfc::BatchJob b{};
            b.u0 = fp32 ? (const void *)j->u0f : (const void *)j->u0;
            b.g = fp32 ? (void *)j->gf : (void *)j->g;
            b.omega = omega;
            b.ns = j->geom.ns;
            b.Vh = j->Vh;
            b.hits = hits;
            std::vector<int4> it[2];
            for (int t = 0; t < nt; t++)
                if (active[t])
                    for (int q = 0; q < 2; q++)
                        for (int h0 = t * S; h0 < (t + 1) * S; h0 += kBlock)
                            it[q].push_back(make_int4(0, h0, q, 0));
            HIP_TRY(hipMemcpyAsync(d_job, &b, sizeof b, hipMemcpyHostToDevice, s));
            for (int q = 0; q < 2; q++) {
                n_items[q] = (int)it[q].size();
                HIP_TRY(hipMemcpyAsync(d_items[q], it[q].data(), sizeof(int4) * it[q].size(),
                                       hipMemcpyHostToDevice, s));
            }
            HIP_TRY(hipStreamSynchronize(s));
            dirty = false;
            desc_fp32 = fp32;
        }
        const int n = n_items[p];
        // cached loads while the active slices' data fit in the cache
        const bool stream = (int64_t)n_active * slice_bytes(j, fp32) > kCacheBytes;
        if (fp32 && stream)
            hipLaunchKernelGGL((gfc::k_sweep_batch<float, true>), dim3(n), dim3(kBlock), 0, s,
                               d_job, d_items[p], n);
        else if (fp32)
            hipLaunchKernelGGL((gfc::k_sweep_batch<float, false>), dim3(n), dim3(kBlock), 0, s,
                               d_job, d_items[p], n);
        else if (stream)
            hipLaunchKernelGGL((gfc::k_sweep_batch<double, true>), dim3(n), dim3(kBlock), 0, s,
                               d_job, d_items[p], n);
        else
            hipLaunchKernelGGL((gfc::k_sweep_batch<double, false>), dim3(n), dim3(kBlock), 0, s,
                               d_job, d_items[p], n);
        LAUNCH_CHECK();
        return GFC_OK;
    }
    // Proxy of every timeslice of the current state; retires those at or below
    // tol.  Returns the global proxy (mean over slices) in *global.
    int probe(gfc_job *j, bool fp32, int which, double tol, double *global) {
        gfc_ctx *c = j->ctx;
        const unsigned b = nblocks(2 * (int64_t)j->Vh);
        if (fp32)
            hipLaunchKernelGGL((gfc::k_obs_factored<double, float>), dim3(b), dim3(kBlock), 0,
                               c->stream, j->u0, j->gf, j->geom.ns, j->Vh, c->part);
        else
            hipLaunchKernelGGL((gfc::k_obs_factored<double, double>), dim3(b), dim3(kBlock), 0,
                               c->stream, j->u0, j->g, j->geom.ns, j->Vh, c->part);
        LAUNCH_CHECK();
        part.resize((size_t)b * 3);
        HIP_TRY(hipMemcpyAsync(part.data(), c->part, sizeof(double) * part.size(),
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        const int bps = S / kBlock, half = j->Vh / kBlock;
        const double vt = 2.0 * S;  // sites per slice
        double sum = 0;
        for (int t = 0; t < nt; t++) {
            double acc = 0;
            for (int q = 0; q < 2; q++)
                for (int k = 0; k < bps; k++) acc += part[3 * (q * half + t * bps + k) + which];
            const double pt = acc / (which == 1 ? 3.0 * vt : vt);
            sum += pt;
            if (active[t] && pt <= tol) {
                active[t] = 0;
                n_active--;
                dirty = true;
            }
        }
        *global = sum / nt;
        return GFC_OK;
    }
};

int run_fix(gfc_job *j, const gfc_params *p, const gfc_options *o_in, gfc_report *rep,
            double *trace, int trace_len) {
    if (j == nullptr || p == nullptr) return fail(GFC_EINVAL, "gfc_job_fix: NULL argument");
    if (!gfc_params_valid(p)) return fail(GFC_EPARAMS, "gfc_job_fix: invalid parameters");
    gfc_options o;
    if (o_in)
        o = *o_in;
    else
        gfc_options_default(&o);
    if (o.precision == GFC_PREC_MIXED && j->precision != GFC_PREC_MIXED)
        return fail(GFC_EINVAL, "gfc_job_fix: MIXED needs a job created with GFC_PREC_MIXED");
    if (o.mixed_probe_every <= 0) o.mixed_probe_every = 50;
    HIP_TRY(hipSetDevice(j->ctx->device));
    const auto t0 = std::chrono::steady_clock::now();
    gfc_report r;
    std::memset(&r, 0, sizeof r);

    RC_TRY(prepare_fix(j, o.precision == GFC_PREC_MIXED));
    const double omega = (double)p->omega_OR;  // float in the reference
    const int hits = (int)p->hits;
    const int which = p->gfix_proxy == GFC_PROXY_THETA ? 1 : 0;
    SweepState st;
    st.fp32 = o.precision == GFC_PREC_MIXED;
    double ob[3];

    SweepTimer timer;
    timer.on = o.time_sweeps != 0;
    timer.s = j->ctx->stream;
    SliceTracker sl;
    if (o.retire_slices) {
        if (j->geom.ns % 8)
            return fail(GFC_EINVAL, "gfc_job_fix: retire_slices needs ns a multiple of 8");
        RC_TRY(sl.init(j));
    }

    RC_TRY(obs_factored(j, false, ob));
    r.residue_evals++;
    double res = ob[which];
    r.residue_initial = res;
    if (o.verbose) printf("Sweeps: %8d.\t residue: %3.2E \n", 0, res);
    int due = gfc_next_check(0, res, p);
    Handoff ho;
    ho.start(res);
    int sweep = 0;
    int rc = 0;
    bool done = false;
    // per-sweep traces need the state after every sweep of the whole lattice;
    // the other sweeps run timeslice-blocked, with slice retirement one
    // timeslice per group so that retired slices are left out
    const bool may_block = trace == nullptr;
    while (res > p->tolerance) {
        int group = may_block ? group_size(j, o, st.fp32) : 0;
        if (sl.on && group > 0) group = 1;
        const bool in32 = st.fp32;
        timer.before();
        if (group > 0) {
            // the next sweep after which the host looks at the state
            int64_t stop = due;
            const int64_t pe = o.mixed_probe_every;
            if ((st.fp32 && !ho.at_floor) || sl.on)
                stop = std::min<int64_t>(stop, (sweep / pe + 1) * pe);
            if (st.fp32 && ho.switch_at > sweep) stop = std::min<int64_t>(stop, ho.switch_at);
            stop = std::min<int64_t>(stop, (int64_t)p->max_sweeps_to_fix);
            if (stop <= sweep) stop = sweep + 1;
            RC_TRY(sweep_blocked(j, ctx_streams(j->ctx), st.fp32, sweep, (int)stop, group,
                                 lane_count(j, o, st.fp32), omega, hits,
                                 p->sweeps_to_reunitarization, sl.on ? &sl.active : nullptr));
            if (sl.on) sl.slice_sweeps += (double)(stop - sweep) * sl.n_active / sl.nt;
            if (in32) r.sweeps_fp32 += (int)stop - sweep;
            sweep = (int)stop;
        } else {
            if (sl.on)
                RC_TRY(sl.sweep(j, st.fp32, sweep & 1, omega, hits));
            else
                RC_TRY(half_sweep(j, st, sweep & 1, omega, hits));
            sweep++;
            if (in32) r.sweeps_fp32++;
        }
        timer.after();
        if (trace && sweep <= trace_len) {
            RC_TRY(obs_factored(j, st.fp32, ob));
            trace[sweep - 1] = ob[which];
        }
        if (((st.fp32 && !ho.at_floor) || sl.on) && sweep != due &&
            sweep % o.mixed_probe_every == 0) {
            double g = 0;
            if (sl.on) {
                RC_TRY(sl.probe(j, st.fp32, which, p->tolerance, &g));
            } else {
                RC_TRY(obs_factored(j, true, ob));
                g = ob[which];
            }
            r.residue_evals++;
            if (st.fp32 && ho.observe(o, p->tolerance, sweep, g, due)) RC_TRY(promote_g(j, st));
        }
        if (sweep == due) {
            const double last = res;
            if (sl.on) {
                // the per-slice measurement also gives the global proxy
                RC_TRY(sl.probe(j, st.fp32, which, p->tolerance, &res));
            } else {
                RC_TRY(obs_factored(j, st.fp32, ob));
                res = ob[which];
            }
            r.residue_evals++;
            if (std::fabs(res - last) / last < 10E-16) {
                if (o.verbose) printf("Sweeps: %8d.\t residue: %3.2E \n", sweep, res);
                rc = GFC_ENOTCONVERGED;
                done = true;
                break;
            } else if (res <= p->tolerance) {
                break;
            } else {
                if (o.verbose) printf("Sweeps: %8d.\t residue: %3.2E \n", sweep, res);
                due = gfc_next_check((unsigned)sweep, res, p);
                if (st.fp32 && ho.observe(o, p->tolerance, sweep, res, due))
                    RC_TRY(promote_g(j, st));
            }
        }
        if (st.fp32 && sweep >= ho.switch_at) RC_TRY(promote_g(j, st));
        if ((unsigned)sweep >= p->max_sweeps_to_fix) {
            rc = GFC_ENOTCONVERGED;
            done = true;
            break;
        }
        if (!((unsigned)sweep % p->sweeps_to_reunitarization))
            RC_TRY(reunit_g(j, st.fp32, 0, 2));
    }
    // left the loop in the FP32 phase (tolerance reached there or max sweeps):
    // hand the state over so that G and U are FP64
    if (st.fp32) RC_TRY(promote_g(j, st));
    // converged: the final reunitarization of G and U (src/gauge_fixing.c:822-823);
    // the max-sweeps and stall exits return before it
    if (!done) RC_TRY(reunit_g(j, false, 0, 2));
    RC_TRY(materialize(j, !done));
    HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    r.seconds_sweeps = timer.seconds();
    r.sweeps = sweep;
    r.slice_sweeps = sl.on ? sl.slice_sweeps : (double)sweep;
    r.residue_final = res;
    r.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rep) *rep = r;
    if (rc == GFC_ENOTCONVERGED) {
        gfcx::last_error = "gfc_job_fix: not converged (max sweeps or stalled residue)";
        return rc;
    }
    return sweep;
}

}  // namespace

extern "C" int gfc_job_fix(gfc_job *j, const gfc_params *p, const gfc_options *o,
                           gfc_report *rep) {
    return run_fix(j, p, o, rep, nullptr, 0);
}

extern "C" int gfc_job_fix_trace(gfc_job *j, const gfc_params *p, const gfc_options *o,
                                 gfc_report *rep, double *trace, int trace_len) {
    return run_fix(j, p, o, rep, trace, trace_len);
}

extern "C" int gfc_job_observables(gfc_job *j, double *e2, double *theta, double *F) {
    if (j == nullptr) return fail(GFC_EINVAL, "gfc_job_observables: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    double ob[3];
    RC_TRY(obs_aos(j, ob));
    if (e2) *e2 = ob[0];
    if (theta) *theta = ob[1];
    if (F) *F = ob[2];
    return GFC_OK;
}

extern "C" int gfc_job_plaquette(gfc_job *j, double *spatial, double *temporal) {
    if (j == nullptr) return fail(GFC_EINVAL, "gfc_job_plaquette: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    gfc_ctx *c = j->ctx;
    const unsigned b = nblocks(j->V);
    hipLaunchKernelGGL(gfc::k_plaquette_aos, dim3(b), dim3(kBlock), 0, c->stream, j->cfg,
                       j->geom.ns, j->geom.nt, j->V, c->part);
    LAUNCH_CHECK();
    double s[3];
    RC_TRY(finish_reduction(c, b, s));
    if (spatial) *spatial = s[0] / (3.0 * (double)j->V);
    if (temporal) *temporal = s[1] / (3.0 * (double)j->V);
    return GFC_OK;
}

extern "C" int gfc_job_download_config(gfc_job *j, double *U) {
    if (j == nullptr || U == nullptr) return fail(GFC_EINVAL, "gfc_job_download_config: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    HIP_TRY(hipMemcpyAsync(U, j->cfg, (size_t)j->V * 4 * 18 * sizeof(double),
                           hipMemcpyDeviceToHost, j->ctx->stream));
    HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    return GFC_OK;
}

extern "C" int gfc_job_save_config(gfc_job *j) {
    if (j == nullptr) return fail(GFC_EINVAL, "gfc_job_save_config: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    const size_t bytes = (size_t)j->V * 4 * 9 * sizeof(cplx<double>);
    if (j->saved == nullptr) {
        hipError_t e = hipMalloc(&j->saved, bytes);
        if (e != hipSuccess) {
            j->saved = nullptr;
            return fail(GFC_ENOMEM, std::string("gfc_job_save_config: ") + hipGetErrorString(e));
        }
    }
    HIP_TRY(hipMemcpyAsync(j->saved, j->cfg, bytes, hipMemcpyDeviceToDevice, j->ctx->stream));
    HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    return GFC_OK;
}

extern "C" int gfc_job_restore_config(gfc_job *j) {
    if (j == nullptr || j->saved == nullptr)
        return fail(GFC_EINVAL, "gfc_job_restore_config: nothing saved");
    HIP_TRY(hipSetDevice(j->ctx->device));
    const size_t bytes = (size_t)j->V * 4 * 9 * sizeof(cplx<double>);
    HIP_TRY(hipMemcpyAsync(j->cfg, j->saved, bytes, hipMemcpyDeviceToDevice, j->ctx->stream));
    int rc = set_identity(j);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    return GFC_OK;
}

extern "C" int gfc_job_get_link(gfc_job *j, int64_t site, int mu, double *out18) {
    if (j == nullptr || out18 == nullptr || site < 0 || site >= j->V || mu < 0 || mu > 3)
        return fail(GFC_EINVAL, "gfc_job_get_link: bad argument");
    HIP_TRY(hipSetDevice(j->ctx->device));
    HIP_TRY(hipMemcpyAsync(out18, j->cfg + (site * 4 + mu) * 9, 18 * sizeof(double),
                           hipMemcpyDeviceToHost, j->ctx->stream));
    HIP_TRY(hipStreamSynchronize(j->ctx->stream));
    return GFC_OK;
}

extern "C" int gfc_job_download_gauge(gfc_job *j, double *G) {
    if (j == nullptr || G == nullptr) return fail(GFC_EINVAL, "gfc_job_download_gauge: NULL");
    HIP_TRY(hipSetDevice(j->ctx->device));
    hipStream_t s = j->ctx->stream;
    // AoS staging in the u0 planes, which the next fix rebuilds anyway
    cplx<double> *stage = j->u0;
    hipLaunchKernelGGL(gfc::k_g_to_aos, dim3(nblocks(j->V)), dim3(kBlock), 0, s, j->g, stage,
                       j->geom.ns, j->Vh);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(G, stage, (size_t)j->V * 18 * sizeof(double), hipMemcpyDeviceToHost,
                           s));
    HIP_TRY(hipStreamSynchronize(s));
    return GFC_OK;
}

extern "C" int gfc_fix_coulomb_or(gfc_ctx *ctx, const gfc_geom *g, const gfc_params *p,
                                  const gfc_options *o, double *U, double *G, gfc_report *rep) {
    if (ctx == nullptr || U == nullptr || G == nullptr || p == nullptr)
        return fail(GFC_EINVAL, "gfc_fix_coulomb_or: NULL argument");
    if (!gfc_params_valid(p)) return fail(GFC_EPARAMS, "gfc_fix_coulomb_or: invalid parameters");
    const int prec = (o && o->precision == GFC_PREC_MIXED) ? GFC_PREC_MIXED : GFC_PREC_FP64;
    gfc_job *j = nullptr;
    RC_TRY(gfc_job_create(ctx, g, prec, &j));
    int rc = gfc_job_upload_config(j, U, 8);
    int sweeps = rc;
    if (rc == GFC_OK) {
        sweeps = gfc_job_fix(j, p, o, rep);
        if (sweeps >= 0 || sweeps == GFC_ENOTCONVERGED) {
            rc = gfc_job_download_config(j, U);
            if (rc == GFC_OK) rc = gfc_job_download_gauge(j, G);
            if (rc != GFC_OK) sweeps = rc;
        }
    }
    gfc_job_destroy(j);
    return sweeps;
}

extern "C" int gfc_reunitarize_field(gfc_ctx *ctx, double *field, int64_t n) {
    if (ctx == nullptr || field == nullptr || n < 0)
        return fail(GFC_EINVAL, "gfc_reunitarize_field: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    cplx<double> *d = nullptr;
    HIP_TRY(hipMalloc(&d, (size_t)n * 9 * sizeof(cplx<double>) + 16));
    int rc = GFC_OK;
    int flag = 0;
    do {
        hipError_t e = hipMemcpyAsync(d, field, (size_t)n * 18 * sizeof(double),
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemsetAsync(ctx->flag_dev, 0, sizeof(int), s);
        if (e != hipSuccess) {
            rc = fail(GFC_EDEVICE, hipGetErrorString(e));
            break;
        }
        if (n > 0)
            hipLaunchKernelGGL(gfc::k_reunit_aos, dim3(nblocks(n)), dim3(kBlock), 0, s, d, n,
                               ctx->flag_dev);
        e = hipGetLastError();
        if (e == hipSuccess)
            e = hipMemcpyAsync(field, d, (size_t)n * 18 * sizeof(double), hipMemcpyDeviceToHost,
                               s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&flag, ctx->flag_dev, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(GFC_EDEVICE, hipGetErrorString(e));
    } while (0);
    hipFree(d);
    if (rc == GFC_OK && flag) rc = fail(GFC_EPROJECT, "gfc_reunitarize_field: null row");
    return rc;
}

extern "C" int gfc_config_observables_noalloc(gfc_ctx *ctx, const gfc_geom *g,
                                              const double *U_dev, double *e2, double *theta,
                                              double *F, double *plaq_spatial,
                                              double *plaq_temporal) {
    if (ctx == nullptr || g == nullptr || U_dev == nullptr)
        return fail(GFC_EINVAL, "gfc_config_observables_noalloc: NULL argument");
    if (gfc_geom_check(g))
        return fail(GFC_EGEOMETRY, "gfc_config_observables_noalloc: bad geometry");
    const int64_t V = gfc_geom_volume(g);
    const unsigned b = nblocks(V);
    if (b > (unsigned)kNoallocBlocks)
        return fail(GFC_EINVAL, "gfc_config_observables_noalloc: lattice larger than the "
                                "scratch allows (at most 11182080 sites)");
    HIP_TRY(hipSetDevice(ctx->device));
    double *ptr = noalloc_scratch[ctx->device];
    if (ptr == nullptr)
        return fail(GFC_EDEVICE, "gfc_config_observables_noalloc: scratch not set up");
    hipStream_t s = ctx->stream;
    const cplx<double> *cfg = reinterpret_cast<const cplx<double> *>(U_dev);
    // the same kernels and block decomposition as gfc_job_observables and
    // gfc_job_plaquette, with the partials and sums in the scratch
    hipLaunchKernelGGL(gfc::k_obs_aos, dim3(b), dim3(kBlock), 0, s, cfg, g->ns, V, ptr);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(gfc::k_reduce_final, dim3(1), dim3(kBlock), 0, s, ptr, (int)b,
                       ptr + kNoallocOut);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(gfc::k_plaquette_aos, dim3(b), dim3(kBlock), 0, s, cfg, g->ns, g->nt, V,
                       ptr);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(gfc::k_reduce_final, dim3(1), dim3(kBlock), 0, s, ptr, (int)b,
                       ptr + kNoallocOut + 3);
    LAUNCH_CHECK();
    HIP_TRY(hipMemcpyAsync(ctx->res_host, ptr + kNoallocOut, 6 * sizeof(double),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const double *r = ctx->res_host, dV = (double)V;
    if (e2) *e2 = r[0] / dV;
    if (theta) *theta = r[1] / (3.0 * dV);
    if (F) *F = r[2] / (9.0 * dV);
    if (plaq_spatial) *plaq_spatial = r[3] / (3.0 * dV);
    if (plaq_temporal) *plaq_temporal = r[4] / (3.0 * dV);
    return GFC_OK;
}

extern "C" int gfc_config_observables(gfc_ctx *ctx, const gfc_geom *g, const double *U,
                                      double *e2, double *theta, double *F,
                                      double *plaq_spatial, double *plaq_temporal) {
    gfc_job *j = nullptr;
    RC_TRY(gfc_job_create(ctx, g, GFC_PREC_FP64, &j));
    int rc = gfc_job_upload_config(j, U, 8);
    if (rc == GFC_OK) rc = gfc_job_observables(j, e2, theta, F);
    if (rc == GFC_OK) rc = gfc_job_plaquette(j, plaq_spatial, plaq_temporal);
    gfc_job_destroy(j);
    return rc;
}
