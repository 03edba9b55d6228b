// internal.hpp -- host-side internals shared by the engine translation units
// (engine.hip: single-lattice jobs; server.hip: the serving scheduler).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "gfc.h"
#include "handoff.hpp"
#include "kernels.hpp"

namespace gfcx {

using gfc::cplx;
using gfc::kBlock;

// Message of the most recent failure of the calling thread (gfc_last_error).
extern thread_local std::string last_error;

inline int fail(int code, const std::string &msg) {
    last_error = msg;
    return code;
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace gfcx

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return gfcx::fail(GFC_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

#define RC_TRY(expr)         \
    do {                     \
        int rc_ = (expr);    \
        if (rc_) return rc_; \
    } while (0)

struct gfc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    double *part = nullptr;  // per-block partial sums
    size_t part_cap = 0;     // in blocks
    double *res_dev = nullptr;
    double *res_host = nullptr;  // pinned, 8 doubles
    int *flag_dev = nullptr;
    // extra streams (default priority) over which independent timeslice
    // groups are spread, and the events that fork them from and join them
    // back into `stream`
    static constexpr int kLanes = 4;
    hipStream_t lane[kLanes] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kLanes] = {};
    // pinned host blocks of the file streams (io_device.hip), made on first use
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t pin_ev[2] = {nullptr, nullptr};
};

struct gfc_job {
    gfc_ctx *ctx = nullptr;
    gfc_geom geom{0, 0};
    int64_t V = 0;
    int Vh = 0;
    int precision = GFC_PREC_FP64;
    gfc::cplx<double> *cfg = nullptr;    // AoS FP64 U, 4V matrices
    gfc::cplx<double> *u0 = nullptr;     // SoA spatial links, FP64 (five numbers each)
    gfc::cplx<double> *g = nullptr;      // SoA G, FP64 (two rows)
    gfc::cplx<float> *u0f = nullptr;     // SoA spatial links, FP32 (MIXED)
    gfc::cplx<float> *gf = nullptr;      // SoA G, FP32 (MIXED)
    gfc::cplx<double> *saved = nullptr;  // gfc_job_save_config copy of cfg
};

namespace gfcx {

// Helpers of engine.hip used by the scheduler.  They launch on the given
// stream, or on j->ctx->stream without one.
int set_identity(gfc_job *j);
int set_identity(gfc_job *j, hipStream_t s);
int build_u0(gfc_job *j);
int build_u0(gfc_job *j, hipStream_t s);
int materialize(gfc_job *j, bool project);
int materialize(gfc_job *j, hipStream_t s, bool project);
int reunit_g(gfc_job *j, bool fp32, int p0, int np);
int reunit_g(gfc_job *j, hipStream_t s, bool fp32, int p0, int np);
int promote_g_only(gfc_job *j);
int promote_g_only(gfc_job *j, hipStream_t s);
int generate_smooth_device(gfc_job *j, uint64_t seed, double eps);
int generate_smooth_device(gfc_job *j, hipStream_t s, uint64_t seed, double eps);
int ensure_partials(gfc_ctx *c, size_t blocks);

// Streams over which timeslice groups run at once; lane[0] is the stream the
// work is forked from and joined back into.
struct StreamSet {
    hipStream_t lane[gfc_ctx::kLanes] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[gfc_ctx::kLanes] = {};
};
StreamSet ctx_streams(gfc_ctx *c);
// Timeslices per group (0: whole-lattice sweeps) and streams for the sweeps
// of precision fp32 under options o (gfc_options.slice_group, slice_lanes).
int group_size(const gfc_job *j, const gfc_options &o, bool fp32);
int lane_count(const gfc_job *j, const gfc_options &o, bool fp32);
// Half-sweeps [from, to) timeslice group by timeslice group on the streams of
// ss, with the reunitarization of G after every sweep s + 1 < to that is a
// multiple of R folded in; bit-identical to whole-lattice sweeps.
int sweep_blocked(gfc_job *j, const StreamSet &ss, bool fp32, int from, int to, int group,
                  int max_lanes, double omega, int hits, unsigned R,
                  const std::vector<char> *active = nullptr);

}  // namespace gfcx
