/*
 * gfc.h -- C ABI of the MI355X Coulomb gauge-fixing engine (libgfc.so).
 *
 * Plain C types only: pointers, sizes, ints and doubles.  No exceptions or
 * device types cross this boundary; every call returns an int status
 * (0 on success, a negative GFC_E* code otherwise) unless stated, and
 * gfc_last_error() describes the most recent failure of the calling thread.
 *
 * Host-side field layout is the reference's memory/file image
 * (include/fields.h:33-43 of jesuel-marques/GaugeFix-Coulomb):
 *   site = ((t*ns + z)*ns + y)*ns + x,  link = site*4 + mu,
 *   mu: 0 = T, 1 = Z, 2 = Y, 3 = X  (settings.h:57-60),
 *   one matrix = 9 complex numbers, row-major, (re, im) interleaved:
 *   18 doubles (144 B) in FP64, 18 floats (72 B) in FP32.
 */
#ifndef GFC_H
#define GFC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GFC_OK 0
#define GFC_ENOTCONVERGED (-1) /* max sweeps reached or residue stalled
                                  (gauge_fixing.c:797,811 return -1)       */
#define GFC_EPARAMS (-2)       /* invalid parameters (gauge_fixing.c:749-751) */
#define GFC_EDEVICE (-3)       /* HIP runtime error or no usable GPU          */
#define GFC_EGEOMETRY (-4)     /* ns, nt not positive and even                */
#define GFC_ENOMEM (-5)        /* host or device allocation failed            */
#define GFC_EINVAL (-6)        /* bad argument (NULL pointer, bad enum, size) */
#define GFC_EPROJECT (-7)      /* SU(3) projection met a null row            */

/* ---- parameters -------------------------------------------------------- */
#define GFC_PROXY_E2 0
#define GFC_PROXY_THETA 1

/* Mirrors ORGaugeFixingParameters (include/gauge_fixing.h:31-54).  omega_OR
 * stays a float as in the reference, so 1.95 acts as 1.9500000476837158. */
typedef struct {
    double tolerance;
    int gfix_proxy;                          /* GFC_PROXY_E2 or GFC_PROXY_THETA */
    unsigned max_sweeps_to_fix;
    unsigned estimate_sweeps_to_gf_progress;
    unsigned sweeps_to_reunitarization;
    int error;                               /* set when the values are invalid */
    float omega_OR;
    unsigned hits;
} gfc_params;

/* Replaces initParametersORDefault (src/gauge_fixing.c:72-84). */
void gfc_params_default(gfc_params *p);
/* Replaces initORGaugeFixingParameters (src/gauge_fixing.c:144-244): reads
 * "key: value" lines (whitespace removed, split at the first ':'); unknown
 * keys are reported on stderr; an unreadable file keeps the defaults; the
 * `error` field is set when the result is invalid.  path " " means "no file".
 * Returns 0, or GFC_EPARAMS when `error` was set. */
int gfc_params_from_file(const char *path, gfc_params *p);
/* Replaces validORGaugeFixingParametersQ (src/gauge_fixing.c:45-70): 1 valid. */
int gfc_params_valid(const gfc_params *p);
/* Replaces printORGaugeFixingParameters (src/gauge_fixing.c:86-104). */
void gfc_params_print(const gfc_params *p);
/* Replaces whenNextConvergenceCheckQ (src/gauge_fixing.c:248-285). */
int gfc_next_check(unsigned sweep, double residue, const gfc_params *p);

/* ---- geometry ---------------------------------------------------------- */
typedef struct {
    int ns; /* spatial extent (x, y, z) */
    int nt; /* temporal extent */
} gfc_geom;

/* Replaces initGeometry/validGeometricParametersQ (src/geometry.c:148-228);
 * additionally requires even ns and nt (red/black independence across the
 * periodic wrap).  Returns 0 or GFC_EGEOMETRY. */
int gfc_geom_check(const gfc_geom *g);
/* Number of sites ns^3 * nt. */
int64_t gfc_geom_volume(const gfc_geom *g);

/* ---- host I/O (src/fields_io.c) ----------------------------------------- */
/* Replaces loadConfig (src/fields_io.c:216-305) and, for in_precision 4,
 * convertCfg_in_work (src/convert_precision_io.c:176-215).  Reads exactly
 * 4*V matrices of in_precision (4 or 8) bytes per real, byte-swapping each
 * real when byte_swap != 0, into U (4*V*18 doubles).  Returns 0, -1 (open),
 * -2 (short file), -3 (trailing bytes), -4 (bad precision) like the
 * reference. */
int gfc_load_config(const char *path, const gfc_geom *g, int in_precision, int byte_swap,
                    double *U);
/* Writes U (4*V*18 doubles) as precision 4 or 8 with optional byte swap
 * (writeConfig, src/fields_io.c:405-489). */
int gfc_write_config(const char *path, const gfc_geom *g, int out_precision, int byte_swap,
                     const double *U);
/* Replaces writeGaugeTransf (src/fields_io.c:493-565): raw native-endian
 * FP64 image of G (V*18 doubles).  Returns 0, -1 (open), -2 (short write). */
int gfc_write_gauge_transf(const char *path, const gfc_geom *g, const double *G);
/* Reads a G file written by gfc_write_gauge_transf (loadGaugeTransf,
 * src/fields_io.c:309-401). */
int gfc_load_gauge_transf(const char *path, const gfc_geom *g, double *G);
/* Replaces writeSweepsToGaugefix (gauge_fix_coulomb.c:41-65): appends
 * "<identifier>\t<sweeps>\n" to sweeps_to_gaugefix_<ns>x<nt>.txt in dir
 * (NULL = current directory). */
int gfc_write_sweeps_record(const char *dir, const gfc_geom *g, const char *identifier,
                            int sweeps);

/* ---- synthetic input (no counterpart in the reference) ------------------ */
/* Writes a "smooth" configuration U_mu(n) = g(n) exp(i eps H_mu(n)) g(n+mu)^dagger
 * into U (4*V*18 doubles): g Haar-random SU(3), H traceless Hermitian with unit
 * variance Gaussian entries, from a counter-based RNG keyed by seed, so the
 * result does not depend on the thread count.  Returns 0 or GFC_EGEOMETRY. */
int gfc_generate_smooth(const gfc_geom *g, uint64_t seed, double eps, double *U);
/* Rounds a FP64 field to FP32 (n matrices). */
void gfc_field_to_fp32(const double *in, float *out, int64_t n_matrices);

/* ---- device engine ------------------------------------------------------ */
#define GFC_PREC_FP64 0  /* every sweep in FP64 (drop-in)                      */
#define GFC_PREC_MIXED 1 /* FP32 sweeps until the FP32 phase is finished
                            (mixed_switch or its rounding floor) and, with
                            mixed_defer, until the FP64 sweeps left before
                            the next check suffice; then FP64 sweeps to the
                            tolerance                                          */

typedef struct {
    int precision;          /* GFC_PREC_FP64 (default) or GFC_PREC_MIXED      */
    double mixed_switch;    /* MIXED: FP32 -> FP64 hand-off residue (1e-9)     */
    int mixed_probe_every;  /* MIXED: half-sweeps between FP32-phase residue
                               probes (default 50); scheduled checks probe too */
    int mixed_defer;        /* MIXED: 1 (default): once the FP32 phase has
                               reached mixed_switch or its rounding floor, keep
                               sweeping in FP32 until the sweep that leaves,
                               before the next scheduled check, the FP64
                               sweeps needed to reach the tolerance at the
                               FP32 phase's convergence rate (x1.1 + 10);
                               0: hand over at once                           */
    int verbose;            /* 1: print "Sweeps: ..." lines like the reference */
    int time_sweeps;        /* 1: bracket every whole-lattice sweep launch, or
                               every timeslice-blocked run of sweeps between
                               two host decisions, with HIP events on the job's
                               stream and report their total in
                               gfc_report.seconds_sweeps                       */
    int retire_slices;      /* 1: per-timeslice convergence tracking.  Every
                               mixed_probe_every sweeps (both precisions) the
                               proxy of each timeslice is measured; a slice at
                               or below the tolerance stops being swept (its
                               proxy is frozen: w, e2 and theta only couple
                               sites of one timeslice).  The global proxy is
                               still checked on the reference schedule.  Needs
                               ns a multiple of 8; not the drop-in default.    */
    int slice_group;        /* timeslice blocking of the sweeps between two
                               host decisions (checks, probes, max sweeps):
                               the lattice is cut into groups of slice_group
                               timeslices, and each group runs all those
                               sweeps (with its reunitarizations) on one of
                               four streams, concurrently with the other
                               groups, so that kernel tails overlap.
                               Timeslices do not interact, so the result is
                               bit-identical to whole-lattice sweeps.
                               0 (default): groups sized so that a wave
                               (one group per stream) stays in 192 MiB of the
                               Infinity Cache between half-sweeps (read with
                               cached loads); a lattice that fits in it whole
                               is swept as three groups of nt/3 timeslices
                               when it has at least 2^17 sites per parity,
                               whole otherwise; single timeslices when a wave
                               of them does not fit; -1: whole-lattice
                               sweeps.  With retire_slices, one timeslice per
                               group.                                         */
    int slice_lanes;        /* streams over which the timeslice groups run
                               concurrently: 1-4, or 0 (default): three for a
                               lattice that fits the cache whole, two for
                               FP64 timeslices too large for a cache-resident
                               wave of four, four otherwise                   */
} gfc_options;

typedef struct {
    int sweeps;             /* half-sweeps performed                          */
    int sweeps_fp32;        /* of which in FP32 (MIXED)                        */
    int residue_evals;      /* residue evaluations (checks + probes)          */
    double residue_initial; /* proxy before the first sweep                    */
    double residue_final;   /* proxy at the last evaluation                    */
    double seconds;         /* wall time of the fixing loop on the device      */
    double seconds_sweeps;  /* of which in sweep kernels (HIP events)          */
    double slice_sweeps;    /* retire_slices: timeslice half-sweeps performed
                               divided by nt (equals sweeps without retirement) */
} gfc_report;

void gfc_options_default(gfc_options *o);

typedef struct gfc_ctx gfc_ctx; /* one HIP device + four streams */
typedef struct gfc_job gfc_job; /* one lattice resident in HBM */

const char *gfc_last_error(void);
/* Library version string. */
const char *gfc_version(void);

int gfc_ctx_create(int device, gfc_ctx **ctx);
/* Destroy a context's jobs and servers first: they use its streams. */
void gfc_ctx_destroy(gfc_ctx *ctx);
/* Device copy of `bytes` bytes (a plain 16-byte-per-lane streaming copy
 * kernel) repeated `reps` times on the context's stream; *seconds receives
 * the mean kernel time from HIP events.  Measures the achievable HBM copy
 * bandwidth and calibrates memory-traffic counters (no counterpart in the
 * reference). */
int gfc_ctx_copy_probe(gfc_ctx *ctx, int64_t bytes, int reps, double *seconds);
/* Streams a zeroed device buffer of `bytes` bytes (rounded down to a whole
 * number of 16-byte elements per workgroup chunk: exact for multiples of
 * 64 KiB) through a 16-byte-per-lane read kernel, many passes in one launch
 * with the chunks rotating over the workgroups so that no pass is served by
 * an L2.  *seconds receives the time of one pass without the launch's ramp-up
 * and drain: the difference of a (2 + reps)-pass and a 2-pass launch over
 * reps (1 to 2^20), from HIP events, best of three.  A buffer of at most 256 MiB is read
 * from the Infinity Cache, a larger one from HBM: the read rates that bound
 * the sweeps' L2-miss traffic (no counterpart in the reference). */
int gfc_ctx_read_probe(gfc_ctx *ctx, int64_t bytes, int reps, double *seconds);
/* Bytes of device memory a job of this geometry needs for the given
 * precision (used for admission by footprint). */
int64_t gfc_job_footprint(const gfc_geom *g, int precision);

/* Allocates the device fields of one lattice. */
int gfc_job_create(gfc_ctx *ctx, const gfc_geom *g, int precision, gfc_job **job);
void gfc_job_destroy(gfc_job *job);
/* Copies a host configuration (4*V matrices; precision 8 = double, 4 = float)
 * to the device; G is set to the identity. */
int gfc_job_upload_config(gfc_job *job, const void *U, int in_precision);
/* Synthesises the smooth configuration of gfc_generate_smooth directly in the
 * job's device memory (same construction; agrees with the host generator to
 * rounding, not bit for bit); G is set to the identity.  No counterpart in the
 * reference: it feeds the serving scheduler without host-side generation. */
int gfc_job_generate_smooth(gfc_job *job, uint64_t seed, double eps);
/* Projects all 4*V links onto SU(3) (main's initial reunitarizeField,
 * gauge_fix_coulomb.c:126-131).  Returns GFC_EPROJECT if a link had a null
 * row, like reunitarizeField's -1. */
int gfc_job_reunitarize(gfc_job *job);
/* e2 (calculate_e2, src/gauge_fixing.c:389-442), theta (calculateTheta,
 * :341-384) and F (calculateF, :289-337) of the job's current U. */
int gfc_job_observables(gfc_job *job, double *e2, double *theta, double *F);
/* averagePlaquette(U, "spatial"/"temporal") (src/measurement.c:108-155). */
int gfc_job_plaquette(gfc_job *job, double *spatial, double *temporal);
/* Replaces gaugefixOverrelaxation (src/gauge_fixing.c:713-826, declared at
 * include/gauge_fixing.h:67-69) on the resident job: G <- 1, red/black
 * overrelaxation sweeps with the reference's check schedule, stall test,
 * max-sweeps test and reunitarization every sweeps_to_reunitarization
 * sweeps; U and G are left gauge-fixed on the device.  Returns the sweep
 * count, GFC_ENOTCONVERGED or GFC_EPARAMS (or a device error). */
int gfc_job_fix(gfc_job *job, const gfc_params *p, const gfc_options *o, gfc_report *rep);
/* gfc_job_fix that also stores the proxy measured after each half-sweep k
 * (before that sweep's reunitarization) in trace[k-1], for k <= trace_len.
 * Costs one residue evaluation per sweep; meant for convergence studies. */
int gfc_job_fix_trace(gfc_job *job, const gfc_params *p, const gfc_options *o,
                      gfc_report *rep, double *trace, int trace_len);
/* Copies the job's current U (4*V*18 doubles) / G (V*18 doubles) to the host. */
int gfc_job_download_config(gfc_job *job, double *U);
int gfc_job_download_gauge(gfc_job *job, double *G);
/* Keeps a device-side copy of the current U (allocated on first use) /
 * restores it and resets G to the identity: re-fixing the same input without
 * a host transfer. */
int gfc_job_save_config(gfc_job *job);
int gfc_job_restore_config(gfc_job *job);
/* Copies one link U_mu(site) (18 doubles) of the job's current U to the host
 * (what gauge_fix_coulomb.c:120,132 prints with printMatrix3x3). */
int gfc_job_get_link(gfc_job *job, int64_t site, int mu, double *out18);

/* ---- file streams between disk and the device ---------------------------- */
/* loadConfig (src/fields_io.c:216-305; FP32 input as convertCfg_in_work,
 * src/convert_precision_io.c:176-215) straight into the job: the file is read
 * in pinned blocks by parallel pread while the previous block is copied to the
 * device, then byte-swapped (byte_swap != 0) and converted on the device.  G
 * is reset to the identity.  Same file rules and return codes as
 * gfc_load_config: 0, -1 (open), -2 (short file or read error), -3 (trailing
 * bytes), -4 (bad precision); regular files only. */
int gfc_job_load_config(gfc_job *job, const char *path, int in_precision, int byte_swap);
/* writeGaugeTransf (src/fields_io.c:493-565) of the job's G: native-endian
 * FP64, V*144 bytes.  0, -1 (create), -2 (write error). */
int gfc_job_write_gauge_transf(gfc_job *job, const char *path);
/* writeConfig (src/fields_io.c:405-489) of the job's current U, FP64 or FP32
 * (out_precision 8 or 4), optionally byte-swapped.  0, -1, -2, -4. */
int gfc_job_write_config(gfc_job *job, const char *path, int out_precision, int byte_swap);

/* The drop-in entry point: gaugefixOverrelaxation(U, G, params) on host
 * arrays (include/gauge_fixing.h:67-69).  U (4*V*18 doubles) and G (V*18
 * doubles) are updated in place with the same return codes.  o and rep may
 * be NULL.  U must hold SU(3) links, as the reference's main leaves them
 * (reunitarizeField before the call, gauge_fix_coulomb.c:126): the engine
 * keeps each link as its first two rows (an FP64 link as row 0 and two
 * elements of row 1) and rebuilds the rest from unitarity. */
int gfc_fix_coulomb_or(gfc_ctx *ctx, const gfc_geom *g, const gfc_params *p,
                       const gfc_options *o, double *U, double *G, gfc_report *rep);
/* reunitarizeField (src/fields.c:144-187) on a host array of n matrices. */
int gfc_reunitarize_field(gfc_ctx *ctx, double *field, int64_t n);
/* calculate_e2 / calculateTheta / calculateF / averagePlaquette of a host
 * configuration (4*V*18 doubles), computed on the device. */
int gfc_config_observables(gfc_ctx *ctx, const gfc_geom *g, const double *U, double *e2,
                           double *theta, double *F, double *plaq_spatial,
                           double *plaq_temporal);

/* No counterpart in the reference.  gfc_config_observables on a configuration
 * already in device memory (U_dev: 4*V*18 doubles, the AoS image above, on
 * ctx's device), for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It allocates nothing: the per-block partial sums live in a fixed device
 * array of the library (1 MiB per GPU, held until the process ends), so V is
 * limited to 11182080 sites (GFC_EINVAL above that).  All calls on one GPU
 * share that array: calls that overlap on one GPU (other threads or
 * contexts) race.  Same values, bit for bit, as gfc_config_observables. */
int gfc_config_observables_noalloc(gfc_ctx *ctx, const gfc_geom *g, const double *U_dev,
                                   double *e2, double *theta, double *F,
                                   double *plaq_spatial, double *plaq_temporal);

/* ---- serving: continuous batching of many lattices (no counterpart in the
 * reference; each request runs exactly gfc_job_fix's algorithm) ------------- */
typedef struct gfc_server gfc_server;

typedef struct {
    int64_t id;
    uint64_t tag;
    int status;             /* sweep count, GFC_ENOTCONVERGED, or GFC_EPROJECT
                               (input not reunitarizable)                       */
    int sweeps;             /* half-sweeps performed                           */
    int sweeps_fp32;        /* of which in FP32                                 */
    int ns, nt;
    double residue_initial; /* proxy before the first sweep                     */
    double residue_final;   /* proxy at the last check                          */
    double F;               /* gauge functional at the last check               */
    double t_submit, t_admit, t_done; /* seconds on the server clock            */
} gfc_completion;

typedef struct {
    int64_t submitted, admitted, completed;
    int64_t chunks, graphs_built, graph_replays, residue_batches;
    int64_t site_updates;
    int queued, active;
    int64_t bytes_used, bytes_budget;
} gfc_server_stats;

/* A scheduler on ctx's four streams: lattices are served in four lanes by
 * size (fewer than 2^18, 2^20, 2^22 sites, and larger), one stream each, so
 * that small lattices do not advance at the pace of large ones; the large
 * lane sweeps each lattice in the cache-resident timeslice waves of
 * gfc_job_fix, on three more streams of its own.  Seven streams need eight
 * hardware queues (GPU_MAX_HW_QUEUES=8 before the HIP runtime starts): with
 * HIP's default of four, some share a queue and run one after the other.  mem_budget
 * (bytes, <= 0: 85% of the free HBM) bounds the summed footprints of resident
 * lattices; max_jobs bounds their number; keep_results keeps finished
 * lattices resident until released.  gfc_server_poll never waits for a busy
 * lane (it sleeps 20 us when no lane can proceed). */
int gfc_server_create(gfc_ctx *ctx, int64_t mem_budget, int max_jobs, int keep_results,
                      gfc_server **server);
void gfc_server_destroy(gfc_server *server);
/* Seconds since the server was created (the clock of gfc_completion). */
double gfc_server_clock(gfc_server *server);
/* Queues a request: a smooth configuration (gfc_generate_smooth's construction,
 * synthesised in HBM from seed and eps) to be gauge-fixed with p and o.
 * Returns the request id (> 0) or an error code. */
int64_t gfc_server_submit(gfc_server *server, const gfc_geom *g, uint64_t seed, double eps,
                          const gfc_params *p, const gfc_options *o, uint64_t tag);
/* Queues a request for the caller's configuration U (4*V matrices of doubles,
 * in_precision 8, or floats, 4, in the reference's image), copied at once so
 * that the caller may free it.  At admission it is copied to HBM and
 * reunitarized like the CLI does after loadConfig (gauge_fix_coulomb.c:126);
 * a link with a null row ends the request with status GFC_EPROJECT and no
 * sweep.  Returns the request id (> 0) or an error code. */
int64_t gfc_server_submit_config(gfc_server *server, const gfc_geom *g, const void *U,
                                 int in_precision, const gfc_params *p, const gfc_options *o,
                                 uint64_t tag);
/* One scheduling round: admit what fits, evaluate the residues that are due,
 * apply the per-lattice decisions, retire finished lattices and launch one
 * chunk of packed half-sweeps.  Writes up to max_out completions; returns
 * their number or an error code. */
int gfc_server_poll(gfc_server *server, gfc_completion *out, int max_out);
/* Requests queued, resident or completed but not yet returned by poll. */
int gfc_server_pending(gfc_server *server);
int gfc_server_get_stats(gfc_server *server, gfc_server_stats *stats);
/* With keep_results: G (V*18 doubles) of a completed request, and release. */
int gfc_server_result_gauge(gfc_server *server, int64_t id, double *G);
int gfc_server_release(gfc_server *server, int64_t id);

#ifdef __cplusplus
}
#endif
#endif /* GFC_H */
