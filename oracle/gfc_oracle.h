/*
 * gfc_oracle.h -- CPU oracle for Coulomb gauge fixing (TEST INFRASTRUCTURE).
 *
 * This is a plain C99 restatement of the reference's double-precision CPU path
 * (jesuel-marques/GaugeFix-Coulomb, src/gauge_fixing.c, src/SU3_ops.c,
 * src/SU2_ops.c, src/four_potential.c, src/fields.c, src/measurement.c).
 * It exists only to check the HIP engine: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.  The product path never calls it.
 *
 * Memory layout is the reference's file/memory image (include/fields.h:33-43):
 *   site  = ((t*ns + z)*ns + y)*ns + x
 *   link  = site*4 + mu,   mu: 0=T, 1=Z, 2=Y, 3=X   (settings.h:57-60)
 *   each matrix = 9 complex doubles, row-major, (re, im) interleaved.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   - every function below is compared with the compiled reference
 *     (oracle/_ref, built from the reference sources) through the committed
 *     fixtures in tests/golden/, and additionally against closed forms and
 *     invariants (SU(3) membership, gauge invariance of the plaquette,
 *     theta = e2/6, w*w^-1 = 1, brute-force maximisation of the SU(2) step).
 */
#ifndef GFC_ORACLE_H
#define GFC_ORACLE_H

#include <complex.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double complex m[9]; } orc_mat;

typedef struct {
    int ns, nt;
} orc_geom;

/* Mirrors ORGaugeFixingParameters (include/gauge_fixing.h:31-54).  omega is a
 * float there, so the effective value is (double)(float)omega. */
typedef struct {
    double tolerance;
    int proxy;                 /* 0 = e2, 1 = theta */
    unsigned max_sweeps;
    unsigned estimate;         /* estimate_sweeps_to_gf_progress */
    unsigned reunit_every;     /* sweeps_to_reunitarization */
    float omega;
    unsigned hits;
} orc_params;

/* --- group algebra (src/SU3_ops.c, src/SU2_ops.c) --- */
void orc_prod(const orc_mat *a, const orc_mat *b, orc_mat *ab);
double complex orc_det(const orc_mat *a);
double complex orc_inverse(const orc_mat *a, orc_mat *inv);
void orc_cm_step(orc_mat *w, int sub);
int orc_project_su3(orc_mat *x);
int orc_power_binomial(const orc_mat *a, double omega, orc_mat *out);
void orc_decompose(const orc_mat *a, double c[9]);
void orc_la_update(orc_mat *w, unsigned hits, orc_mat *g);

/* --- lattice operations (src/gauge_fixing.c, src/fields.c, src/measurement.c) --- */
void orc_site_update(orc_mat *U, orc_mat *G, const orc_geom *g, int t, int z, int y,
                     int x, double omega, unsigned hits);
void orc_half_sweep(orc_mat *U, orc_mat *G, const orc_geom *g, int sweep, double omega,
                    unsigned hits);
double orc_e2(const orc_mat *U, const orc_geom *g);
double orc_theta(const orc_mat *U, const orc_geom *g);
double orc_functional(const orc_mat *U, const orc_geom *g);
double orc_plaquette(const orc_mat *U, const orc_geom *g, int which); /* 0 spatial, 1 temporal */
int orc_reunitarize(orc_mat *field, int64_t n);
void orc_set_identity(orc_mat *field, int64_t n);
void orc_apply_gauge(orc_mat *U, const orc_mat *G, const orc_geom *g);
int orc_next_check(unsigned sweep, double res, const orc_params *p);
int orc_params_valid(const orc_params *p);
int orc_gaugefix(orc_mat *U, orc_mat *G, const orc_geom *g, const orc_params *p,
                 double *trace, int trace_len, int verbose);

#ifdef __cplusplus
}
#endif
#endif
