/*
 * gfc_oracle.c -- CPU oracle for Coulomb gauge fixing (TEST INFRASTRUCTURE).
 *
 * A C99 restatement of the reference's double-precision algorithm, written so
 * that the floating-point operations happen in the same order as in the
 * reference (C99 complex arithmetic, no FMA contraction: build with
 * -std=c99 -ffp-contract=off).  Each function cites the reference code it
 * restates.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load this; the HIP engine never calls it.
 *
 * Every exported function is pinned by tests/test_oracle_golden.py (fixtures
 * written by the compiled reference, oracle/ref_golden.c) and by
 * tests/test_oracle_math.py (closed forms and invariants).
 */
#include "gfc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define M(a, b) ((a) * 3 + (b))

static inline int64_t site_index(const orc_geom *g, int t, int z, int y, int x) {
    return (((int64_t)t * g->ns + z) * g->ns + y) * g->ns + x;
}

/* Periodic wrap of one coordinate (src/geometry.c:34-59, 62-103). */
static inline int wrap(int a, int n) {
    int r = a % n;
    return r >= 0 ? r : r + n;
}

/* Site index of the neighbour of (t,z,y,x) one step along mu, dir = +1 front,
 * -1 rear (src/geometry.c:108-144). */
static inline int64_t neighbour(const orc_geom *g, int t, int z, int y, int x, int mu,
                                int dir) {
    switch (mu) {
    case 0: t = wrap(t + dir, g->nt); break;
    case 1: z = wrap(z + dir, g->ns); break;
    case 2: y = wrap(y + dir, g->ns); break;
    default: x = wrap(x + dir, g->ns); break;
    }
    return site_index(g, t, z, y, x);
}

/* prod3x3: src/SU3_ops.c:442-476 */
void orc_prod(const orc_mat *u, const orc_mat *v, orc_mat *uv) {
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            uv->m[M(a, b)] = 0.0;
            for (int c = 0; c < 3; c++) uv->m[M(a, b)] += u->m[M(a, c)] * v->m[M(c, b)];
        }
}

/* hermConj3x3: src/SU3_ops.c:284-319 */
static void herm_conj(const orc_mat *u, orc_mat *ud) {
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) ud->m[M(b, a)] = conj(u->m[M(a, b)]);
}

static void set_identity(orc_mat *u) {
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) u->m[M(a, b)] = a != b ? 0.0 : 1.0;
}

/* accumLeftProd3x3 / accumRightProd3x3: src/SU3_ops.c:570-632 */
static void left_mul(const orc_mat *g, orc_mat *acc) {
    orc_mat aux = *acc;
    orc_prod(g, &aux, acc);
}
static void right_mul(orc_mat *acc, const orc_mat *g) {
    orc_mat aux = *acc;
    orc_prod(&aux, g, acc);
}

/* determinant3x3: src/SU3_ops.c:251-281 (Leibniz expansion along row 0) */
double complex orc_det(const orc_mat *u) {
    const double complex *m = u->m;
    return m[M(0, 0)] * (m[M(1, 1)] * m[M(2, 2)] - m[M(1, 2)] * m[M(2, 1)]) +
           m[M(0, 1)] * (m[M(1, 2)] * m[M(2, 0)] - m[M(1, 0)] * m[M(2, 2)]) +
           m[M(0, 2)] * (m[M(1, 0)] * m[M(2, 1)] - m[M(1, 1)] * m[M(2, 0)]);
}

/* inverse3x3: src/SU3_ops.c:843-900 (cofactors divided by the determinant;
 * a zero determinant gives the zero matrix). */
double complex orc_inverse(const orc_mat *u, orc_mat *inv) {
    const double complex *a = u->m;
    double complex *r = inv->m;
    double complex d = orc_det(u);
    if (d != 0.0) {
        r[M(0, 0)] = (a[M(1, 1)] * a[M(2, 2)] - a[M(1, 2)] * a[M(2, 1)]) / d;
        r[M(0, 1)] = (a[M(0, 2)] * a[M(2, 1)] - a[M(0, 1)] * a[M(2, 2)]) / d;
        r[M(0, 2)] = (a[M(0, 1)] * a[M(1, 2)] - a[M(0, 2)] * a[M(1, 1)]) / d;
        r[M(1, 0)] = (a[M(1, 2)] * a[M(2, 0)] - a[M(1, 0)] * a[M(2, 2)]) / d;
        r[M(1, 1)] = (a[M(0, 0)] * a[M(2, 2)] - a[M(0, 2)] * a[M(2, 0)]) / d;
        r[M(1, 2)] = (a[M(0, 2)] * a[M(1, 0)] - a[M(0, 0)] * a[M(1, 2)]) / d;
        r[M(2, 0)] = (a[M(1, 0)] * a[M(2, 1)] - a[M(1, 1)] * a[M(2, 0)]) / d;
        r[M(2, 1)] = (a[M(0, 1)] * a[M(2, 0)] - a[M(0, 0)] * a[M(2, 1)]) / d;
        r[M(2, 2)] = (a[M(0, 0)] * a[M(1, 1)] - a[M(0, 1)] * a[M(1, 0)]) / d;
    } else {
        for (int k = 0; k < 9; k++) r[k] = 0.0;
    }
    return d;
}

/* One Cabibbo-Marinari SU(2) step (updateSubLosAlamos, src/SU3_ops.c:776-840,
 * with projectSU2 src/SU2_ops.c:369-402, convertFromCK :94-130 and
 * accumProdSU2_3x3 src/SU3_ops.c:636-685).  sub: 0=R(0,1), 1=S(0,2), 2=T(1,2).
 * Left-multiplies rows a, b of w by the SU(2) element that maximises ReTr. */
void orc_cm_step(orc_mat *w, int sub) {
    int a = sub == 2 ? 1 : 0;
    int b = sub == 0 ? 1 : 2;
    double ck[4];
    ck[0] = creal(w->m[M(a, a)] + w->m[M(b, b)]);
    ck[1] = -cimag(w->m[M(a, b)] + w->m[M(b, a)]);
    ck[2] = -creal(w->m[M(a, b)] - w->m[M(b, a)]);
    ck[3] = -cimag(w->m[M(a, a)] - w->m[M(b, b)]);

    double complex det = 0.0;
    for (int k = 0; k < 4; k++) det += ck[k] * ck[k];
    if (det == 0.0) return; /* subgroup skipped */
    double complex alpha = 1.0 / csqrt(det);
    for (int k = 0; k < 4; k++) ck[k] = creal(alpha * ck[k]);

    double complex x00 = ck[0] + I * ck[3], x01 = ck[2] + I * ck[1];
    double complex x10 = -ck[2] + I * ck[1], x11 = ck[0] - I * ck[3];
    for (int c = 0; c < 3; c++) {
        double complex r1 = x00 * w->m[M(a, c)] + x01 * w->m[M(b, c)];
        double complex r2 = x10 * w->m[M(a, c)] + x11 * w->m[M(b, c)];
        w->m[M(a, c)] = r1;
        w->m[M(b, c)] = r2;
    }
}

/* normalize3Vec / crossproductConj / projectSU3: src/SU3_ops.c:903-1049.
 * Row 0 normalised; row 2 = conj(r0 x r1) normalised; row 1 = conj(r2 x r0).
 * The squared norm is the square of cabs (tgmath fabs on complex). */
static int normalize3(double complex v[3]) {
    double r = 0.0;
    for (int a = 0; a < 3; a++) r += cabs(v[a]) * cabs(v[a]);
    if (r != 0) {
        r = 1.0 / sqrt(r);
        for (int a = 0; a < 3; a++) v[a] *= r;
        return 0;
    }
    return -1;
}
static void cross_conj(const double complex u[3], const double complex v[3],
                       double complex out[3]) {
    out[0] = conj(u[1] * v[2] - u[2] * v[1]);
    out[1] = conj(u[2] * v[0] - u[0] * v[2]);
    out[2] = conj(u[0] * v[1] - u[1] * v[0]);
}
int orc_project_su3(orc_mat *x) {
    double complex v0[3], v1[3], v2[3];
    for (int b = 0; b < 3; b++) {
        v0[b] = x->m[M(0, b)];
        v1[b] = x->m[M(1, b)];
    }
    int st = normalize3(v0);
    cross_conj(v0, v1, v2);
    st |= normalize3(v2);
    cross_conj(v2, v0, v1);
    for (int b = 0; b < 3; b++) {
        x->m[M(0, b)] = v0[b];
        x->m[M(1, b)] = v1[b];
        x->m[M(2, b)] = v2[b];
    }
    return st == 0 ? 0 : -1;
}

/* power3x3Binomial: src/SU3_ops.c:734-771.  Proj_SU3(1 + omega (A - 1)). */
int orc_power_binomial(const orc_mat *A, double omega, orc_mat *out) {
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            out->m[M(a, b)] = a == b ? 1.0 + omega * (A->m[M(a, b)] - 1.0)
                                     : omega * (A->m[M(a, b)]);
    return orc_project_su3(out);
}

/* decomposeAlgebraSU3: src/SU3_ops.c:1055-1107 (Gell-Mann components). */
void orc_decompose(const orc_mat *A, double c[9]) {
    const double complex *a = A->m;
    c[0] = 0.0;
    c[1] = creal(a[M(0, 1)]) + creal(a[M(1, 0)]);
    c[2] = -cimag(a[M(0, 1)]) + cimag(a[M(1, 0)]);
    c[3] = creal(a[M(0, 0)]) - creal(a[M(1, 1)]);
    c[4] = creal(a[M(0, 2)]) + creal(a[M(2, 0)]);
    c[5] = -cimag(a[M(0, 2)]) + cimag(a[M(2, 0)]);
    c[6] = creal(a[M(1, 2)]) + creal(a[M(2, 1)]);
    c[7] = -cimag(a[M(1, 2)]) + cimag(a[M(2, 1)]);
    c[8] = (creal(a[M(0, 0)]) + creal(a[M(1, 1)]) - 2.0 * creal(a[M(2, 2)])) / sqrt(3.0);
}

/* updateLosAlamos: src/gauge_fixing.c:537-596.  g = w_after_hits * w^-1.
 * inverse3x3 returns its complex determinant converted to double (its real
 * part, src/SU3_ops.c:843), so the hits run only when Re det(w) != 0; when
 * det(w) = 0 the inverse is the zero matrix, so g = 0 (the identity set at
 * :589 is overwritten by the product at :593). */
void orc_la_update(orc_mat *w, unsigned hits, orc_mat *g) {
    orc_mat winv;
    if (creal(orc_inverse(w, &winv)) != 0.0) {
        for (unsigned h = 1; h <= hits; h++)
            for (int sub = 0; sub < 3; sub++) orc_cm_step(w, sub);
    } else {
        set_identity(g);
    }
    orc_prod(w, &winv, g);
}

/* gaugefixOverrelaxationLocal + calculate_w + updateLocalUG:
 * src/gauge_fixing.c:445-531, 600-708. */
void orc_site_update(orc_mat *U, orc_mat *G, const orc_geom *g, int t, int z, int y,
                     int x, double omega, unsigned hits) {
    int64_t s = site_index(g, t, z, y, x);
    orc_mat w;
    for (int k = 0; k < 9; k++) w.m[k] = 0.0;
    for (int mu = 1; mu <= 3; mu++) {
        const orc_mat *uf = &U[s * 4 + mu];
        const orc_mat *ur = &U[neighbour(g, t, z, y, x, mu, -1) * 4 + mu];
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) w.m[M(a, b)] += uf->m[M(a, b)] + conj(ur->m[M(b, a)]);
    }
    orc_mat gla, gor;
    orc_la_update(&w, hits, &gla);
    if (orc_power_binomial(&gla, omega, &gor)) set_identity(&gor);

    left_mul(&gor, &G[s]);
    orc_mat gd;
    herm_conj(&gor, &gd);
    for (int mu = 0; mu < 4; mu++) {
        left_mul(&gor, &U[s * 4 + mu]);
        right_mul(&U[neighbour(g, t, z, y, x, mu, -1) * 4 + mu], &gd);
    }
}

/* The red/black half-sweep of src/gauge_fixing.c:772-787: sites whose parity
 * (x^y^z^t)&1 equals sweep&1 are updated (BLACKRED_SWEEP, :42). */
void orc_half_sweep(orc_mat *U, orc_mat *G, const orc_geom *g, int sweep, double omega,
                    unsigned hits) {
    const int par = sweep & 1;
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < g->nt; t++)
        for (int z = 0; z < g->ns; z++)
            for (int y = 0; y < g->ns; y++)
                for (int x = 0; x < g->ns; x++)
                    if (((x ^ y ^ z ^ t) & 1) == par)
                        orc_site_update(U, G, g, t, z, y, x, omega, hits);
}

/* subtractionHermConjTrless3x3 (src/SU3_ops.c:322-368) followed by
 * multByScalar3x3(-i/2) (calculateA, src/four_potential.c:35-67). */
static void calc_A(const orc_mat *u, orc_mat *A) {
    orc_mat d;
    double complex tr = 0.0;
    for (int a = 0; a < 3; a++) tr += d.m[M(a, a)] = 2 * I * cimag(u->m[M(a, a)]);
    tr /= 3.0;
    for (int a = 0; a < 3; a++) d.m[M(a, a)] -= tr;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            if (a != b) {
                d.m[M(a, b)] = u->m[M(a, b)] - conj(u->m[M(b, a)]);
                d.m[M(b, a)] = u->m[M(b, a)] - conj(u->m[M(a, b)]);
            }
    double complex alpha = -0.5 * I;
    for (int k = 0; k < 9; k++) A->m[k] = alpha * d.m[k];
}

/* divergenceA: src/four_potential.c:70-113 */
static void div_A(const orc_mat *U, const orc_geom *g, int t, int z, int y, int x,
                  orc_mat *dv) {
    int64_t s = site_index(g, t, z, y, x);
    for (int k = 0; k < 9; k++) dv->m[k] = 0.0;
    for (int mu = 1; mu <= 3; mu++) {
        orc_mat a1, a2;
        calc_A(&U[s * 4 + mu], &a1);
        calc_A(&U[neighbour(g, t, z, y, x, mu, -1) * 4 + mu], &a2);
        for (int k = 0; k < 9; k++) dv->m[k] += a1.m[k] - a2.m[k];
    }
}

/* calculate_e2: src/gauge_fixing.c:389-442.  Per-timeslice partial sums are
 * added in t order, so the result does not depend on the thread count. */
double orc_e2(const orc_mat *U, const orc_geom *g) {
    double *part = calloc((size_t)g->nt, sizeof(double));
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < g->nt; t++) {
        double acc = 0.0;
        for (int z = 0; z < g->ns; z++)
            for (int y = 0; y < g->ns; y++)
                for (int x = 0; x < g->ns; x++) {
                    orc_mat dv;
                    double c[9];
                    div_A(U, g, t, z, y, x, &dv);
                    orc_decompose(&dv, c);
                    for (int a = 1; a <= 8; a++) acc += c[a] * c[a];
                }
        part[t] = acc;
    }
    double e2 = 0.0;
    for (int t = 0; t < g->nt; t++) e2 += part[t];
    free(part);
    double vol = (double)g->ns * g->ns * g->ns * g->nt;
    return e2 / vol;
}

/* calculateTheta: src/gauge_fixing.c:341-384 (a serial loop there). */
double orc_theta(const orc_mat *U, const orc_geom *g) {
    double theta = 0.0;
    for (int t = 0; t < g->nt; t++)
        for (int z = 0; z < g->ns; z++)
            for (int y = 0; y < g->ns; y++)
                for (int x = 0; x < g->ns; x++) {
                    orc_mat dv, dvd, p;
                    div_A(U, g, t, z, y, x, &dv);
                    herm_conj(&dv, &dvd);
                    orc_prod(&dv, &dvd, &p);
                    theta += creal(p.m[0] + p.m[4] + p.m[8]);
                }
    double vol = (double)g->ns * g->ns * g->ns * g->nt;
    return theta / (3.0 * vol);
}

/* calculateF: src/gauge_fixing.c:289-337 (serial accumulation of all spatial
 * links into one matrix, then Re Tr / (9 V)). */
double orc_functional(const orc_mat *U, const orc_geom *g) {
    orc_mat acc;
    for (int k = 0; k < 9; k++) acc.m[k] = 0.0;
    for (int t = 0; t < g->nt; t++)
        for (int z = 0; z < g->ns; z++)
            for (int y = 0; y < g->ns; y++)
                for (int x = 0; x < g->ns; x++) {
                    int64_t s = site_index(g, t, z, y, x);
                    for (int mu = 1; mu <= 3; mu++)
                        for (int k = 0; k < 9; k++) acc.m[k] += U[s * 4 + mu].m[k];
                }
    double vol = (double)g->ns * g->ns * g->ns * g->nt;
    return creal(acc.m[0] + acc.m[4] + acc.m[8]) / (vol * 3 * 3);
}

/* TrPlaquette / averagePlaquettegeneric / averagePlaquette:
 * src/measurement.c:38-155.  which: 0 = spatial (XY, XZ, YZ), 1 = temporal
 * (TZ, TY, TX). */
static double complex plaq_plane(const orc_mat *U, const orc_geom *g, int mu, int nu) {
    double complex *part = calloc((size_t)g->nt, sizeof(double complex));
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < g->nt; t++) {
        double complex acc = 0.0;
        for (int z = 0; z < g->ns; z++)
            for (int y = 0; y < g->ns; y++)
                for (int x = 0; x < g->ns; x++) {
                    int64_t s = site_index(g, t, z, y, x);
                    int64_t s_mu = neighbour(g, t, z, y, x, mu, +1);
                    int64_t s_nu = neighbour(g, t, z, y, x, nu, +1);
                    orc_mat ua = U[s * 4 + mu], ub = U[s_mu * 4 + nu], uc, ud;
                    /* uc = U_mu(n + nu)^dagger, ud = U_nu(n)^dagger */
                    herm_conj(&U[s_nu * 4 + mu], &uc);
                    herm_conj(&U[s * 4 + nu], &ud);
                    double complex tr = 0.0;
                    for (int a = 0; a < 3; a++) {
                        double complex r = 0.0;
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++)
                                for (int k = 0; k < 3; k++)
                                    r += ua.m[M(a, i)] * ub.m[M(i, j)] * uc.m[M(j, k)] *
                                         ud.m[M(k, a)];
                        tr += r;
                    }
                    acc += tr / 3;
                }
        part[t] = acc;
    }
    double complex sum = 0.0;
    for (int t = 0; t < g->nt; t++) sum += part[t];
    free(part);
    double vol = (double)g->ns * g->ns * g->ns * g->nt;
    return sum / vol;
}

double orc_plaquette(const orc_mat *U, const orc_geom *g, int which) {
    double avg = 0.0;
    int count = 0;
    if (which == 1) {
        for (int i = 1; i <= 3; i++) {
            avg += creal(plaq_plane(U, g, 0, i));
            count++;
        }
    } else {
        avg += creal(plaq_plane(U, g, 3, 2)); count++;
        avg += creal(plaq_plane(U, g, 3, 1)); count++;
        avg += creal(plaq_plane(U, g, 2, 1)); count++;
    }
    return avg / (double)count;
}

/* reunitarizeField: src/fields.c:144-187 */
int orc_reunitarize(orc_mat *field, int64_t n) {
    int st = 0;
#pragma omp parallel for reduction(| : st) schedule(static)
    for (int64_t i = 0; i < n; i++) st |= orc_project_su3(&field[i]);
    return st != 0 ? -1 : 0;
}

/* setFieldToIdentity: src/fields.c:64-100 */
void orc_set_identity(orc_mat *field, int64_t n) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) set_identity(&field[i]);
}

/* applyGaugeTransformationU + prod_vuwdagger3x3: src/fields.c:375-435,
 * src/SU3_ops.c:688-730.  U_mu(n) <- G(n) U_mu(n) G(n+mu)^dagger. */
void orc_apply_gauge(orc_mat *U, const orc_mat *G, const orc_geom *g) {
#pragma omp parallel for schedule(dynamic)
    for (int t = 0; t < g->nt; t++)
        for (int z = 0; z < g->ns; z++)
            for (int y = 0; y < g->ns; y++)
                for (int x = 0; x < g->ns; x++) {
                    int64_t s = site_index(g, t, z, y, x);
                    for (int mu = 0; mu < 4; mu++) {
                        const orc_mat *v = &G[s], *w = &G[neighbour(g, t, z, y, x, mu, +1)];
                        orc_mat *u = &U[s * 4 + mu], r;
                        for (int a = 0; a < 3; a++)
                            for (int b = 0; b < 3; b++) {
                                r.m[M(a, b)] = 0.0;
                                for (int c = 0; c < 3; c++)
                                    for (int d = 0; d < 3; d++)
                                        r.m[M(a, b)] += v->m[M(a, c)] * u->m[M(c, d)] *
                                                        conj(w->m[M(b, d)]);
                            }
                        *u = r;
                    }
                }
}

/* whenNextConvergenceCheckQ: src/gauge_fixing.c:248-285 */
int orc_next_check(unsigned sweep, double res, const orc_params *p) {
    if (res < p->tolerance) return (int)sweep;
    return (int)(sweep + 10 +
                 (unsigned)(p->estimate * (1.0 - log10(res) / log10(p->tolerance))));
}

/* validORGaugeFixingParametersQ: src/gauge_fixing.c:45-70 */
int orc_params_valid(const orc_params *p) {
    return p->omega >= 1 && p->omega < 2 && p->tolerance > 0 && p->reunit_every > 0;
}

static double proxy(const orc_mat *U, const orc_geom *g, int which) {
    return which == 1 ? orc_theta(U, g) : orc_e2(U, g);
}

/* gaugefixOverrelaxation: src/gauge_fixing.c:713-826.  When trace is not NULL,
 * trace[k-1] receives the proxy after half-sweep k (before that sweep's
 * reunitarization), for k <= trace_len. */
int orc_gaugefix(orc_mat *U, orc_mat *G, const orc_geom *g, const orc_params *p,
                 double *trace, int trace_len, int verbose) {
    if (!orc_params_valid(p)) return -2;
    const int64_t vol = (int64_t)g->ns * g->ns * g->ns * g->nt;
    const double omega = (double)p->omega;
    orc_set_identity(G, vol);
    int sweep = 0;
    double res = proxy(U, g, p->proxy);
    if (verbose) printf("Sweeps: %8d.\t residue: %3.2E \n", sweep, res);
    int due = orc_next_check((unsigned)sweep, res, p);
    while (res > p->tolerance) {
        orc_half_sweep(U, G, g, sweep, omega, p->hits);
        sweep++;
        if (trace && sweep <= trace_len) trace[sweep - 1] = proxy(U, g, p->proxy);
        if (sweep == due) {
            double last = res;
            res = proxy(U, g, p->proxy);
            if (fabs(res - last) / last < 10E-16) {
                if (verbose) printf("Sweeps: %8d.\t residue: %3.2E \n", sweep, res);
                return -1;
            } else if (res <= p->tolerance) {
                break;
            } else {
                if (verbose) printf("Sweeps: %8d.\t residue: %3.2E \n", sweep, res);
                due = orc_next_check((unsigned)sweep, res, p);
            }
        }
        if ((unsigned)sweep >= p->max_sweeps) return -1;
        if (!((unsigned)sweep % p->reunit_every)) {
            orc_reunitarize(U, 4 * vol);
            orc_reunitarize(G, vol);
        }
    }
    orc_reunitarize(U, 4 * vol);
    orc_reunitarize(G, vol);
    return sweep;
}
