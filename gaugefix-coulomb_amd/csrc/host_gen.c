/*
 * host_gen.c -- seeded synthetic "smooth" SU(3) configurations.
 *
 * U_mu(n) = g(n) . exp(i eps H_mu(n)) . g(n + mu)^dagger   (periodic)
 *
 *  - Random numbers: draw k of site n is the splitmix64 finaliser of
 *    base + (64 n + k + 1) * 0x9E3779B97F4A7C15 with base = mix64(seed + that
 *    constant), so every site draws the same numbers whatever the thread
 *    count.  Uniforms are
 *    ((z >> 11) + 1) * 2^-53 in (0, 1]; normals come from Box-Muller pairs.
 *  - g(n): Gram-Schmidt on the rows of a 3x3 matrix of complex normals (Haar
 *    on U(3)), then multiplied by det^(-1/3) (principal branch) to land in
 *    SU(3).  Draws k = 0..17.
 *  - H_mu(n): traceless Hermitian, off-diagonal H_ab = (N + iN')/sqrt 2 for
 *    a < b (unit variance), diagonal N_a minus the mean of the three.
 *    Draws k = 18 + 10 mu .. 27 + 10 mu (the tenth normal is unused).
 *  - exp: scaling and squaring with a 16-term Taylor polynomial, followed by
 *    an SU(3) projection (row 0 normalised, row 2 = conj(r0 x r1) normalised,
 *    row 1 = conj(r2 x r0)).
 */
#define _POSIX_C_SOURCE 200809L
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gfc.h"

typedef struct { double complex m[9]; } mat;

static inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

#define GOLDEN 0x9E3779B97F4A7C15ULL

static inline double uniform(uint64_t base, uint64_t site, unsigned k) {
    uint64_t z = mix64(base + (site * 64 + k + 1) * GOLDEN);
    return (double)((z >> 11) + 1) * 0x1.0p-53;
}

/* normals for draw slots k, k+1 of a site (Box-Muller) */
static inline void normal_pair(uint64_t base, uint64_t site, unsigned k, double *a,
                               double *b) {
    double u1 = uniform(base, site, k), u2 = uniform(base, site, k + 1);
    double r = sqrt(-2.0 * log(u1));
    double ph = 6.283185307179586 * u2;
    *a = r * cos(ph);
    *b = r * sin(ph);
}

static void mprod(const mat *a, const mat *b, mat *c) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double complex s = 0;
            for (int k = 0; k < 3; k++) s += a->m[3 * i + k] * b->m[3 * k + j];
            c->m[3 * i + j] = s;
        }
}

static void project(mat *x) {
    double complex *r0 = x->m, *r1 = x->m + 3, *r2 = x->m + 6;
    double n = 0;
    for (int a = 0; a < 3; a++) n += creal(r0[a] * conj(r0[a]));
    n = 1.0 / sqrt(n);
    for (int a = 0; a < 3; a++) r0[a] *= n;
    r2[0] = conj(r0[1] * r1[2] - r0[2] * r1[1]);
    r2[1] = conj(r0[2] * r1[0] - r0[0] * r1[2]);
    r2[2] = conj(r0[0] * r1[1] - r0[1] * r1[0]);
    n = 0;
    for (int a = 0; a < 3; a++) n += creal(r2[a] * conj(r2[a]));
    n = 1.0 / sqrt(n);
    for (int a = 0; a < 3; a++) r2[a] *= n;
    r1[0] = conj(r2[1] * r0[2] - r2[2] * r0[1]);
    r1[1] = conj(r2[2] * r0[0] - r2[0] * r0[2]);
    r1[2] = conj(r2[0] * r0[1] - r2[1] * r0[0]);
}

static void haar_su3(uint64_t base, uint64_t site, mat *g) {
    double v[18];
    for (unsigned k = 0; k < 18; k += 2) normal_pair(base, site, k, &v[k], &v[k + 1]);
    double complex r[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r[i][j] = v[6 * i + 2 * j] + I * v[6 * i + 2 * j + 1];
    for (int i = 0; i < 3; i++) {
        for (int p = 0; p < i; p++) {
            double complex ip = 0;
            for (int j = 0; j < 3; j++) ip += conj(r[p][j]) * r[i][j];
            for (int j = 0; j < 3; j++) r[i][j] -= ip * r[p][j];
        }
        double n = 0;
        for (int j = 0; j < 3; j++) n += creal(r[i][j] * conj(r[i][j]));
        n = 1.0 / sqrt(n);
        for (int j = 0; j < 3; j++) r[i][j] *= n;
    }
    double complex det = r[0][0] * (r[1][1] * r[2][2] - r[1][2] * r[2][1]) +
                         r[0][1] * (r[1][2] * r[2][0] - r[1][0] * r[2][2]) +
                         r[0][2] * (r[1][0] * r[2][1] - r[1][1] * r[2][0]);
    double complex ph = cexp(-I * carg(det) / 3.0);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) g->m[3 * i + j] = r[i][j] * ph;
}

static void expi_hermitian(uint64_t base, uint64_t site, int mu, double eps, mat *out) {
    double d[10];
    unsigned k0 = 18 + 10 * (unsigned)mu;
    for (unsigned k = 0; k < 10; k += 2) normal_pair(base, site, k0 + k, &d[k], &d[k + 1]);
    const double s2 = 0.7071067811865476;
    mat H;
    double mean = (d[0] + d[1] + d[2]) / 3.0;
    H.m[0] = d[0] - mean;
    H.m[4] = d[1] - mean;
    H.m[8] = d[2] - mean;
    H.m[1] = (d[3] + I * d[4]) * s2;
    H.m[2] = (d[5] + I * d[6]) * s2;
    H.m[5] = (d[7] + I * d[8]) * s2;
    H.m[3] = conj(H.m[1]);
    H.m[6] = conj(H.m[2]);
    H.m[7] = conj(H.m[5]);
    /* A = i eps H, scaled so that its max-row-sum norm is <= 0.5 */
    mat A;
    double nrm = 0;
    for (int i = 0; i < 3; i++) {
        double rs = 0;
        for (int j = 0; j < 3; j++) rs += cabs(H.m[3 * i + j]);
        if (rs > nrm) nrm = rs;
    }
    nrm *= fabs(eps);
    int sq = 0;
    while (nrm > 0.5) {
        nrm *= 0.5;
        sq++;
    }
    double scale = ldexp(eps, -sq);
    for (int k = 0; k < 9; k++) A.m[k] = I * scale * H.m[k];
    /* Taylor: E = sum_{k<=16} A^k / k!  (Horner form) */
    mat E, T;
    for (int k = 0; k < 9; k++) E.m[k] = (k % 4 == 0) ? 1.0 : 0.0;
    for (int n = 16; n >= 1; n--) {
        mprod(&A, &E, &T);
        for (int k = 0; k < 9; k++) E.m[k] = T.m[k] / n + ((k % 4 == 0) ? 1.0 : 0.0);
    }
    for (int s = 0; s < sq; s++) {
        mprod(&E, &E, &T);
        E = T;
    }
    project(&E);
    *out = E;
}

int gfc_generate_smooth(const gfc_geom *g, uint64_t seed, double eps, double *U) {
    if (gfc_geom_check(g) || U == NULL) return GFC_EGEOMETRY;
    const int ns = g->ns, nt = g->nt;
    const int64_t vol = gfc_geom_volume(g);
    const uint64_t base = mix64(seed + GOLDEN);
    mat *gs = (mat *)malloc((size_t)vol * sizeof(mat));
    if (gs == NULL) return GFC_ENOMEM;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < vol; s++) haar_su3(base, (uint64_t)s, &gs[s]);
    mat *out = (mat *)U;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < vol; s++) {
        int x = (int)(s % ns), y = (int)((s / ns) % ns), z = (int)((s / ns / ns) % ns);
        int t = (int)(s / ((int64_t)ns * ns * ns));
        for (int mu = 0; mu < 4; mu++) {
            int c[4] = {t, z, y, x};
            c[mu] += 1;
            if (c[mu] == (mu == 0 ? nt : ns)) c[mu] = 0;
            int64_t sn = (((int64_t)c[0] * ns + c[1]) * ns + c[2]) * ns + c[3];
            mat X, T1, gd;
            expi_hermitian(base, (uint64_t)s, mu, eps, &X);
            mprod(&gs[s], &X, &T1);
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) gd.m[3 * i + j] = conj(gs[sn].m[3 * j + i]);
            mprod(&T1, &gd, &out[s * 4 + mu]);
        }
    }
    free(gs);
    return 0;
}
