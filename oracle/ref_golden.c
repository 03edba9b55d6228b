/*
 * ref_golden.c -- golden-vector harness around the compiled reference.
 *
 * Built by oracle/Makefile (target `ref`) against objects compiled from the
 * reference's own sources into oracle/_ref/, only where the reference tree is
 * present.  tests/golden/make_golden.py runs it to write the committed
 * fixtures.  It calls the reference's exported functions; its own code only
 * moves data in and out.
 *
 * Modes (all outputs are raw little-endian doubles appended to OUT):
 *   prims N SEED OUT
 *       N random complex 3x3 matrices A (entries uniform in [-1,1]^2), and for
 *       each: A(18) det(2) inverse(18) projectSU3(A)(18) status(1)
 *       power3x3Binomial(A, omega=(double)1.95f)(18) status(1)
 *       power3x3Binomial(A, 1.0)(18) status(1) updateSubLosAlamos R,S,T(3x18)
 *       decomposeAlgebraSU3(A)(9) prod3x3(A, B)(18) with B the next matrix.
 *   obs CFG NS NT OUT
 *       loads CFG (little-endian FP64 build) and writes e2, theta, F,
 *       plaquette spatial, temporal; then reunitarizes (status) and writes
 *       the same five values and the reunitarized U (4V x 18).
 *   traj CFG NS NT R K1 [K2 ...] OUT
 *       reunitarize, then for each K: gaugefixOverrelaxation on a fresh copy
 *       with max_sweeps_to_fix = K (estimate 1e8, tol 1e-300 so no check is
 *       due), hits 2, omega 1.95f, R = sweeps_to_reunitarization; writes K,
 *       return code, e2, theta, F, plaq spatial, plaq temporal, then G (V x 18)
 *       and U (4V x 18) when V <= 4096 else G at sites 0, 7, 13, ... (stride 7,
 *       up to 4096 sites).
 *   trace CFG NS NT R KMAX OUT
 *       the same trajectory as traj, observed after every half-sweep
 *       k = 1..KMAX at the cost of one run (R even): U only evolves through
 *       U (w and the site update never read G), so from the state after k
 *       sweeps (k even) one gaugefixOverrelaxation call with
 *       max_sweeps_to_fix = 1 on a copy gives the state after k + 1 and one
 *       with max_sweeps_to_fix = 2 the state after k + 2 (both calls start
 *       at sweep 0, even parity, and return before their own
 *       reunitarization); the harness then reunitarizes U when k + 2 is a
 *       multiple of R, as the continuous run does after that sweep.  Writes
 *       K, e2, theta, F, plaq spatial, plaq temporal for each K.
 *   full CFG NS NT R EST TOL OUT
 *       reunitarize and fix to tolerance TOL with estimate EST; writes return
 *       code, e2, theta, F, plaq spatial, plaq temporal and G as in traj.
 *   load CFG NS NT OUT
 *       loadConfig only (this build's byte-swap / precision settings); writes
 *       the return code and, on success, U (4V x 18).
 *   params PAR OUT
 *       initORGaugeFixingParameters(PAR): writes tolerance, proxy (0 e2,
 *       1 theta), max_sweeps, estimate, reunit, error, omega, hits.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fields.h>
#include <fields_io.h>
#include <gauge_fixing.h>
#include <geometry.h>
#include <measurement.h>
#include <SU2_ops.h>
#include <SU3_ops.h>

static FILE *out_f;

static void put(double v) { fwrite(&v, sizeof v, 1, out_f); }
static void put_m(const Mtrx3x3 *m) {
    for (int k = 0; k < 9; k++) {
        put(creal(m->m[k]));
        put(cimag(m->m[k]));
    }
}

static unsigned long long rng_state;
static double urand(void) {
    unsigned long long z = (rng_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}

static void rand_m(Mtrx3x3 *m) {
    for (int k = 0; k < 9; k++) {
        double re = urand(), im = urand();
        m->m[k] = re + I * im;
    }
}

static void mode_prims(int n, unsigned long long seed) {
    rng_state = seed;
    Mtrx3x3 A, B, tmp;
    rand_m(&B);
    for (int i = 0; i < n; i++) {
        A = B;
        rand_m(&B);
        /* every 8th input is near SU(3): a projected matrix plus noise */
        if (i % 8 == 7) {
            projectSU3(&A);
            for (int k = 0; k < 9; k++) A.m[k] += 1e-3 * urand();
        }
        put_m(&A);
        Scalar d = determinant3x3(&A);
        put(creal(d));
        put(cimag(d));
        inverse3x3(&A, &tmp);
        put_m(&tmp);
        tmp = A;
        put(projectSU3(&tmp));
        put_m(&tmp);
        put(power3x3Binomial(&A, (double)1.95f, &tmp));
        put_m(&tmp);
        put(power3x3Binomial(&A, 1.0, &tmp));
        put_m(&tmp);
        for (int s = R; s <= T; s++) {
            tmp = A;
            updateSubLosAlamos(&tmp, (Submtrx)s);
            put_m(&tmp);
        }
        MtrxSU3Alg c;
        decomposeAlgebraSU3(&A, &c);
        for (int k = 0; k < 9; k++) put(creal(c.m[k]));
        prod3x3(&A, &B, &tmp);
        put_m(&tmp);
    }
}

static Mtrx3x3 *load_or_die(const char *cfg) {
    Mtrx3x3 *U = allocate3x3Field(DIM * lattice_param.volume);
    int rc = loadConfig(U, (char *)cfg);
    if (rc) {
        fprintf(stderr, "loadConfig failed (%d)\n", rc);
        exit(2);
    }
    return U;
}

static void put_obs(Mtrx3x3 *U) {
    put(calculate_e2(U));
    put(calculateTheta(U));
    put(calculateF(U));
    put(creal(averagePlaquette(U, "spatial")));
    put(creal(averagePlaquette(U, "temporal")));
}

static void put_fields(Mtrx3x3 *U, Mtrx3x3 *G, int with_u) {
    const unsigned V = lattice_param.volume;
    if (V <= 4096) {
        for (unsigned s = 0; s < V; s++) put_m(&G[s]);
        if (with_u)
            for (unsigned s = 0; s < 4 * V; s++) put_m(&U[s]);
    } else {
        int cnt = 0;
        for (unsigned s = 0; s < V && cnt < 4096; s += 7, cnt++) put_m(&G[s]);
    }
}

/* omega 1.95 unless REF_OMEGA is set (fixtures of the degenerate branches of
   the local update need omega = 1.5 or 1.0, where they stay exact). */
static ORGaugeFixingParameters base_params(unsigned R) {
    ORGaugeFixingParameters p = initParametersORDefault();
    const char *om = getenv("REF_OMEGA");
    p.omega_OR = om ? (float)atof(om) : 1.95f;
    p.hits = 2;
    p.generic_gf.sweeps_to_reunitarization = R;
    return p;
}

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: see header of ref_golden.c\n");
        return 1;
    }
    const char *mode = argv[1];
    out_f = fopen(argv[argc - 1], "wb");
    if (!out_f) return 1;
    if (!strcmp(mode, "prims")) {
        mode_prims(atoi(argv[2]), strtoull(argv[3], NULL, 10));
    } else if (!strcmp(mode, "params")) {
        ORGaugeFixingParameters p = initORGaugeFixingParameters(argv[2]);
        put(p.generic_gf.tolerance);
        put(p.generic_gf.gfix_proxy == calculateTheta ? 1 : 0);
        put(p.generic_gf.max_sweeps_to_fix);
        put(p.generic_gf.estimate_sweeps_to_gf_progress);
        put(p.generic_gf.sweeps_to_reunitarization);
        put(p.generic_gf.error);
        put(p.omega_OR);
        put(p.hits);
    } else {
        initGeometry(atoi(argv[3]), atoi(argv[4]));
        const unsigned V = lattice_param.volume;
        if (!strcmp(mode, "load")) {
            Mtrx3x3 *U = allocate3x3Field(DIM * V);
            int rc = loadConfig(U, argv[2]);
            put(rc);
            if (rc == 0)
                for (unsigned s = 0; s < 4 * V; s++) put_m(&U[s]);
            free(U);
        } else if (!strcmp(mode, "obs")) {
            Mtrx3x3 *U = load_or_die(argv[2]);
            put_obs(U);
            put(reunitarizeField(U, DIM * V));
            put_obs(U);
            for (unsigned s = 0; s < 4 * V; s++) put_m(&U[s]);
            free(U);
        } else if (!strcmp(mode, "traj")) {
            Mtrx3x3 *U0 = load_or_die(argv[2]);
            reunitarizeField(U0, DIM * V);
            unsigned R = (unsigned)atoi(argv[5]);
            Mtrx3x3 *U = allocate3x3Field(DIM * V), *G = allocate3x3Field(V);
            for (int a = 6; a < argc - 1; a++) {
                int K = atoi(argv[a]);
                memcpy(U, U0, sizeof(Mtrx3x3) * DIM * V);
                ORGaugeFixingParameters p = base_params(R);
                p.generic_gf.tolerance = 1e-300;
                p.generic_gf.estimate_sweeps_to_gf_progress = 100000000;
                p.generic_gf.max_sweeps_to_fix = (unsigned)K;
                int rc = gaugefixOverrelaxation(U, G, p);
                put(K);
                put(rc);
                put_obs(U);
                put_fields(U, G, 1);
            }
        } else if (!strcmp(mode, "trace")) {
            Mtrx3x3 *U = load_or_die(argv[2]);
            reunitarizeField(U, DIM * V);
            const unsigned R = (unsigned)atoi(argv[5]);
            const int kmax = atoi(argv[6]);
            if (R % 2) {
                fprintf(stderr, "trace needs an even R\n");
                return 1;
            }
            Mtrx3x3 *Uc = allocate3x3Field(DIM * V), *G = allocate3x3Field(V);
            ORGaugeFixingParameters p = base_params(R);
            p.generic_gf.tolerance = 1e-300;
            p.generic_gf.estimate_sweeps_to_gf_progress = 100000000;
            for (int k = 0; k < kmax; k += 2) {
                memcpy(Uc, U, sizeof(Mtrx3x3) * DIM * V);
                p.generic_gf.max_sweeps_to_fix = 1;
                gaugefixOverrelaxation(Uc, G, p);
                put(k + 1);
                put_obs(Uc);
                if (k + 2 > kmax) break;
                p.generic_gf.max_sweeps_to_fix = 2;
                gaugefixOverrelaxation(U, G, p);
                put(k + 2);
                put_obs(U);
                if ((unsigned)(k + 2) % R == 0) reunitarizeField(U, DIM * V);
            }
        } else if (!strcmp(mode, "full")) {
            Mtrx3x3 *U = load_or_die(argv[2]);
            reunitarizeField(U, DIM * V);
            Mtrx3x3 *G = allocate3x3Field(V);
            ORGaugeFixingParameters p = base_params((unsigned)atoi(argv[5]));
            p.generic_gf.estimate_sweeps_to_gf_progress = (unsigned)atoi(argv[6]);
            p.generic_gf.tolerance = atof(argv[7]);
            int rc = gaugefixOverrelaxation(U, G, p);
            put(rc);
            put_obs(U);
            put_fields(U, G, 0);
        } else {
            fprintf(stderr, "unknown mode %s\n", mode);
            return 1;
        }
    }
    fclose(out_f);
    return 0;
}
