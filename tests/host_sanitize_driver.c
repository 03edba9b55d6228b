/*
 * Exercises the host C code of libgfc (host_params.c, host_io.c, host_gen.c)
 * for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (tests/test_host_sanitize.py): parameter files with long, empty, malformed
 * and unknown lines, the check schedule at its edges, geometry checks, the
 * synthetic generator into exactly sized buffers, configuration and gauge
 * transformation files in every precision and byte order with short, long
 * and missing files, the sweeps record and the FP32 rounding.  Every buffer
 * is allocated to its exact size so that an overrun is reported.  Prints
 * "OK" and exits 0 when every check holds.
 *
 * usage: host_sanitize_driver TMPDIR
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gfc.h"

static int failures = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            failures++;                                                   \
        }                                                                 \
    } while (0)

static char dir[4096];

static const char *path_of(const char *name) {
    static char buf[8192];
    snprintf(buf, sizeof buf, "%s/%s", dir, name);
    return buf;
}

static void write_text(const char *name, const char *text) {
    FILE *f = fopen(path_of(name), "w");
    if (f) {
        fputs(text, f);
        fclose(f);
    }
}

static void params(void) {
    char longline[400];
    memset(longline, 'x', sizeof longline - 2);
    longline[sizeof longline - 2] = '\n';
    longline[sizeof longline - 1] = '\0';
    const char *texts[] = {
        "tolerance: 1e-12\nhits: 3\nomega_OR: 1.5\nmax_sweeps_to_fix: 7\n"
        "estimate_sweeps_to_gf_progress: 0\nsweeps_to_reunitarization: 10\ngfix_proxy: theta\n",
        "", "no colon\n::\n:\nkey_only:\n: value\nomega_OR:1.2:3\n", longline,
        "omega_OR: 2.0\n", "tolerance: -1\n", "gfix_proxy: e3\n", "sweeps_to_reunitarization: 0\n",
        "hits: 99999999999999999999\nmax_sweeps_to_fix: -5\n",
        "  t o l e r a n c e : 1e-15\n\n\n",
    };
    for (size_t i = 0; i < sizeof texts / sizeof texts[0]; i++) {
        char name[32];
        snprintf(name, sizeof name, "p%zu.par", i);
        write_text(name, texts[i]);
        gfc_params p;
        const int rc = gfc_params_from_file(path_of(name), &p);
        CHECK(rc == 0 || rc == GFC_EPARAMS);
        CHECK((rc == 0) == (gfc_params_valid(&p) != 0));
        gfc_params_print(&p);
    }
    gfc_params p;
    CHECK(gfc_params_from_file(path_of("missing.par"), &p) == 0);
    CHECK(gfc_params_from_file(" ", &p) == 0);
    gfc_params_default(&p);
    const double res[] = {16.0, 1.0, p.tolerance, p.tolerance / 10, 1e-300, 1e300};
    for (size_t i = 0; i < sizeof res / sizeof res[0]; i++)
        for (unsigned est = 0; est <= 1000; est += 1000) {
            p.estimate_sweeps_to_gf_progress = est;
            CHECK(gfc_next_check(0, res[i], &p) >= 0);
            CHECK(gfc_next_check(1084, res[i], &p) >= 1084);
        }
}

static void geometry(void) {
    const int bad[][2] = {{0, 4}, {4, 0}, {3, 4}, {4, 3}, {-2, 4}, {4, -2}};
    for (size_t i = 0; i < sizeof bad / sizeof bad[0]; i++) {
        gfc_geom g = {bad[i][0], bad[i][1]};
        CHECK(gfc_geom_check(&g) == GFC_EGEOMETRY);
    }
    gfc_geom g = {2, 2};
    CHECK(gfc_geom_check(&g) == 0 && gfc_geom_volume(&g) == 16);
}

static void io(void) {
    gfc_geom g = {4, 6};
    const size_t V = (size_t)gfc_geom_volume(&g), nU = V * 4 * 18, nG = V * 18;
    double *U = malloc(nU * sizeof(double)), *U2 = malloc(nU * sizeof(double));
    double *G = malloc(nG * sizeof(double)), *G2 = malloc(nG * sizeof(double));
    float *U32 = malloc(nU * sizeof(float));
    CHECK(gfc_generate_smooth(&g, 9, 0.3, U) == 0);
    gfc_field_to_fp32(U, U32, (int64_t)V * 4);
    for (int prec = 4; prec <= 8; prec += 4)
        for (int swap = 0; swap <= 1; swap++) {
            CHECK(gfc_write_config(path_of("cfg.bin"), &g, prec, swap, U) == 0);
            CHECK(gfc_load_config(path_of("cfg.bin"), &g, prec, swap, U2) == 0);
            for (size_t i = 0; i < nU; i++)
                CHECK(U2[i] == (prec == 8 ? U[i] : (double)U32[i]));
            /* one byte short, one byte long, missing */
            FILE *f = fopen(path_of("cfg.bin"), "rb");
            const size_t bytes = nU * (size_t)prec;
            unsigned char *raw = malloc(bytes + 1);
            CHECK(f && fread(raw, 1, bytes, f) == bytes);
            if (f) fclose(f);
            raw[bytes] = 0;
            f = fopen(path_of("short.bin"), "wb");
            fwrite(raw, 1, bytes - 1, f);
            fclose(f);
            f = fopen(path_of("long.bin"), "wb");
            fwrite(raw, 1, bytes + 1, f);
            fclose(f);
            free(raw);
            CHECK(gfc_load_config(path_of("short.bin"), &g, prec, swap, U2) == -2);
            CHECK(gfc_load_config(path_of("long.bin"), &g, prec, swap, U2) == -3);
            CHECK(gfc_load_config(path_of("none.bin"), &g, prec, swap, U2) == -1);
        }
    CHECK(gfc_load_config(path_of("cfg.bin"), &g, 2, 0, U2) == -4);
    for (size_t s = 0; s < V; s++)
        for (int k = 0; k < 18; k++) G[s * 18 + k] = U[s * 72 + k];
    CHECK(gfc_write_gauge_transf(path_of("gt.bin"), &g, G) == 0);
    CHECK(gfc_load_gauge_transf(path_of("gt.bin"), &g, G2) == 0);
    CHECK(memcmp(G, G2, nG * sizeof(double)) == 0);
    CHECK(gfc_write_sweeps_record(dir, &g, "cfg.bin", 1085) == 0);
    CHECK(gfc_write_sweeps_record(dir, &g, "cfg.bin", -1) == 0);
    free(U);
    free(U2);
    free(G);
    free(G2);
    free(U32);
}

static void generator(void) {
    const int shapes[][2] = {{2, 2}, {4, 8}, {6, 4}};
    for (size_t i = 0; i < sizeof shapes / sizeof shapes[0]; i++) {
        gfc_geom g = {shapes[i][0], shapes[i][1]};
        const size_t n = (size_t)gfc_geom_volume(&g) * 4 * 18;
        double *U = malloc(n * sizeof(double));
        CHECK(gfc_generate_smooth(&g, 5, 0.25, U) == 0);
        for (size_t k = 0; k < n; k++) CHECK(isfinite(U[k]));
        free(U);
    }
    gfc_geom odd = {3, 4};
    CHECK(gfc_generate_smooth(&odd, 1, 0.1, NULL) == GFC_EGEOMETRY);
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    snprintf(dir, sizeof dir, "%s", argv[1]);
    params();
    geometry();
    generator();
    io();
    if (failures) return 1;
    printf("OK\n");
    return 0;
}
