/*
 * host_io.c -- headerless SCIDAC binary I/O and the sweeps record file.
 *
 * File formats follow src/fields_io.c of the reference: a configuration is the
 * raw image of 4*V 3x3 complex matrices (site-major, mu = T,Z,Y,X), a gauge
 * transformation the raw image of V matrices; reals are 4 or 8 bytes.  The
 * configuration reader checks that the file ends exactly after 4*V matrices.
 * Files are streamed in blocks so that FP32 input and byte swapping never
 * need a second full-size host buffer.
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gfc.h"

#define BLOCK_REALS (1 << 20)

static void swap_bytes(void *buf, size_t n_elems, size_t size) {
    unsigned char *p = (unsigned char *)buf;
    for (size_t i = 0; i < n_elems; i++, p += size)
        for (size_t a = 0, b = size - 1; a < b; a++, b--) {
            unsigned char c = p[a];
            p[a] = p[b];
            p[b] = c;
        }
}

/* loadConfig (src/fields_io.c:216-305) with the FP32 input conversion of
 * convertCfg_in_work (src/convert_precision_io.c:176-215). */
int gfc_load_config(const char *path, const gfc_geom *g, int in_precision, int byte_swap,
                    double *U) {
    if (in_precision != 4 && in_precision != 8) return -4;
    if (gfc_geom_check(g) || U == NULL) return GFC_EINVAL;
    FILE *f = fopen(path, "rb");
    if (f == NULL) {
        fprintf(stderr, "Error: Problem opening config file %s.\n", path);
        return -1;
    }
    const size_t total = (size_t)gfc_geom_volume(g) * 4 * 18;
    void *blk = malloc((size_t)BLOCK_REALS * in_precision);
    if (blk == NULL) {
        fclose(f);
        return GFC_ENOMEM;
    }
    size_t done = 0;
    while (done < total) {
        size_t n = total - done < BLOCK_REALS ? total - done : BLOCK_REALS;
        if (fread(blk, in_precision, n, f) != n) {
            fprintf(stderr, "Error: Reading from file %s.\n", path);
            free(blk);
            fclose(f);
            return -2;
        }
        if (byte_swap) swap_bytes(blk, n, in_precision);
        if (in_precision == 8) {
            memcpy(U + done, blk, n * sizeof(double));
        } else {
            const float *src = (const float *)blk;
            for (size_t i = 0; i < n; i++) U[done + i] = (double)src[i];
        }
        done += n;
    }
    free(blk);
    fgetc(f);
    if (!feof(f)) {
        fprintf(stderr, "Error: File has not been read till the end.\n Check lattice sizes.\n");
        fclose(f);
        return -3;
    }
    fclose(f);
    return 0;
}

/* writeConfig (src/fields_io.c:405-489). */
int gfc_write_config(const char *path, const gfc_geom *g, int out_precision, int byte_swap,
                     const double *U) {
    if (out_precision != 4 && out_precision != 8) return -4;
    if (gfc_geom_check(g) || U == NULL) return GFC_EINVAL;
    FILE *f = fopen(path, "wb");
    if (f == NULL) {
        fprintf(stderr, "Error: Problem creating file %s for config.\n", path);
        return -1;
    }
    const size_t total = (size_t)gfc_geom_volume(g) * 4 * 18;
    void *blk = malloc((size_t)BLOCK_REALS * out_precision);
    if (blk == NULL) {
        fclose(f);
        return GFC_ENOMEM;
    }
    int rc = 0;
    for (size_t done = 0; done < total;) {
        size_t n = total - done < BLOCK_REALS ? total - done : BLOCK_REALS;
        if (out_precision == 8) {
            memcpy(blk, U + done, n * sizeof(double));
        } else {
            float *dst = (float *)blk;
            for (size_t i = 0; i < n; i++) dst[i] = (float)U[done + i];
        }
        if (byte_swap) swap_bytes(blk, n, out_precision);
        if (fwrite(blk, out_precision, n, f) != n) {
            rc = -2;
            break;
        }
        done += n;
    }
    free(blk);
    fclose(f);
    return rc;
}

/* writeGaugeTransf (src/fields_io.c:493-565): native-endian FP64, never
 * swapped. */
int gfc_write_gauge_transf(const char *path, const gfc_geom *g, const double *G) {
    if (gfc_geom_check(g) || G == NULL) return GFC_EINVAL;
    FILE *f = fopen(path, "wb");
    if (f == NULL) return -1;
    const size_t n = (size_t)gfc_geom_volume(g) * 18;
    if (fwrite(G, sizeof(double), n, f) != n) {
        fclose(f);
        return -2;
    }
    fclose(f);
    return 0;
}

/* loadGaugeTransf (src/fields_io.c:309-401), FP64 native-endian. */
int gfc_load_gauge_transf(const char *path, const gfc_geom *g, double *G) {
    if (gfc_geom_check(g) || G == NULL) return GFC_EINVAL;
    FILE *f = fopen(path, "rb");
    if (f == NULL) {
        fprintf(stderr, "Error: Problem opening file %s to load gauge transformation.\n",
                path);
        return -1;
    }
    const size_t n = (size_t)gfc_geom_volume(g) * 18;
    if (fread(G, sizeof(double), n, f) != n) {
        fprintf(stderr, "Error: Problem reading file %s to load gauge transformation.\n",
                path);
        fclose(f);
        return -2;
    }
    fgetc(f);
    if (!feof(f)) {
        fprintf(stderr, "Error: File has not been read till the end.\nCheck lattice sizes.\n");
        fclose(f);
        return -2;
    }
    fclose(f);
    return 0;
}

/* writeSweepsToGaugefix (gauge_fix_coulomb.c:41-65). */
int gfc_write_sweeps_record(const char *dir, const gfc_geom *g, const char *identifier,
                            int sweeps) {
    if (gfc_geom_check(g) || identifier == NULL) return GFC_EINVAL;
    char name[4096];
    if (dir && *dir)
        snprintf(name, sizeof name, "%s/sweeps_to_gaugefix_%dx%d.txt", dir, g->ns, g->nt);
    else
        snprintf(name, sizeof name, "sweeps_to_gaugefix_%dx%d.txt", g->ns, g->nt);
    FILE *f = fopen(name, "a+");
    if (f == NULL) {
        fprintf(stderr, "Error opening file %s to record sweeps needed to gaugefix.\n", name);
        return -1;
    }
    fprintf(f, "%s\t%d\n", identifier, sweeps);
    fclose(f);
    return 0;
}

void gfc_field_to_fp32(const double *in, float *out, int64_t n_matrices) {
    const int64_t n = n_matrices * 18;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) out[i] = (float)in[i];
}
