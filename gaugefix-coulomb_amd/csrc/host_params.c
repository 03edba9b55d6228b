/*
 * host_params.c -- parameters.par handling, check schedule and geometry checks.
 *
 * Behaviour follows src/gauge_fixing.c:45-285 of the reference: same keys,
 * same defaults, same validity rule, same "whitespace removed, split at the
 * first ':'" line format, same stderr messages for unknown keys.
 */
#define _POSIX_C_SOURCE 200809L
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gfc.h"

/* initParametersORDefault, src/gauge_fixing.c:72-84 */
void gfc_params_default(gfc_params *p) {
    p->omega_OR = 1.95f;
    p->hits = 2;
    p->error = 0;
    p->gfix_proxy = GFC_PROXY_E2;
    p->tolerance = 1E-16;
    p->max_sweeps_to_fix = 100000;
    p->estimate_sweeps_to_gf_progress = 1000;
    p->sweeps_to_reunitarization = 250;
}

/* validORGaugeFixingParametersQ, src/gauge_fixing.c:45-70 */
int gfc_params_valid(const gfc_params *p) {
    return p->omega_OR >= 1 && p->omega_OR < 2 && p->tolerance > 0 && !p->error &&
           p->sweeps_to_reunitarization > 0;
}

/* printORGaugeFixingParameters, src/gauge_fixing.c:86-104 */
void gfc_params_print(const gfc_params *p) {
    printf("\n");
    printf("max_hits: %u\n", p->hits);
    printf("omega_OR: %lf\n", (double)p->omega_OR);
    printf("\n");
    printf("tolerance: %3.2E\n", p->tolerance);
    printf("max_sweeps_to_fix: %u\n", p->max_sweeps_to_fix);
    printf("estimate_sweeps_to_gfixprogress: %u\n", p->estimate_sweeps_to_gf_progress);
    printf("\n");
    printf("sweeps_to_reunitarization: %u\n", p->sweeps_to_reunitarization);
    printf("\n");
}

static void strip_whitespace(char *s) {
    size_t k = 0;
    for (size_t i = 0; s[i]; i++)
        if (!isspace((unsigned char)s[i])) s[k++] = s[i];
    s[k] = '\0';
}

/* atof into an unsigned field, as the reference's implicit conversion does. */
static unsigned as_unsigned(const char *v) {
    double d = atof(v);
    return d > 0 ? (unsigned)d : 0u;
}

/* initORGaugeFixingParameters, src/gauge_fixing.c:144-244.  Lines are read
 * in 100-byte pieces (fgets with a 100-byte buffer); a piece without ':' or
 * without a value is ignored. */
int gfc_params_from_file(const char *path, gfc_params *p) {
    gfc_params_default(p);
    if (path == NULL || !strcmp(path, " ")) {
        printf("No parameter-file detected, using default values.\n");
        return 0;
    }
    FILE *f = fopen(path, "r");
    if (f == NULL) {
        fprintf(stderr,
                "Gauge-fixing parameter file %s could not be opened. "
                "Using default parameters instead.\n",
                path);
        gfc_params_print(p);
    } else {
        char buf[100];
        while (fgets(buf, sizeof buf, f)) {
            strip_whitespace(buf);
            char *colon = strchr(buf, ':');
            if (colon == NULL) continue;
            /* strtok semantics: leading ':' are skipped before the key */
            char *key = buf;
            while (*key == ':') key++;
            colon = strchr(key, ':');
            if (colon == NULL) continue;
            *colon = '\0';
            char *value = colon + 1;
            if (*value == '\0') continue;
            if (!strcmp(key, "tolerance")) {
                p->tolerance = atof(value);
            } else if (!strcmp(key, "hits")) {
                p->hits = as_unsigned(value);
            } else if (!strcmp(key, "omega_OR")) {
                p->omega_OR = (float)atof(value);
            } else if (!strcmp(key, "max_sweeps_to_fix")) {
                p->max_sweeps_to_fix = as_unsigned(value);
            } else if (!strcmp(key, "estimate_sweeps_to_gf_progress")) {
                p->estimate_sweeps_to_gf_progress = as_unsigned(value);
            } else if (!strcmp(key, "sweeps_to_reunitarization")) {
                p->sweeps_to_reunitarization = as_unsigned(value);
            } else if (!strcmp(key, "gfix_proxy")) {
                if (!strcmp(value, "e2"))
                    p->gfix_proxy = GFC_PROXY_E2;
                else if (!strcmp(value, "theta"))
                    p->gfix_proxy = GFC_PROXY_THETA;
                else
                    fprintf(stderr, "Unrecognized gfix_proxy: %s.\n", value);
            } else {
                fprintf(stderr, "Unrecognized parameter: %s.\n", key);
            }
        }
        fclose(f);
    }
    if (!gfc_params_valid(p)) {
        p->error = 1;
        return GFC_EPARAMS;
    }
    return 0;
}

/* whenNextConvergenceCheckQ, src/gauge_fixing.c:248-285 */
int gfc_next_check(unsigned sweep, double residue, const gfc_params *p) {
    if (residue < p->tolerance) return (int)sweep;
    return (int)(sweep + 10 +
                 (unsigned)(p->estimate_sweeps_to_gf_progress *
                            (1.0 - log10(residue) / log10(p->tolerance))));
}

/* initGeometry / validGeometricParametersQ, src/geometry.c:148-228, plus the
 * even-extent requirement of the red/black decomposition. */
int gfc_geom_check(const gfc_geom *g) {
    if (g == NULL || g->ns <= 0 || g->nt <= 0) return GFC_EGEOMETRY;
    if ((g->ns & 1) || (g->nt & 1)) return GFC_EGEOMETRY;
    if (g->ns > 32767 || g->nt > 32767) return GFC_EGEOMETRY;
    return 0;
}

int64_t gfc_geom_volume(const gfc_geom *g) {
    return (int64_t)g->ns * g->ns * g->ns * g->nt;
}
