/*
 * gauge_fix_coulomb -- drop-in command line for the MI355X engine.
 *
 *   gauge_fix_coulomb config_path gtransformation_path ns nt [parameters.par]
 *
 * Same flow, messages, outputs and exit codes as the reference's main
 * (gauge_fix_coulomb.c:67-211 of jesuel-marques/GaugeFix-Coulomb): load the
 * SCIDAC configuration, reunitarize it, print the plaquettes, gauge-fix,
 * print the plaquettes again, append the sweep count to
 * sweeps_to_gaugefix_<ns>x<nt>.txt and write G.  Exit 1 (and no G file) when
 * loading, reunitarization or fixing fails.
 *
 * The reference's compile-time switches (settings.h) are runtime
 * environment variables here, with the same defaults:
 *   GFC_BYTE_SWAP_IN      1  swap the byte order of the input reals
 *                            (NEED_BYTE_SWAP_IN, defined in settings.h)
 *   GFC_INPUT_PRECISION   8  4 reads FP32 input (CONV_CFG_TO_WORKING_PRECISION)
 * Engine options:
 *   GFC_PRECISION   fp64     "mixed": FP32 sweeps, then FP64 to the tolerance
 *   GFC_MIXED_SWITCH 1e-9    residue at the FP32 -> FP64 hand-off
 *   GFC_DEVICE      0        HIP device index
 *   GFC_METRICS     0        1: print one JSON metrics line at the end
 */
#define _POSIX_C_SOURCE 200809L
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gfc.h"

static long env_long(const char *name, long dflt) {
    const char *v = getenv(name);
    return (v && *v) ? strtol(v, NULL, 10) : dflt;
}

static double env_double(const char *name, double dflt) {
    const char *v = getenv(name);
    return (v && *v) ? strtod(v, NULL) : dflt;
}

/* printMatrix3x3, src/SU3_ops.c:30-46 */
static void print_matrix(const double *m) {
    printf("{");
    for (int a = 0; a < 3; a++) {
        printf("{");
        for (int b = 0; b < 3; b++) {
            printf("%.*lf + I*(%.*lf)", 18, m[2 * (3 * a + b)], 18, m[2 * (3 * a + b) + 1]);
            if (b != 2) printf(", ");
        }
        if (a != 2) printf("},\n");
    }
    printf("}}\n\n");
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char *argv[]) {
    if (argc < 5) {
        fprintf(stderr,
                "Bad input.\n"
                "Usage: gauge_fix_coulomb config_path gtransformation_path ns nt "
                "[parameters.par]\n");
        return EXIT_FAILURE;
    }
    const char *config_filename = argv[1];
    const char *gt_filename = argv[2];
    gfc_geom geom = {atoi(argv[3]), atoi(argv[4])};

    gfc_params p;
    if (argc == 6)
        gfc_params_from_file(argv[5], &p);
    else
        gfc_params_from_file(" ", &p);

    if ((argc != 6 && argc != 5) || !gfc_params_valid(&p) || gfc_geom_check(&geom)) {
        fprintf(stderr,
                "Bad input.\n"
                "Usage: Input config filename, gt filename, n_SPC, n_T, "
                "tolerance and omega_OR in this order.\n"
                "Check that:\n"
                "1 < omega_OR < 2\n"
                "n_SPC > 0 and n_T >0\n"
                "tolerance > 0 \n");
        fprintf(stderr,
                "\n Provided parameters: \n"
                "n_SPC: %d \nn_T: %d \n"
                "omega_OR : %lf \n"
                "tolerance : %3.2E \n",
                geom.ns, geom.nt, (double)p.omega_OR, p.tolerance);
        fprintf(stderr, "Initializing gauge-fixing parameters to default instead.\n");
        gfc_params_default(&p);
        if (gfc_geom_check(&geom)) {
            fprintf(stderr, "ns and nt must be positive and even.\n");
            return EXIT_FAILURE;
        }
    }
    printf("\nGauge-fixing parameters provided:\n");
    gfc_params_print(&p);

    const int swap = (int)env_long("GFC_BYTE_SWAP_IN", 1);
    const int in_prec = (int)env_long("GFC_INPUT_PRECISION", 8);
    gfc_options opt;
    gfc_options_default(&opt);
    const char *prec = getenv("GFC_PRECISION");
    if (prec && !strcmp(prec, "mixed")) opt.precision = GFC_PREC_MIXED;
    opt.mixed_switch = env_double("GFC_MIXED_SWITCH", opt.mixed_switch);
    opt.verbose = 1;
    const int device = (int)env_long("GFC_DEVICE", 0);

    const double t_start = now_s();
    const int64_t vol = gfc_geom_volume(&geom);
    gfc_ctx *ctx = NULL;
    gfc_job *job = NULL;
    if (gfc_ctx_create(device, &ctx) || gfc_job_create(ctx, &geom, opt.precision, &job)) {
        fprintf(stderr, "Device setup failed: %s\n", gfc_last_error());
        gfc_job_destroy(job);
        gfc_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
    /* the file streams straight into HBM (byte swap and conversion on the
     * device); the loader's rules and messages are those of loadConfig */
    const double t_load = now_s();
    if (gfc_job_load_config(job, config_filename, in_prec, swap)) {
        fprintf(stderr, "Loading of file %s failed.\n", config_filename);
        gfc_job_destroy(job);
        gfc_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
    const double load_seconds = now_s() - t_load;
    printf("Config from file %s loaded.\n", config_filename);
    double link0[18];
    gfc_job_get_link(job, 0, 0, link0);
    print_matrix(link0);

    /* Reunitarizing right away (gauge_fix_coulomb.c:124-131) */
    if (gfc_job_reunitarize(job)) {
        fprintf(stderr, "Configuration in file %s could not be reunitarized.\n",
                config_filename);
        gfc_job_destroy(job);
        gfc_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
    double link[18];
    gfc_job_get_link(job, 0, 0, link);
    print_matrix(link);

    double ps, pt;
    gfc_job_plaquette(job, &ps, &pt);
    printf("Average trace of spatial plaquettes before gauge transformation: %.10lf\n", ps);
    printf("Average trace of temporal plaquettes before gauge transformation: %.10lf\n", pt);

    gfc_report rep;
    const int sweeps = gfc_job_fix(job, &p, &opt, &rep);

    gfc_job_plaquette(job, &ps, &pt);
    printf("Average trace of spatial plaquettes after gauge transformation: %.10lf\n", ps);
    printf("Average trace of temporal plaquettes after gauge transformation: %.10lf\n", pt);

    if (sweeps < 0) {
        switch (sweeps) {
        case GFC_ENOTCONVERGED:
            fprintf(stderr,
                    "Configuration in file %s could not be gauge-fixed "
                    "within the maximum number of sweeps passed.\n"
                    "SOR algorithm seems not to work or be slower than "
                    "what the user expected for this particular config. \n",
                    config_filename);
            break;
        case GFC_EPARAMS:
            fprintf(stderr, "Error in parameters for gauge fixing\n");
            break;
        default:
            fprintf(stderr, "Device error: %s\n", gfc_last_error());
            break;
        }
        gfc_job_destroy(job);
        gfc_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }

    double e2, theta, F;
    gfc_job_observables(job, &e2, &theta, &F);
    printf("Sweeps needed to gauge-fix config from file %s: %d. residue: %3.2E \n",
           config_filename, sweeps, p.gfix_proxy == GFC_PROXY_THETA ? theta : e2);
    gfc_write_sweeps_record(NULL, &geom, config_filename, sweeps);

    const double t_write = now_s();
    const int wrc = gfc_job_write_gauge_transf(job, gt_filename);
    const double write_seconds = now_s() - t_write;
    if (wrc) {
        fprintf(stderr,
                "Gauge transformation writing to file %s"
                "failed for configuration %s .\n",
                gt_filename, config_filename);
    } else {
        printf("G for config %s written to file %s.\n", config_filename, gt_filename);
    }
    gfc_job_destroy(job);
    gfc_ctx_destroy(ctx);

    if (env_long("GFC_METRICS", 0)) {
        const double upd = (double)sweeps * (double)(vol / 2);
        printf("{\"sweeps\": %d, \"sweeps_fp32\": %d, \"fix_seconds\": %.6f, "
               "\"site_updates_per_s\": %.6e, \"wall_seconds\": %.6f, "
               "\"residue_final\": %.6e, \"F\": %.15f, \"load_seconds\": %.6f, "
               "\"write_seconds\": %.6f}\n",
               sweeps, rep.sweeps_fp32, rep.seconds, rep.seconds > 0 ? upd / rep.seconds : 0.0,
               now_s() - t_start, rep.residue_final, F, load_seconds, write_seconds);
    }
    return EXIT_SUCCESS;
}
