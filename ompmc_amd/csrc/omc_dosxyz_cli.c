/*
 * omc_dosxyz -- drop-in replacement of the reference CLI
 * (ucodes/omc_dosxyz/omc_dosxyz.c main, :1072-1307) running the transport on
 * the GPU through the C ABI.
 *
 *   omc_dosxyz -i <stem> -o <name> [--verbose|--brief]
 *              [--fp64] [--history] [--seed N]
 *
 * Reads <stem>.inp with the reference's keys, the egsphant phantom, PEGS4 and
 * spectrum files, and writes <output folder><name>.3ddose.  Exit status 0, or
 * EXIT_FAILURE with a message on stderr, as the reference.
 */
#include <getopt.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "omc_gpu.h"
#include "omc_host.h"

static double now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static int need(const omc_inp *inp, const char *key, char *buf) {
    if (omc_inp_get(inp, key, buf, OMC_INP_LEN) != 1) {
        fprintf(stderr, "Can not find '%s' key on input file.\n", key);
        return 0;
    }
    return 1;
}

#define DIE(...) do { fprintf(stderr, __VA_ARGS__); fputc('\n', stderr); goto fail; } while (0)

int main(int argc, char **argv) {
    double t0 = now();
    static int verbose = 0;
    int fp64 = 0, history = 0;
    uint64_t seed = 0;
    int have_seed = 0;
    const char *in = NULL, *out = NULL;
    static struct option opts[] = {{"verbose", no_argument, &verbose, 1}, {"brief", no_argument, &verbose, 0},
                                   {"input", required_argument, 0, 'i'},  {"output", required_argument, 0, 'o'},
                                   {"fp64", no_argument, 0, 'd'},         {"history", no_argument, 0, 'h'},
                                   {"seed", required_argument, 0, 's'},   {0, 0, 0, 0}};
    int ch, idx;
    while ((ch = getopt_long(argc, argv, "i:o:", opts, &idx)) != -1) {
        if (ch == 'i') in = optarg;
        else if (ch == 'o') out = optarg;
        else if (ch == 'd') fp64 = 1;
        else if (ch == 'h') history = 1;
        else if (ch == 's') { seed = strtoull(optarg, NULL, 10); have_seed = 1; }
        else if (ch != 0) return EXIT_FAILURE;
    }
    if (!in || !out) {
        fprintf(stderr, "usage: %s -i <input stem> -o <output name> [--fp64] [--history] [--seed N]\n", argv[0]);
        return EXIT_FAILURE;
    }
    omc_inp inp;
    omc_phantom ph;
    memset(&ph, 0, sizeof ph);
    omc_ctx *ctx = NULL;
    double *sum = NULL, *sum2 = NULL, *dose = NULL, *unc = NULL;
    char buf[OMC_INP_LEN], path[OMC_INP_LEN], pegs[OMC_INP_LEN], form[OMC_INP_LEN], folder[OMC_INP_LEN], outdir[OMC_INP_LEN];
    if (omc_inp_parse(&inp, in)) DIE("%s", omc_host_last_error());
    if (!need(&inp, "phantom file", buf)) goto fail;
    omc_strip_spaces(path, buf, sizeof path);
    if (omc_read_egsphant(&ph, path)) DIE("%s", omc_host_last_error());
    if (!need(&inp, "pegs file", buf)) goto fail;
    omc_strip_spaces(pegs, buf, sizeof pegs);
    if (!need(&inp, "pgs4form file", buf)) goto fail;
    omc_strip_spaces(form, buf, sizeof form);
    if (!need(&inp, "data folder", buf)) goto fail;
    omc_strip_spaces(folder, buf, sizeof folder);
    const char *names[OMC_MXMED];
    for (int i = 0; i < ph.nmed; i++) names[i] = ph.med_names[i];
    if (omc_ctx_create(&ctx, ph.nmed, names, pegs, form, folder, 0)) DIE("%s", omc_last_error());

    omc_spectrum spec;
    if (omc_inp_get(&inp, "spectrum file", buf, sizeof buf) == 1) {
        omc_strip_spaces(path, buf, sizeof path);
        if (omc_spectrum_read(&spec, path)) DIE("%s", omc_host_last_error());
    } else {
        if (!need(&inp, "mono energy", buf)) goto fail;
        omc_spectrum_mono(&spec, atof(buf));
    }
    omc_dosxyz_source_desc src;
    memset(&src, 0, sizeof src);
    src.spec.spectrum = spec.spectrum;
    src.spec.energy = spec.energy;
    src.spec.cdfinv1 = spec.cdfinv1;
    src.spec.cdfinv2 = spec.cdfinv2;
    if (!need(&inp, "collimator bounds", buf)) goto fail;
    if (sscanf(buf, "%lf %lf %lf %lf", &src.collimator[0], &src.collimator[1], &src.collimator[2], &src.collimator[3]) != 4)
        DIE("malformed 'collimator bounds'");
    if (!need(&inp, "charge", buf)) goto fail;
    src.charge = atoi(buf);
    if (!need(&inp, "ssd", buf)) goto fail;
    src.ssd = atof(buf);
    if (!need(&inp, "global ecut", buf)) goto fail;
    double ecut = atof(buf);
    if (!need(&inp, "global pcut", buf)) goto fail;
    double pcut = atof(buf);
    omc_phantom_desc pdsc = {ph.nx, ph.ny, ph.nz, ph.xb, ph.yb, ph.zb, ph.med, ph.rho, ecut, pcut};
    int pid;
    if (omc_phantom_upload(ctx, &pdsc, &pid)) DIE("%s", omc_last_error());
    omc_run_opts ro;
    omc_run_opts_default(&ro);
    if (!need(&inp, "nsplit", buf)) goto fail;
    ro.nsplit = atoi(buf);
    if (ro.nsplit < 1) ro.nsplit = 1;
    if (!need(&inp, "rng seeds", buf)) goto fail;
    int ixx = 1802, jxx = 9373;
    sscanf(buf, "%d %d", &ixx, &jxx);
    ro.seed = have_seed ? seed : ((uint64_t)(unsigned)ixx << 32) | (unsigned)jxx;
    ro.precision = fp64 ? OMC_FP64 : OMC_FP32;
    ro.form = history ? OMC_FORM_HISTORY : OMC_FORM_EVENT;
    if (ro.nsplit > 1) printf("Photon splitting enabled, nsplit = %d\n", ro.nsplit);
    if (!need(&inp, "ncase", buf)) goto fail;
    long nhist = atol(buf);
    if (!need(&inp, "nbatch", buf)) goto fail;
    int nbatch = atoi(buf);
    if (nbatch < 2) DIE("nbatch must be at least 2");
    if (nhist / nbatch == 0) nhist = nbatch;
    long nper = nhist / nbatch;
    long nreg = (long)ph.nx * ph.ny * ph.nz + 1;
    sum = calloc(nreg, sizeof(double));
    sum2 = calloc(nreg, sizeof(double));
    dose = calloc(nreg, sizeof(double));
    unc = calloc(nreg, sizeof(double));
    printf("Total number of particle histories: %ld\n", nper * nbatch);
    printf("Number of statistical batches: %d\n", nbatch);
    printf("Histories per batch: %ld\n", nper);
    printf("Execution time up to this point : %8.2f seconds\n", now() - t0);
    omc_run_stats st;
    if (omc_run_dosxyz(ctx, pid, &src, nper, nbatch, &ro, sum, sum2, &st)) DIE("%s", omc_last_error());
    printf("Simulation finished\n");
    printf("Execution time up to this point : %8.2f seconds\n", now() - t0);
    printf("Transport: %.3f s on the GPU, %.4g histories/s\n", st.gpu_seconds, st.histories / st.gpu_seconds);
    if (verbose) {
        double etot = 0.0;
        for (long i = 1; i < nreg; i++) etot += sum[i];
        printf("Deposited energy in the phantom: %g MeV per batch\n", etot / nbatch);
    }
    omc_dosxyz_dose(ctx, pid, sum, sum2, nper, nbatch, dose, unc);
    if (!need(&inp, "output folder", buf)) goto fail;
    omc_strip_spaces(outdir, buf, sizeof outdir);
    char file[2 * OMC_INP_LEN];
    snprintf(file, sizeof file, "%s%s.3ddose", outdir, out);
    if (omc_write_3ddose_file(ctx, pid, file, dose, unc)) DIE("%s", omc_last_error());
    printf("Total execution time : %8.5f seconds\n", now() - t0);
    omc_ctx_destroy(ctx);
    omc_phantom_free(&ph);
    free(sum); free(sum2); free(dose); free(unc);
    return EXIT_SUCCESS;
fail:
    if (ctx) omc_ctx_destroy(ctx);
    omc_phantom_free(&ph);
    free(sum); free(sum2); free(dose); free(unc);
    return EXIT_FAILURE;
}
