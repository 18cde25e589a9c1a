/*
 * dmi_run: a C host program of the C ABI (include/decaes_amd.h), driving it the way a compiled caller does (a Julia
 * ccall shim, a C or C++ pipeline); the reference's own callers are Julia: T2mapSEcorr (src/T2mapSEcorr.jl:148-149)
 * and T2partSEcorr (src/T2partSEcorr.jl:40-71).
 *
 * It synthesizes a seeded multi-echo image (dmi_synth_mse), runs T2mapSEcorr and T2partSEcorr on host arrays
 * (dmi_t2map_f64, dmi_t2part_f64) into NaN-filled outputs, and writes the image, the distribution and the maps as
 * raw little-endian float64 files, column-major in the reference's shapes, into OUTDIR, with one JSON line of the
 * call's counters on stdout.
 *
 *   dmi_run OUTDIR X Y Z nTE TE nT2 REG nRefAngles seed [mask_fraction]
 *
 * REG: none | chi2 (Chi2Factor 1.02) | gcv | lcurve | mdp (NoiseLevel 1e-3). T2Range (10 ms, 2 s), T1 1 s, and the
 * reference's defaults for every other option; T2part with SPWin (10 ms, 40 ms), MPWin (40 ms, 2 s).
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "decaes_amd.h"

static int write_f64(const char *dir, const char *name, const double *p, size_t n) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s.f64", dir, name);
    FILE *f = fopen(path, "wb");
    if (!f) {
        fprintf(stderr, "dmi_run: cannot write %s\n", path);
        return 1;
    }
    size_t w = fwrite(p, sizeof(double), n, f);
    const int closed = fclose(f) == 0;
    if (w != n || !closed) {
        fprintf(stderr, "dmi_run: short write to %s\n", path);
        return 1;
    }
    return 0;
}

static double *nan_array(size_t n) {
    double *p = (double *)malloc(sizeof(double) * (n ? n : 1));
    if (p)
        for (size_t i = 0; i < n; ++i) p[i] = NAN;
    return p;
}

static double now_ms(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

static int fail(dmi_ctx *ctx, const char *what, int rc) {
    fprintf(stderr, "dmi_run: %s failed (%d): %s\n", what, rc, dmi_last_error(ctx));
    return 1;
}

int main(int argc, char **argv) {
    if (argc < 11) {
        fprintf(stderr, "usage: dmi_run OUTDIR X Y Z nTE TE nT2 REG nRefAngles seed [mask_fraction]\n"
                        "       REG: none | chi2 | gcv | lcurve | mdp\n");
        return 2;
    }
    const char *dir = argv[1];
    const int64_t X = atoll(argv[2]), Y = atoll(argv[3]), Z = atoll(argv[4]);
    const int nTE = atoi(argv[5]);
    const double TE = atof(argv[6]);
    const int nT2 = atoi(argv[7]);
    const char *reg = argv[8];
    const int K = atoi(argv[9]);
    const unsigned long long seed = strtoull(argv[10], NULL, 10);
    const double frac = argc > 11 ? atof(argv[11]) : 0.7;
    if (X < 1 || Y < 1 || Z < 1 || nTE < 4 || nT2 < 2 || K < 2 || !(TE > 0)) {
        fprintf(stderr, "dmi_run: invalid sizes\n");
        return 2;
    }

    dmi_t2map_opts o;
    memset(&o, 0, sizeof o);
    o.MatrixSize[0] = X; o.MatrixSize[1] = Y; o.MatrixSize[2] = Z;
    o.nTE = nTE; o.nT2 = nT2; o.TE = TE;
    o.T2Range[0] = 10e-3; o.T2Range[1] = 2.0;
    o.T1 = 1.0; o.Threshold = 0.0; o.MinRefAngle = 50.0;
    o.nRefAngles = K; o.nRefAnglesMin = K < 5 ? K : 5;
    o.Chi2Factor = NAN; o.NoiseLevel = NAN; o.RefConAngle = 180.0; o.SetFlipAngle = NAN;
    o.SaveRegParam = 1; o.SaveResidualNorm = 1;
    o.precision = DMI_PREC_FP64; o.schedule = DMI_SCHED_AUTO;
    if (!strcmp(reg, "none")) o.Reg = DMI_REG_NONE;
    else if (!strcmp(reg, "chi2")) { o.Reg = DMI_REG_CHI2; o.Chi2Factor = 1.02; }
    else if (!strcmp(reg, "gcv")) o.Reg = DMI_REG_GCV;
    else if (!strcmp(reg, "lcurve")) o.Reg = DMI_REG_LCURVE;
    else if (!strcmp(reg, "mdp")) { o.Reg = DMI_REG_MDP; o.NoiseLevel = 1e-3; }
    else {
        fprintf(stderr, "dmi_run: unknown REG %s\n", reg);
        return 2;
    }

    const size_t V = (size_t)(X * Y * Z);
    double *image = (double *)malloc(sizeof(double) * V * (size_t)nTE);
    if (!image) return 1;
    const int64_t nin = dmi_synth_mse(image, X, Y, Z, nTE, TE, o.T1, 3, frac, seed, 8);
    if (nin < 0) {
        fprintf(stderr, "dmi_run: dmi_synth_mse rejected the sizes\n");
        return 2;
    }

    dmi_ctx *ctx = NULL;
    int rc = dmi_create(&ctx, 0, 0);
    if (rc) return fail(NULL, "dmi_create", rc);

    const char *names[] = {"dist", "gdn", "ggm", "gva", "fnr", "snr", "alpha", "resnorm", "mu", "chi2factor",
                           "sfr", "sgm", "mfr", "mgm"};
    enum { NOUT = sizeof names / sizeof names[0] };
    double *buf[NOUT];
    for (int i = 0; i < NOUT; ++i) {
        buf[i] = nan_array(i == 0 ? V * (size_t)nT2 : V);
        if (!buf[i]) return 1;
    }
    dmi_t2map_out out;
    memset(&out, 0, sizeof out);
    out.dist = buf[0]; out.gdn = buf[1]; out.ggm = buf[2]; out.gva = buf[3]; out.fnr = buf[4]; out.snr = buf[5];
    out.alpha = buf[6]; out.resnorm = buf[7]; out.mu = buf[8]; out.chi2factor = buf[9];

    const double t0 = now_ms();
    if ((rc = dmi_t2map_f64(ctx, image, &o, NULL, &out))) return fail(ctx, "dmi_t2map_f64", rc);
    const double t1 = now_ms();
    dmi_stats st;
    if ((rc = dmi_last_stats(ctx, &st))) return fail(ctx, "dmi_last_stats", rc);

    dmi_t2part_opts po;
    memset(&po, 0, sizeof po);
    po.MatrixSize[0] = X; po.MatrixSize[1] = Y; po.MatrixSize[2] = Z;
    po.nT2 = nT2;
    po.T2Range[0] = o.T2Range[0]; po.T2Range[1] = o.T2Range[1];
    po.SPWin[0] = 10e-3; po.SPWin[1] = 40e-3;
    po.MPWin[0] = 40e-3; po.MPWin[1] = 2.0;
    po.Sigmoid = NAN;
    if ((rc = dmi_t2part_f64(ctx, buf[0], &po, buf[10], buf[11], buf[12], buf[13])))
        return fail(ctx, "dmi_t2part_f64", rc);
    const double t2 = now_ms();

    int werr = write_f64(dir, "image", image, V * (size_t)nTE);
    for (int i = 0; i < NOUT; ++i) werr |= write_f64(dir, names[i], buf[i], i == 0 ? V * (size_t)nT2 : V);
    double mwf = 0.0;
    int64_t nmwf = 0;
    for (size_t v = 0; v < V; ++v)
        if (!isnan(buf[10][v])) { mwf += buf[10][v]; ++nmwf; }
    printf("{\"voxels_in_mask\": %lld, \"voxels\": %lld, \"nnls_solves\": %lld, \"kernel_ms\": %.3f, "
           "\"t2map_wall_ms\": %.3f, \"t2part_wall_ms\": %.3f, \"mean_sfr\": %.17g, \"lds_rows\": %d, "
           "\"workgroups_per_cu\": %d}\n",
           (long long)nin, (long long)st.voxels, (long long)st.nnls_solves, st.kernel_ms, t1 - t0, t2 - t1,
           nmwf ? mwf / (double)nmwf : NAN, st.lds_rows, st.workgroups_per_cu);
    for (int i = 0; i < NOUT; ++i) free(buf[i]);
    free(image);
    dmi_destroy(ctx);
    return werr;
}
