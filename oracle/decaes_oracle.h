/*
 * decaes_oracle.h -- FP64 CPU restatement of DECAES.jl's voxelwise T2mapSEcorr / T2partSEcorr path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker and the timed CPU baseline. It is never
 * linked into, loaded by, or called from the product library (decaes.jl_amd/). Only tests/, the smoke()
 * entry and bench.py's CPU-baseline leg may load it.
 *
 * Every function cites the reference file:line it restates (paths relative to the DECAES.jl v0.6.1 tree).
 * Julia is 1-based and column-major; this restatement keeps column-major storage and converts indices.
 * It is compiled with -ffp-contract=off and calls fma() exactly where the reference uses muladd/@muladd.
 *
 * What pins each part (tests/test_oracle_*.py; fixtures in tests/golden/ written by numpy/scipy only):
 *   orc_epg_*            closed form at alpha = beta = 180; full-state complex EPG simulation; cross-form agreement
 *   orc_epg_dalpha       central finite differences
 *   orc_nnls, _tikh      scipy.optimize.nnls vectors; brute force over all active sets; KKT; Tikhonov gradient
 *   orc_brent_*, bracket sin / sin^2 roots and minima (the reference's test/optimization.jl cases)
 *   orc_hermite_minimize exact minimum of cubics
 *   orc_surrogate_search the reference's hand-checked bounding boxes; evaluation sequences derived by hand on a
 *                        quadratic loss (K = 32, 64; minimum left of the grid)
 *   orc_lcurve_corner    mock (mu, 1/mu) curve: log mu* = 0 within 1e-3 (the reference's test/nnls.jl case); the
 *                        first golden-section moves by hand; corners recovered across the bracket; the Ctol rule;
 *                        backtracking on/off identical
 *   orc_spline_*         scipy's FITPACK (splrep / splev, s = 0) bit for bit; exact on cubics
 *   orc_jlrange          Julia's documented range lengths and decimal endpoints; correctly rounded k / 1000
 *   chi2 legacy          the doubling search and the sampled root rebuilt from orc_nnls_tikh and scipy's spline
 *   orc_svdvals          numpy SVD of EPG bases; orc_gcv_dof: trace formula
 *   orc_lsqnonneg_*      chi2 / mdp targets met; gcv optimum is a local minimum; lcurve solution is the
 *                        Tikhonov solution at its mu; none is plain NNLS
 *   orc_t2part, erfinv   closed forms on hand-built distributions; scipy.special.erfinv
 *   orc_t2map            noiseless on-grid voxels recover alpha; threshold and NaN semantics
 * Parity unpinned against the reference itself (no reference-held vectors; Julia cannot run here): mu beyond
 * solver tolerance in orc_lsqnonneg_{chi2,mdp,gcv}, LAPACK's singular values beyond SVD equivalence, and hence the
 * end-to-end maps of orc_t2map against Julia DECAES.
 */
#ifndef DECAES_ORACLE_H
#define DECAES_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Regularization methods (src/T2mapSEcorr.jl:440-449). */
enum { ORC_REG_NONE = 0, ORC_REG_LCURVE = 1, ORC_REG_GCV = 2, ORC_REG_CHI2 = 3, ORC_REG_MDP = 4 };

/* Mirrors T2mapOptions (src/types.jl:19-100). NaN in SetFlipAngle means "nothing". */
typedef struct {
    int64_t nx, ny, nz;
    int32_t nTE;
    int32_t nT2;
    double TE;
    double T2Range[2];
    double T1;
    double Threshold;
    double MinRefAngle;
    int32_t nRefAngles;
    int32_t nRefAnglesMin;
    int32_t Reg;
    int32_t legacy;
    double Chi2Factor;
    double NoiseLevel;
    double RefConAngle;
    double SetFlipAngle;
} orc_t2map_opts;

/* Mirrors T2partOptions (src/types.jl:139-172). NaN in Sigmoid means "nothing". */
typedef struct {
    int64_t nx, ny, nz;
    int32_t nT2;
    int32_t _pad0;
    double T2Range[2];
    double SPWin[2];
    double MPWin[2];
    double Sigmoid;
} orc_t2part_opts;

/* ---- grids (src/utils.jl:7, src/types.jl:108-109) ---- */
void orc_logrange(double a, double b, int n, double *out);
void orc_linrange(double a, double b, int n, double *out);
double orc_sind(double x);
double orc_cosd(double x);
double orc_erfinv(double x);

/* ---- EPG (src/EPGdecaycurve.jl) ---- */
/* Constant-flip-angle train (EPGWork_ReIm_DualFlat_Split_Dynamic, :936-1028). */
void orc_epg_const(double *dc, int ETL, double alpha, double TE, double T2, double T1);
/* General refocusing control angle beta (EPGWork_ReIm_DualVector_Split_Dynamic, :722-818). */
void orc_epg_general(double *dc, int ETL, double alpha, double TE, double T2, double T1, double beta);
/* Readable complex rotation-matrix form (EPGWork_Basic_Cplx, :271-322), the reference's own test reference. */
void orc_epg_basic(double *dc, int ETL, double alpha, double TE, double T2, double T1, double beta);
/* Value and d/d(alpha) (alpha in degrees), forward mode as ForwardDiff does (:224-248).
 * beta >= 180 selects the constant-flip recursion, otherwise the general one (src/T2mapSEcorr.jl:616-620). */
void orc_epg_dalpha(double *dc, double *ddc, int ETL, double alpha, double TE, double T2, double T1, double beta);

/* ---- NNLS (src/NNLS.jl, src/lsqnonneg.jl:30-164) ---- */
/* Unregularized NNLSProblem solve with the reference warm start. A is m x n column-major.
 * Outputs x (n), returns rnorm; nsetp, mode and idx (1-based, n) optional (may be NULL). */
double orc_nnls(const double *A, const double *b, int m, int n, double *x, int *nsetp, int *mode, int *idx,
                double *w_out);
/* Tikhonov NNLSProblem solve: min ||Ax-b||^2 + mu^2||x||^2. Returns rnorm of the augmented system. */
double orc_nnls_tikh(const double *A, const double *b, int m, int n, double mu, double *x, int *nsetp, int *mode);

/* ---- regularization (src/lsqnonneg.jl) ---- */
/* Each returns x (n) and writes mu, chi2factor. */
void orc_lsqnonneg_none(const double *A, const double *b, int m, int n, double *x, double *mu, double *chi2);
void orc_lsqnonneg_chi2(const double *A, const double *b, int m, int n, double chi2_target, double *x, double *mu,
                        double *chi2);
void orc_lsqnonneg_mdp(const double *A, const double *b, int m, int n, double delta, double *x, double *mu,
                       double *chi2);
void orc_lsqnonneg_gcv(const double *A, const double *b, int m, int n, double *x, double *mu, double *chi2);
void orc_lsqnonneg_lcurve(const double *A, const double *b, int m, int n, double *x, double *mu, double *chi2);
/* Singular values, descending, min(m,n) of them (one-sided Jacobi; stands in for LAPACK dgesdd). */
void orc_svdvals(const double *A, int m, int n, double *s);
/* Householder bidiagonalization + Golub-Kahan tridiagonal bisection (the device kernel's route) */
void orc_svdvals_bidiag(const double *A, int m, int n, double *s);
/* GCV singular-value route: 0 = one-sided Jacobi on A (default), 1 = Gram-matrix eigenvalues (diagnostic),
 * 2 = orc_svdvals_bidiag */
void orc_set_svd_route(int route);
/* NNLS dot-product summation order: 0 = in order (default), 1 = reversed, 2 = two interleaved partial sums */
void orc_set_sum_order(int order);
/* mu = exp(log mu) moved by dir ulps (dir in {-1, 0, 1}; default 0) */
void orc_set_mu_ulp(int dir);
/* per-voxel output (volume-sized, indexed by linear voxel index) of the L-curve final-argmax margin: the relative
 * gap between the best and second-best cached curvature; NULL (default) disables it */
void orc_set_margin_output(double *margin);
double orc_gcv_dof(int m, int n, const double *gamma, int ngamma, double lambda);

/* ---- scalar optimizers (src/optimization.jl) with a C callback ---- */
typedef double (*orc_fun1)(double x, void *ctx);
void orc_brent_root(orc_fun1 f, void *ctx, double x0, double x1, double xatol, double xrtol, double ftol,
                    int maxiters, double *xout, double *fout);
void orc_bracket_root_monotonic(orc_fun1 f, void *ctx, double a, double delta, double dilate, double mono,
                                int maxiters, double out[4]);
void orc_brent_minimize(orc_fun1 f, void *ctx, double x1, double x2, double xrtol, double xatol, int maxiters,
                        double *xout, double *fout);
/* lcurve_corner on a caller-supplied curve f(logmu) -> (xi, eta) (src/lsqnonneg.jl:872-972). */
typedef void (*orc_fun2)(double x, void *ctx, double out[2]);
double orc_lcurve_corner(orc_fun2 f, void *ctx, double xlow, double xhigh, double xtol, double Ptol, double Ctol);
double orc_lcurve_corner_bt(orc_fun2 f, void *ctx, double xlow, double xhigh, double xtol, double Ptol, double Ctol,
                            int backtracking);
/* Largest L-curve cache sizes (function, point, state caches) seen since the previous call. */
void orc_lcurve_cache_stats(int out[3]);
/* CubicHermiteInterpolator + minimize (src/splines.jl:62-110). out = (x, u). */
void orc_hermite_minimize(double a, double b, double u0, double u1, double m0, double m1, double out[2]);

/* ---- flip-angle search on one voxel (src/T2mapSEcorr.jl:409-423, src/splines.jl:664-850,975-1051) ---- */
/* As: m x n x K grid bases, dAs: their d/dalpha. Returns alpha_opt; nevals optional. */
double orc_flip_search(const double *As, const double *dAs, const double *alphas, int m, int n, int K, int mineval,
                       const double *b, int *nevals);
/* The same searcher on any loss (u, du at 1-based grid index I); trace receives the evaluated indices in order. */
double orc_surrogate_search(const double *alphas, int K, int mineval,
                            void (*loss)(void *ctx, int I, double *u, double *du), void *ctx, int *trace,
                            int *nevals);
/* The same searcher with the legacy surrogate (CubicSplineSurrogate + spline_opt_legacy, src/splines.jl:494-500). */
double orc_surrogate_search_legacy(const double *alphas, int K, int mineval,
                                   void (*loss)(void *ctx, int I, double *u, double *du), void *ctx, int *trace,
                                   int *nevals);
/* minimal_bounding_box / sorted_corners of the 1-D searcher; seen is indexed 1..K */
void orc_minimal_bounding_box(const unsigned char *seen, const double *alphas, int K, double x, int *lo, int *hi);
void orc_sorted_corners(const double *alphas, int lo, int hi, double x, int out[2]);

/* ---- legacy path (legacy = true; src/splines.jl:312-315,414-446; src/lsqnonneg.jl:521-532,595-636) ---- */
/* Dierckx.Spline1D(x, y; k, s = 0): FITPACK's interpolating spline (knots of fpcurf's s = 0 case, Givens
 * rotations, back substitution). t receives m + k + 1 knots, c m coefficients; returns the knot count. */
int orc_spline_interp(const double *x, const double *y, int m, int k, double *t, double *c);
/* splev at one point (bc = "extrapolate") */
double orc_spline_eval(const double *t, int n, const double *c, int k, double x);
/* Julia's Float64 range start:step:stop: returns its length, writes min(length, cap) elements */
int64_t orc_jlrange(double start, double step, double stop, double *out, int64_t cap);
/* spline_opt_legacy: out = (x, y) of the least sampled value; spline_root_legacy: the sample nearest value */
void orc_spline_opt_legacy(const double *x, const double *y, int m, double out[2]);
double orc_spline_root_legacy(const double *x, const double *y, int m, double value);
/* lsqnonneg_chi2! with legacy = true; a search still short of the target after 24 doublings gives NaN */
void orc_lsqnonneg_chi2_legacy(const double *A, const double *b, int m, int n, double chi2_target, double *x,
                               double *mu, double *chi2);

/* ---- whole pipeline (src/T2mapSEcorr.jl:148-238,512-591) ---- */
/* image: (nx,ny,nz,nTE) column-major. alpha_in: B1 map (nx,ny,nz) or NULL. Outputs are caller-allocated and
 * must be pre-filled with NaN; voxels below threshold are left untouched. Optional outputs may be NULL.
 * decaybasisset (nTE x nT2 x K, or nTE x nT2 with SetFlipAngle) may be NULL.
 * Returns the number of processed voxels, or a negative error code. */
int64_t orc_t2map(const orc_t2map_opts *o, const double *image, const double *alpha_in, double *dist, double *gdn,
                  double *ggm, double *gva, double *fnr, double *snr, double *alpha, double *resnorm,
                  double *decaycurve, double *mu, double *chi2factor, double *decaybasis, double *decaybasisset,
                  int nthreads);
/* Pipeline over an explicit list of linear voxel indices (used for bounded CPU-baseline samples). */
int64_t orc_t2map_subset(const orc_t2map_opts *o, const double *image, const int64_t *vox, int64_t nvox,
                         double *dist, double *gdn, double *ggm, double *gva, double *fnr, double *snr,
                         double *alpha, double *mu, double *chi2factor, int nthreads);

/* src/T2partSEcorr.jl:40-164. dist: (nx,ny,nz,nT2). Outputs pre-filled with NaN. */
int orc_t2part(const orc_t2part_opts *o, const double *dist, double *sfr, double *sgm, double *mfr, double *mgm);

#ifdef __cplusplus
}
#endif
#endif
