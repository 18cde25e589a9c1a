// C ABI of the gfx950 T2mapSEcorr / T2partSEcorr path (declared in include/decaes_amd.h).
//
// The host side mirrors what the reference does around its voxel loop: option validation with the same
// @assert messages (src/types.jl:19-100,139-172), the T2 and flip-angle grids (src/types.jl:108-109,
// src/utils.jl:7), the threshold filter (src/T2mapSEcorr.jl:177), output conventions (NaN-filled maps,
// voxels below threshold untouched) and T2part window ranges (src/T2partSEcorr.jl:143-164). All voxel work
// runs on the device; there is no CPU fallback path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "decaes_amd.h"
#include "dmi_internal.h"

namespace {

thread_local std::string g_create_error;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        size_t want = std::max<size_t>(n, 256);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct GridKey {
    int m = -1, n = -1, K = -1, general = -1;
    double TE = 0, T1 = 0, beta = 0, T2lo = 0, T2hi = 0, minref = 0, setflip = 0;
    bool operator==(const GridKey &o) const {
        auto same = [](double a, double b) { return a == b || (std::isnan(a) && std::isnan(b)); };
        return m == o.m && n == o.n && K == o.K && general == o.general && same(TE, o.TE) && same(T1, o.T1) &&
               same(beta, o.beta) && same(T2lo, o.T2lo) && same(T2hi, o.T2hi) && same(minref, o.minref) &&
               same(setflip, o.setflip);
    }
};

}  // namespace

struct dmi_ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;
    std::mutex mu;
    dmi_stats stats{};
    // grids
    GridKey key;
    DevBuf E2, logT2, alphas, As, dAs, Afixed, AsT, dAsT, AfixedT;
    // per-call buffers
    DevBuf vox, counter, scratch, counters, status;
    // host-array staging
    DevBuf img, ain, out[12], t2p[5];
    DevBuf weights, tlog;
};

namespace {

int fail(dmi_ctx *c, int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    else g_create_error = buf;
    return code;
}

#define HIPCHECK(c, expr)                                                                                          \
    do {                                                                                                           \
        hipError_t _e = (expr);                                                                                    \
        if (_e != hipSuccess) return fail((c), DMI_EHIP, "%s failed: %s", #expr, hipGetErrorString(_e));         \
    } while (0)

// range(a, b; length = n) as Julia evaluates it (TwicePrecision, correctly rounded); endpoints exact.
void linrange(double a, double b, int n, double *out) {
    if (n == 1) { out[0] = a; return; }
    for (int j = 0; j < n; ++j)
        out[j] = double(((long double)(n - 1 - j) * a + (long double)j * b) / (long double)(n - 1));
    out[0] = a;
    out[n - 1] = b;
}
// logrange (src/utils.jl:7)
void logrange(double a, double b, int n, double *out) {
    linrange(std::log(a), std::log(b), n, out);
    for (int j = 0; j < n; ++j) out[j] = std::exp(out[j]);
    out[0] = a;
    out[n - 1] = b;
}

// Inverse error function: Winitzki's (2008) closed-form start refined by Halley steps, on erf near zero and on
// erfc against 1 - |y| (exact for |y| >= 1/2) in the tails.
double erfinv(double y) {
    if (y <= -1.0) return y == -1.0 ? -INFINITY : NAN;
    if (y >= 1.0) return y == 1.0 ? INFINITY : NAN;
    if (y == 0.0) return y;
    const double a = 0.147, pi = 3.141592653589793;
    const double ay = std::fabs(y), q = 1.0 - ay;
    double l = std::log1p(-y * y), t = 2.0 / (pi * a) + 0.5 * l;
    double x = std::sqrt(std::sqrt(t * t - l / a) - t);
    for (int it = 0; it < 5; ++it) {
        double f = ay < 0.5 ? std::erf(x) - ay : q - std::erfc(x);
        double fp = 2.0 / std::sqrt(pi) * std::exp(-x * x);
        double s = f / fp;
        x -= s / (1.0 + x * s);
    }
    return std::copysign(x, y);
}

// T2mapOptions @assert checks, in the reference's order and wording (src/types.jl:28-84).
int validate_t2map(dmi_ctx *c, const dmi_t2map_opts *o) {
    if (!o) return fail(c, DMI_EINVAL, "opts is NULL");
    if (o->legacy) return fail(c, DMI_EUNSUPPORTED, "legacy = true is not supported by the MI355X backend");
    if (!(o->MatrixSize[0] >= 1 && o->MatrixSize[1] >= 1 && o->MatrixSize[2] >= 1))
        return fail(c, DMI_EINVAL, "MatrixSize must be a tuple of 3 positive integers, but MatrixSize = (%lld, %lld, %lld).",
                    (long long)o->MatrixSize[0], (long long)o->MatrixSize[1], (long long)o->MatrixSize[2]);
    if (!(o->nTE >= 4)) return fail(c, DMI_EINVAL, "At least four echoes are required for T2 mapping, but nTE = %d.", o->nTE);
    if (!(o->TE > 0.0)) return fail(c, DMI_EINVAL, "Echo spacing must be positive, but TE = %g.", o->TE);
    if (!(o->nT2 >= 2)) return fail(c, DMI_EINVAL, "At least two T2 components are required for T2 mapping, but nT2 = %d.", o->nT2);
    if (!(0.0 < o->T2Range[0] && o->T2Range[0] < o->T2Range[1]))
        return fail(c, DMI_EINVAL, "T2Range must a sorted 2-tuple of positive values, but T2Range = (%g, %g).", o->T2Range[0], o->T2Range[1]);
    if (!(o->T1 > 0.0)) return fail(c, DMI_EINVAL, "T1 must be positive, but T1 = %g.", o->T1);
    if (!(o->Threshold >= 0.0 || o->Threshold == -INFINITY))
        return fail(c, DMI_EINVAL, "First echo signal threshold must be non-negative or -Inf to force processing of every voxel, but Threshold = %g.", o->Threshold);
    if (!(0.0 <= o->MinRefAngle && o->MinRefAngle <= 180.0))
        return fail(c, DMI_EINVAL, "Minimum refocusing angle must be in the range [0, 180], but MinRefAngle = %g.", o->MinRefAngle);
    if (!(o->nRefAngles >= 2))
        return fail(c, DMI_EINVAL, "Maximum number of angles to check during flip angle optimization must be at least 2, but nRefAngles = %d.", o->nRefAngles);
    if (!(2 <= o->nRefAnglesMin && o->nRefAnglesMin <= o->nRefAngles))
        return fail(c, DMI_EINVAL, "Minimum number of angles to check during flip angle optimization must be in the range [2, nRefAngles], but nRefAngles = %d and nRefAnglesMin = %d.", o->nRefAngles, o->nRefAnglesMin);
    if (o->Reg < DMI_REG_NONE || o->Reg > DMI_REG_MDP) return fail(c, DMI_EINVAL, "Reg must be one of none, lcurve, gcv, chi2, mdp");
    if (o->Reg == DMI_REG_CHI2 && !(o->Chi2Factor > 1.0))
        return fail(c, DMI_EINVAL, "Chi2Factor must be greater than 1.0, but Chi2Factor = %g.", o->Chi2Factor);
    if (o->Reg == DMI_REG_MDP && !(o->NoiseLevel > 0.0))
        return fail(c, DMI_EINVAL, "Noise level must be positive, but NoiseLevel = %g.", o->NoiseLevel);
    if (!(0.0 <= o->RefConAngle && o->RefConAngle <= 180.0))
        return fail(c, DMI_EINVAL, "Refocusing control angle must be in the range [0, 180], but RefConAngle = %g.", o->RefConAngle);
    if (!std::isnan(o->SetFlipAngle) && !(0.0 <= o->SetFlipAngle && o->SetFlipAngle <= 180.0))
        return fail(c, DMI_EINVAL, "Fixed flip angle must be in the range [0, 180], but SetFlipAngle = %g.", o->SetFlipAngle);
    if (o->nT2 > dmi::kMaxT2) return fail(c, DMI_EUNSUPPORTED, "nT2 = %d exceeds the device limit of %d", o->nT2, dmi::kMaxT2);
    if (o->nTE > dmi::kMaxEcho) return fail(c, DMI_EUNSUPPORTED, "nTE = %d exceeds the device limit of %d", o->nTE, dmi::kMaxEcho);
    if (std::isnan(o->SetFlipAngle) && o->nRefAngles > dmi::kMaxAngles)
        return fail(c, DMI_EUNSUPPORTED, "nRefAngles = %d exceeds the device limit of %d", o->nRefAngles, dmi::kMaxAngles);
    if (o->precision != DMI_PREC_FP64 && o->precision != DMI_PREC_MIXED) return fail(c, DMI_EINVAL, "unknown precision %d", o->precision);
    return DMI_OK;
}

// Grid bases for the search (or the fixed basis), cached on the context by their defining options.
int ensure_grids(dmi_ctx *c, const dmi_t2map_opts *o, bool search) {
    GridKey k;
    k.m = o->nTE; k.n = o->nT2; k.K = search ? o->nRefAngles : 1;
    k.general = o->RefConAngle != 180.0;
    k.TE = o->TE; k.T1 = o->T1; k.beta = o->RefConAngle; k.T2lo = o->T2Range[0]; k.T2hi = o->T2Range[1];
    k.minref = search ? o->MinRefAngle : NAN;
    k.setflip = search ? NAN : o->SetFlipAngle;
    if (k == c->key) return DMI_OK;
    const int m = k.m, n = k.n, K = k.K;
    std::vector<double> T2(n), E2(n), lg(n), al(K);
    logrange(o->T2Range[0], o->T2Range[1], n, T2.data());
    for (int j = 0; j < n; ++j) { E2[j] = std::exp(-(o->TE / 2) / T2[j]); lg[j] = std::log(T2[j]); }
    if (search) linrange(o->MinRefAngle, 180.0, K, al.data());
    else al[0] = o->SetFlipAngle;
    HIPCHECK(c, c->E2.ensure(sizeof(double) * n));
    HIPCHECK(c, c->logT2.ensure(sizeof(double) * n));
    HIPCHECK(c, c->alphas.ensure(sizeof(double) * K));
    HIPCHECK(c, hipMemcpyAsync(c->E2.p, E2.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->logT2.p, lg.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipMemcpyAsync(c->alphas.p, al.data(), sizeof(double) * K, hipMemcpyHostToDevice, c->stream));
    dmi::GridParams g{};
    g.m = m; g.n = n; g.K = K; g.general = k.general; g.beta = o->RefConAngle;
    g.E1 = std::exp(-(o->TE / 2) / o->T1);
    g.E2 = c->E2.as<double>();
    g.alphas = c->alphas.as<double>();
    size_t bytes = sizeof(double) * (size_t)m * n * K;
    if (search) {
        HIPCHECK(c, c->As.ensure(bytes));
        HIPCHECK(c, c->dAs.ensure(bytes));
        HIPCHECK(c, c->AsT.ensure(bytes));
        HIPCHECK(c, c->dAsT.ensure(bytes));
        g.As = c->As.as<double>();
        g.dAs = c->dAs.as<double>();
        g.AsT = c->AsT.as<double>();
    } else {
        HIPCHECK(c, c->Afixed.ensure(bytes));
        HIPCHECK(c, c->AfixedT.ensure(bytes));
        g.As = c->Afixed.as<double>();
        g.dAs = nullptr;
        g.AsT = c->AfixedT.as<double>();
    }
    HIPCHECK(c, dmi::launch_grid_basis(g, c->stream));
    if (search) HIPCHECK(c, dmi::launch_transpose(c->dAs.as<double>(), c->dAsT.as<double>(), m, n, K, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    c->key = k;
    return DMI_OK;
}

int run_t2map(dmi_ctx *c, const double *d_image, const dmi_t2map_opts *o, const double *d_alpha_in,
              const dmi_t2map_out *d) {
    const int m = o->nTE, n = o->nT2;
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    if (!d || !d->dist || !d->gdn || !d->ggm || !d->gva || !d->fnr || !d->snr || !d->alpha)
        return fail(c, DMI_EINVAL, "dist, gdn, ggm, gva, fnr, snr and alpha outputs are required");
    if (o->SaveRegParam && (!d->mu || !d->chi2factor)) return fail(c, DMI_EINVAL, "SaveRegParam requires mu and chi2factor outputs");
    if (o->SaveResidualNorm && !d->resnorm) return fail(c, DMI_EINVAL, "SaveResidualNorm requires the resnorm output");
    if (o->SaveDecayCurve && !d->decaycurve) return fail(c, DMI_EINVAL, "SaveDecayCurve requires the decaycurve output");
    if (o->SaveNNLSBasis && !d->decaybasis) return fail(c, DMI_EINVAL, "SaveNNLSBasis requires the decaybasis output");
    const bool fixed = !std::isnan(o->SetFlipAngle);
    const bool b1 = d_alpha_in != nullptr;
    HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    int rc = ensure_grids(c, o, !fixed);
    if (rc) return rc;
    if (fixed && o->SaveNNLSBasis)
        HIPCHECK(c, hipMemcpyAsync(d->decaybasis, c->Afixed.p, sizeof(double) * m * n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(c, c->vox.ensure(sizeof(int64_t) * std::max<int64_t>(V, 1)));
    HIPCHECK(c, c->counter.ensure(sizeof(unsigned) * 4));
    HIPCHECK(c, c->counters.ensure(sizeof(unsigned long long) * 4));
    HIPCHECK(c, c->status.ensure(sizeof(int) * 4));
    HIPCHECK(c, hipMemsetAsync(c->counter.p, 0, sizeof(unsigned) * 4, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->counters.p, 0, sizeof(unsigned long long) * 4, c->stream));
    HIPCHECK(c, hipMemsetAsync(c->status.p, 0, sizeof(int) * 4, c->stream));
    HIPCHECK(c, dmi::launch_select(d_image, V, o->Threshold, c->vox.as<int64_t>(), c->counter.as<unsigned>(), c->stream));
    unsigned nvox = 0;
    HIPCHECK(c, hipMemcpyAsync(&nvox, c->counter.p, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMemsetAsync(c->counter.p, 0, sizeof(unsigned), c->stream));
    c->stats = dmi_stats{};
    c->stats.voxels = nvox;
    if (nvox == 0) {
        HIPCHECK(c, hipEventRecord(c->ev[3], c->stream));
        HIPCHECK(c, hipEventSynchronize(c->ev[3]));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[3]);
        c->stats.total_ms = ms;
        return DMI_OK;
    }
    dmi::T2mapParams p{};
    p.m = m; p.n = n; p.K = fixed ? 1 : o->nRefAngles; p.mineval = o->nRefAnglesMin; p.reg = o->Reg;
    p.mode = b1 ? dmi::kModeB1 : (fixed ? dmi::kModeFixed : dmi::kModeSearch);
    p.general = o->RefConAngle != 180.0;
    p.lda = dmi::t2map_lda(m, n);
    p.mw = m + n;
    p.save_resnorm = o->SaveResidualNorm; p.save_decaycurve = o->SaveDecayCurve;
    p.save_regparam = o->SaveRegParam; p.save_basis = o->SaveNNLSBasis;
    p.V = V;
    p.nvox = nvox;
    p.E1 = std::exp(-(o->TE / 2) / o->T1);
    p.beta = o->RefConAngle;
    p.chi2factor = o->Chi2Factor;
    p.noiselevel = o->NoiseLevel;
    p.setflip = o->SetFlipAngle;
    p.image = d_image;
    p.vox = c->vox.as<int64_t>();
    p.counter = c->counter.as<unsigned>();
    p.E2 = c->E2.as<double>();
    p.logT2 = c->logT2.as<double>();
    p.alphas = c->alphas.as<double>();
    p.As = fixed ? nullptr : c->As.as<double>();
    p.AsT = fixed ? nullptr : c->AsT.as<double>();
    p.dAsT = fixed ? nullptr : c->dAsT.as<double>();
    p.Afixed = fixed ? c->Afixed.as<double>() : nullptr;
    p.AfixedT = fixed ? c->AfixedT.as<double>() : nullptr;
    p.alpha_in = d_alpha_in;
    p.dist = d->dist; p.gdn = d->gdn; p.ggm = d->ggm; p.gva = d->gva; p.fnr = d->fnr; p.snr = d->snr; p.alpha = d->alpha;
    p.resnorm = o->SaveResidualNorm ? d->resnorm : nullptr;
    p.decaycurve = o->SaveDecayCurve ? d->decaycurve : nullptr;
    p.mu = o->SaveRegParam ? d->mu : nullptr;
    p.chi2f = o->SaveRegParam ? d->chi2factor : nullptr;
    p.decaybasis = (o->SaveNNLSBasis && !fixed) ? d->decaybasis : nullptr;
    p.counters = c->counters.as<unsigned long long>();
    p.status = c->status.as<int>();
    size_t lds = dmi::t2map_lds_bytes(m, n, p.lda, p.mw);
    int per_cu = 0;
    HIPCHECK(c, dmi::t2map_occupancy(p.reg, &per_cu, lds));
    if (per_cu < 1) return fail(c, DMI_EUNSUPPORTED, "voxel kernel does not fit on a CU (LDS %zu bytes)", lds);
    int grid = (int)std::min<int64_t>((int64_t)c->num_cu * per_cu, (int64_t)nvox);
    p.slot_stride = dmi::t2map_slot_doubles(m, n);
    HIPCHECK(c, c->scratch.ensure(sizeof(double) * (size_t)p.slot_stride * grid));
    p.scratch = c->scratch.as<double>();
    HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHECK(c, dmi::launch_t2map(p, grid, lds, c->stream));
    HIPCHECK(c, hipEventRecord(c->ev[2], c->stream));
    unsigned long long cnt[2] = {0, 0};
    int status = 0;
    HIPCHECK(c, hipMemcpyAsync(cnt, c->counters.p, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(&status, c->status.p, sizeof status, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipEventRecord(c->ev[3], c->stream));
    HIPCHECK(c, hipEventSynchronize(c->ev[3]));
    float kms = 0, tms = 0;
    HIPCHECK(c, hipEventElapsedTime(&kms, c->ev[1], c->ev[2]));
    HIPCHECK(c, hipEventElapsedTime(&tms, c->ev[0], c->ev[3]));
    c->stats.nnls_solves = (int64_t)cnt[1];
    c->stats.fp64_flops = (int64_t)cnt[0];
    c->stats.kernel_ms = kms;
    c->stats.total_ms = tms;
    if (status & dmi::kStatusLcurveOverflow)
        return fail(c, DMI_ECAPACITY, "L-curve cache exceeded %d entries for at least one voxel", dmi::kLcurveCap);
    return DMI_OK;
}

int t2part_windows(dmi_ctx *c, const dmi_t2part_opts *o, std::vector<double> &lg, std::vector<double> &wts,
                   dmi::T2partParams &p) {
    if (!o) return fail(c, DMI_EINVAL, "opts is NULL");
    if (o->legacy) return fail(c, DMI_EUNSUPPORTED, "legacy = true is not supported by the MI355X backend");
    if (!(o->MatrixSize[0] >= 1 && o->MatrixSize[1] >= 1 && o->MatrixSize[2] >= 1)) return fail(c, DMI_EINVAL, "all(MatrixSize .>= 1)");
    if (!(o->nT2 >= 2)) return fail(c, DMI_EINVAL, "nT2 >= 2");
    if (!(0.0 < o->T2Range[0] && o->T2Range[0] < o->T2Range[1])) return fail(c, DMI_EINVAL, "0.0 < T2Range[1] < T2Range[2]");
    if (!(o->SPWin[0] < o->SPWin[1])) return fail(c, DMI_EINVAL, "SPWin[1] < SPWin[2]");
    if (!(o->MPWin[0] < o->MPWin[1])) return fail(c, DMI_EINVAL, "MPWin[1] < MPWin[2]");
    if (!(std::isnan(o->Sigmoid) || o->Sigmoid > 0)) return fail(c, DMI_EINVAL, "Sigmoid === nothing || Sigmoid > 0");
    const int n = o->nT2;
    std::vector<double> T2(n);
    logrange(o->T2Range[0], o->T2Range[1], n, T2.data());
    lg.resize(n);
    for (int j = 0; j < n; ++j) lg[j] = std::log(T2[j]);
    auto ffirst = [&](double lo) { for (int j = 0; j < n; ++j) if (T2[j] >= lo) return j; return -1; };
    auto flast = [&](double hi) { for (int j = n - 1; j >= 0; --j) if (T2[j] <= hi) return j; return -1; };
    p.sp0 = ffirst(o->SPWin[0]); p.sp1 = flast(o->SPWin[1]);
    p.mp0 = ffirst(o->MPWin[0]); p.mp1 = flast(o->MPWin[1]);
    if (p.sp0 < 0 || p.sp1 < 0 || p.mp0 < 0 || p.mp1 < 0)
        return fail(c, DMI_EINVAL, "T2 window outside T2Range: findfirst/findlast returned nothing");
    wts.clear();
    if (!std::isnan(o->Sigmoid)) {
        // sigmoid_weights (src/T2partSEcorr.jl:155-164)
        const double k = 0.1, sigma = std::fabs(o->Sigmoid / (std::sqrt(2.0) * erfinv(2 * k - 1)));
        wts.resize(n);
        for (int j = 0; j < n; ++j) {
            double v = std::erfc(((T2[j] - o->SPWin[1]) / sigma) / std::sqrt(2.0)) / 2;
            wts[j] = v <= 2.220446049250313e-16 ? 0.0 : v;
        }
    }
    p.n = n;
    p.V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    return DMI_OK;
}

int run_t2part(dmi_ctx *c, const double *d_dist, const dmi_t2part_opts *o, double *sfr, double *sgm, double *mfr,
               double *mgm) {
    std::vector<double> lg, wts;
    dmi::T2partParams p{};
    int rc = t2part_windows(c, o, lg, wts, p);
    if (rc) return rc;
    HIPCHECK(c, c->tlog.ensure(sizeof(double) * p.n));
    HIPCHECK(c, hipMemcpyAsync(c->tlog.p, lg.data(), sizeof(double) * p.n, hipMemcpyHostToDevice, c->stream));
    p.logT2 = c->tlog.as<double>();
    p.weights = nullptr;
    if (!wts.empty()) {
        HIPCHECK(c, c->weights.ensure(sizeof(double) * p.n));
        HIPCHECK(c, hipMemcpyAsync(c->weights.p, wts.data(), sizeof(double) * p.n, hipMemcpyHostToDevice, c->stream));
        p.weights = c->weights.as<double>();
    }
    p.dist = d_dist; p.sfr = sfr; p.sgm = sgm; p.mfr = mfr; p.mgm = mgm;
    HIPCHECK(c, dmi::launch_t2part(p, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return DMI_OK;
}

}  // namespace

extern "C" {

int dmi_create(dmi_ctx **out, int device, uint32_t flags) {
    (void)flags;
    if (!out) return fail(nullptr, DMI_EINVAL, "out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return fail(nullptr, DMI_ENODEV, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return fail(nullptr, DMI_ENODEV, "device %d out of range (%d devices)", device, ndev);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return fail(nullptr, DMI_EHIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, DMI_ENODEV, "device %d is %s; this library is built for gfx950 only", device, prop.gcnArchName);
    auto *c = new dmi_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    if ((e = hipSetDevice(device)) != hipSuccess || (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete c;
        return fail(nullptr, DMI_EHIP, "stream creation failed: %s", hipGetErrorString(e));
    }
    for (auto &ev : c->ev) (void)hipEventCreate(&ev);
    *out = c;
    return DMI_OK;
}

int dmi_destroy(dmi_ctx *c) {
    if (!c) return DMI_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    DevBuf *bufs[] = {&c->E2, &c->logT2, &c->alphas, &c->As, &c->dAs, &c->Afixed, &c->AsT, &c->dAsT, &c->AfixedT,
                      &c->vox, &c->counter, &c->scratch,
                      &c->counters, &c->status, &c->img, &c->ain, &c->weights, &c->tlog};
    for (auto *b : bufs) b->release();
    for (auto &b : c->out) b.release();
    for (auto &b : c->t2p) b.release();
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DMI_OK;
}

const char *dmi_last_error(const dmi_ctx *c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int dmi_last_stats(const dmi_ctx *c, dmi_stats *out) {
    if (!c || !out) return DMI_EINVAL;
    *out = c->stats;
    return DMI_OK;
}

int dmi_t2map_f64_device(dmi_ctx *c, const double *d_image, const dmi_t2map_opts *o, const double *d_alpha_in,
                         const dmi_t2map_out *d_out) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHECK(c, hipSetDevice(c->device));
    int rc = validate_t2map(c, o);
    if (rc) return rc;
    if (!d_image) return fail(c, DMI_EINVAL, "image is NULL");
    return run_t2map(c, d_image, o, d_alpha_in, d_out);
}

int dmi_t2map_f64(dmi_ctx *c, const double *image, const dmi_t2map_opts *o, const double *alpha_in,
                  const dmi_t2map_out *out) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHECK(c, hipSetDevice(c->device));
    int rc = validate_t2map(c, o);
    if (rc) return rc;
    if (!image || !out) return fail(c, DMI_EINVAL, "image and out are required");
    const int m = o->nTE, n = o->nT2;
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    const bool fixed = !std::isnan(o->SetFlipAngle);
    HIPCHECK(c, c->img.ensure(sizeof(double) * (size_t)V * m));
    HIPCHECK(c, hipMemcpyAsync(c->img.p, image, sizeof(double) * (size_t)V * m, hipMemcpyHostToDevice, c->stream));
    if (alpha_in) {
        HIPCHECK(c, c->ain.ensure(sizeof(double) * (size_t)V));
        HIPCHECK(c, hipMemcpyAsync(c->ain.p, alpha_in, sizeof(double) * (size_t)V, hipMemcpyHostToDevice, c->stream));
    }
    // Stage every output on the device with the caller's current contents, so voxels at or below the
    // threshold keep whatever the caller put there (NaN, or the B1 map in alpha).
    double *hp[12] = {out->dist, out->gdn, out->ggm, out->gva, out->fnr, out->snr, out->alpha,
                      o->SaveResidualNorm ? out->resnorm : nullptr, o->SaveDecayCurve ? out->decaycurve : nullptr,
                      o->SaveRegParam ? out->mu : nullptr, o->SaveRegParam ? out->chi2factor : nullptr,
                      o->SaveNNLSBasis ? out->decaybasis : nullptr};
    size_t cnt[12] = {(size_t)V * n, (size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)V,
                      (size_t)V * m, (size_t)V, (size_t)V, fixed ? (size_t)m * n : (size_t)V * m * n};
    double *dp[12] = {nullptr};
    for (int i = 0; i < 12; ++i) {
        if (!hp[i]) continue;
        HIPCHECK(c, c->out[i].ensure(sizeof(double) * cnt[i]));
        dp[i] = c->out[i].as<double>();
        HIPCHECK(c, hipMemcpyAsync(dp[i], hp[i], sizeof(double) * cnt[i], hipMemcpyHostToDevice, c->stream));
    }
    dmi_t2map_out d{dp[0], dp[1], dp[2], dp[3], dp[4], dp[5], dp[6], dp[7], dp[8], dp[9], dp[10], dp[11]};
    rc = run_t2map(c, c->img.as<double>(), o, alpha_in ? c->ain.as<double>() : nullptr, &d);
    if (rc && rc != DMI_ECAPACITY) return rc;
    for (int i = 0; i < 12; ++i)
        if (hp[i]) HIPCHECK(c, hipMemcpyAsync(hp[i], dp[i], sizeof(double) * cnt[i], hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return rc;
}

int dmi_epg_basis_set(dmi_ctx *c, const dmi_t2map_opts *o, double *out) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHECK(c, hipSetDevice(c->device));
    int rc = validate_t2map(c, o);
    if (rc) return rc;
    if (!out) return fail(c, DMI_EINVAL, "out is NULL");
    const bool search = std::isnan(o->SetFlipAngle);
    rc = ensure_grids(c, o, search);
    if (rc) return rc;
    const int m = o->nTE, n = o->nT2, K = search ? o->nRefAngles : 1;
    std::vector<double> rm((size_t)m * n * K);
    HIPCHECK(c, hipMemcpy(rm.data(), search ? c->As.p : c->Afixed.p, sizeof(double) * rm.size(), hipMemcpyDeviceToHost));
    // device layout K x m x n (row-major per angle) -> Julia (nTE, nT2, K) column-major
    for (int k = 0; k < K; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < m; ++i) out[i + (size_t)m * (j + (size_t)n * k)] = rm[((size_t)k * m + i) * n + j];
    return DMI_OK;
}

int dmi_t2part_f64_device(dmi_ctx *c, const double *d_dist, const dmi_t2part_opts *o, double *sfr, double *sgm,
                          double *mfr, double *mgm) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHECK(c, hipSetDevice(c->device));
    if (!d_dist || !sfr || !sgm || !mfr || !mgm) return fail(c, DMI_EINVAL, "dist and all four outputs are required");
    return run_t2part(c, d_dist, o, sfr, sgm, mfr, mgm);
}

int dmi_t2part_f64(dmi_ctx *c, const double *dist, const dmi_t2part_opts *o, double *sfr, double *sgm, double *mfr,
                   double *mgm) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    HIPCHECK(c, hipSetDevice(c->device));
    if (!o) return fail(c, DMI_EINVAL, "opts is NULL");
    if (!dist || !sfr || !sgm || !mfr || !mgm) return fail(c, DMI_EINVAL, "dist and all four outputs are required");
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    HIPCHECK(c, c->t2p[0].ensure(sizeof(double) * (size_t)V * std::max(o->nT2, 1)));
    HIPCHECK(c, hipMemcpyAsync(c->t2p[0].p, dist, sizeof(double) * (size_t)V * o->nT2, hipMemcpyHostToDevice, c->stream));
    double *hp[4] = {sfr, sgm, mfr, mgm};
    double *dp[4];
    for (int i = 0; i < 4; ++i) {
        HIPCHECK(c, c->t2p[1 + i].ensure(sizeof(double) * (size_t)V));
        dp[i] = c->t2p[1 + i].as<double>();
        HIPCHECK(c, hipMemcpyAsync(dp[i], hp[i], sizeof(double) * (size_t)V, hipMemcpyHostToDevice, c->stream));
    }
    int rc = run_t2part(c, c->t2p[0].as<double>(), o, dp[0], dp[1], dp[2], dp[3]);
    if (rc) return rc;
    for (int i = 0; i < 4; ++i)
        HIPCHECK(c, hipMemcpyAsync(hp[i], dp[i], sizeof(double) * (size_t)V, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return DMI_OK;
}

}  // extern "C"
