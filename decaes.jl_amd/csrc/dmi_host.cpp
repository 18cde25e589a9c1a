// Host-side building blocks shared by the synchronous C ABI and the request loop (see dmi_host.h).
//
// Mirrors what the reference does around its voxel loop: option validation with the same @assert messages
// (src/types.jl:19-100), the T2 and flip-angle grids (src/types.jl:108-109, src/utils.jl:7), and the T2part
// window ranges and weights (src/T2partSEcorr.jl:143-164).
#include "dmi_host.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <cstdarg>
#include <cstdio>

namespace dmi_host {

int errorf(std::string &err, int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
}

hipError_t DevBuf::ensure(size_t n) {
    if (n <= bytes && p) return hipSuccess;
    if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
    size_t want = std::max<size_t>(n, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) bytes = want;
    return e;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

bool GridKey::operator==(const GridKey &o) const {
    auto same = [](double a, double b) { return a == b || (std::isnan(a) && std::isnan(b)); };
    return m == o.m && n == o.n && K == o.K && general == o.general && fp32 == o.fp32 && same(TE, o.TE) && same(T1, o.T1) &&
           same(beta, o.beta) && same(T2lo, o.T2lo) && same(T2hi, o.T2hi) && same(minref, o.minref) &&
           same(setflip, o.setflip);
}

GridKey GridKey::of(const dmi_t2map_opts *o) {
    const bool search = std::isnan(o->SetFlipAngle);
    GridKey k;
    k.m = o->nTE; k.n = o->nT2; k.K = search ? o->nRefAngles : 1;
    k.general = o->RefConAngle != 180.0;
    k.fp32 = o->precision == DMI_PREC_MIXED;
    k.TE = o->TE; k.T1 = o->T1; k.beta = o->RefConAngle; k.T2lo = o->T2Range[0]; k.T2hi = o->T2Range[1];
    k.minref = search ? o->MinRefAngle : NAN;
    k.setflip = search ? NAN : o->SetFlipAngle;
    return k;
}

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

int validate_t2map(std::string &err, const dmi_t2map_opts *o) {
    if (!o) return errorf(err, DMI_EINVAL, "opts is NULL");
    if (!(o->MatrixSize[0] >= 1 && o->MatrixSize[1] >= 1 && o->MatrixSize[2] >= 1))
        return errorf(err, DMI_EINVAL, "MatrixSize must be a tuple of 3 positive integers, but MatrixSize = (%lld, %lld, %lld).",
                      (long long)o->MatrixSize[0], (long long)o->MatrixSize[1], (long long)o->MatrixSize[2]);
    if (!(o->nTE >= 4)) return errorf(err, DMI_EINVAL, "At least four echoes are required for T2 mapping, but nTE = %d.", o->nTE);
    if (!(o->TE > 0.0)) return errorf(err, DMI_EINVAL, "Echo spacing must be positive, but TE = %g.", o->TE);
    if (!(o->nT2 >= 2)) return errorf(err, DMI_EINVAL, "At least two T2 components are required for T2 mapping, but nT2 = %d.", o->nT2);
    if (!(0.0 < o->T2Range[0] && o->T2Range[0] < o->T2Range[1]))
        return errorf(err, DMI_EINVAL, "T2Range must a sorted 2-tuple of positive values, but T2Range = (%g, %g).", o->T2Range[0], o->T2Range[1]);
    if (!(o->T1 > 0.0)) return errorf(err, DMI_EINVAL, "T1 must be positive, but T1 = %g.", o->T1);
    if (!(o->Threshold >= 0.0 || o->Threshold == -INFINITY))
        return errorf(err, DMI_EINVAL, "First echo signal threshold must be non-negative or -Inf to force processing of every voxel, but Threshold = %g.", o->Threshold);
    if (!(0.0 <= o->MinRefAngle && o->MinRefAngle <= 180.0))
        return errorf(err, DMI_EINVAL, "Minimum refocusing angle must be in the range [0, 180], but MinRefAngle = %g.", o->MinRefAngle);
    if (!(o->nRefAngles >= 2))
        return errorf(err, DMI_EINVAL, "Maximum number of angles to check during flip angle optimization must be at least 2, but nRefAngles = %d.", o->nRefAngles);
    if (!(2 <= o->nRefAnglesMin && o->nRefAnglesMin <= o->nRefAngles))
        return errorf(err, DMI_EINVAL, "Minimum number of angles to check during flip angle optimization must be in the range [2, nRefAngles], but nRefAngles = %d and nRefAnglesMin = %d.", o->nRefAngles, o->nRefAnglesMin);
    if (o->Reg < DMI_REG_NONE || o->Reg > DMI_REG_MDP) return errorf(err, DMI_EINVAL, "Reg must be one of none, lcurve, gcv, chi2, mdp");
    if (o->Reg == DMI_REG_CHI2 && !(o->Chi2Factor > 1.0))
        return errorf(err, DMI_EINVAL, "Chi2Factor must be greater than 1.0, but Chi2Factor = %g.", o->Chi2Factor);
    if (o->Reg == DMI_REG_MDP && !(o->NoiseLevel > 0.0))
        return errorf(err, DMI_EINVAL, "Noise level must be positive, but NoiseLevel = %g.", o->NoiseLevel);
    if (!(0.0 <= o->RefConAngle && o->RefConAngle <= 180.0))
        return errorf(err, DMI_EINVAL, "Refocusing control angle must be in the range [0, 180], but RefConAngle = %g.", o->RefConAngle);
    if (!std::isnan(o->SetFlipAngle) && !(0.0 <= o->SetFlipAngle && o->SetFlipAngle <= 180.0))
        return errorf(err, DMI_EINVAL, "Fixed flip angle must be in the range [0, 180], but SetFlipAngle = %g.", o->SetFlipAngle);
    if (o->nT2 > dmi::kMaxT2) return errorf(err, DMI_EUNSUPPORTED, "nT2 = %d exceeds the device limit of %d", o->nT2, dmi::kMaxT2);
    if (o->nTE > dmi::kMaxEcho) return errorf(err, DMI_EUNSUPPORTED, "nTE = %d exceeds the device limit of %d", o->nTE, dmi::kMaxEcho);
    if (std::isnan(o->SetFlipAngle) && o->nRefAngles > dmi::kMaxAngles)
        return errorf(err, DMI_EUNSUPPORTED, "nRefAngles = %d exceeds the device limit of %d", o->nRefAngles, dmi::kMaxAngles);
    if (o->precision != DMI_PREC_FP64 && o->precision != DMI_PREC_MIXED) return errorf(err, DMI_EINVAL, "unknown precision %d", o->precision);
    if (o->schedule < DMI_SCHED_AUTO || o->schedule > DMI_SCHED_SPLIT) return errorf(err, DMI_EINVAL, "unknown schedule %d", o->schedule);
    return DMI_OK;
}

int validate_outputs(std::string &err, const dmi_t2map_opts *o, const dmi_t2map_out *d) {
    if (!d || !d->dist || !d->gdn || !d->ggm || !d->gva || !d->fnr || !d->snr || !d->alpha)
        return errorf(err, DMI_EINVAL, "dist, gdn, ggm, gva, fnr, snr and alpha outputs are required");
    if (o->SaveRegParam && (!d->mu || !d->chi2factor)) return errorf(err, DMI_EINVAL, "SaveRegParam requires mu and chi2factor outputs");
    if (o->SaveResidualNorm && !d->resnorm) return errorf(err, DMI_EINVAL, "SaveResidualNorm requires the resnorm output");
    if (o->SaveDecayCurve && !d->decaycurve) return errorf(err, DMI_EINVAL, "SaveDecayCurve requires the decaycurve output");
    if (o->SaveNNLSBasis && !d->decaybasis) return errorf(err, DMI_EINVAL, "SaveNNLSBasis requires the decaybasis output");
    return DMI_OK;
}

int build_grids(std::string &err, const dmi_t2map_opts *o, const GridPtrs &gp, hipStream_t s) {
    const GridKey k = GridKey::of(o);
    const bool search = std::isnan(o->SetFlipAngle);
    const int m = k.m, n = k.n, K = k.K;
    std::vector<double> T2(n), e2(n), lg(n), al(K);
    logrange(o->T2Range[0], o->T2Range[1], n, T2.data());
    for (int j = 0; j < n; ++j) { e2[j] = std::exp(-(o->TE / 2) / T2[j]); lg[j] = std::log(T2[j]); }
    if (search) linrange(o->MinRefAngle, 180.0, K, al.data());
    else al[0] = o->SetFlipAngle;
    DMI_HIP(err, hipMemcpyAsync(gp.E2, e2.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    DMI_HIP(err, hipMemcpyAsync(gp.logT2, lg.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
    DMI_HIP(err, hipMemcpyAsync(gp.alphas, al.data(), sizeof(double) * K, hipMemcpyHostToDevice, s));
    dmi::GridParams g{};
    g.m = m; g.n = n; g.K = K; g.general = k.general; g.fp32 = k.fp32; g.beta = o->RefConAngle;
    g.E1 = std::exp(-(o->TE / 2) / o->T1);
    g.E2 = gp.E2;
    g.alphas = gp.alphas;
    // column-major copies: columns of stride basis_ld(m); the pad row of an odd m stays zero
    const size_t bytesT = sizeof(double) * (size_t)dmi::basis_ld(m) * n * K;
    DMI_HIP(err, dmi::launch_fill(gp.AsT, bytesT, 0u, s));
    if (search) DMI_HIP(err, dmi::launch_fill(gp.dAsT, bytesT, 0u, s));
    g.As = gp.As;
    g.dAs = search ? gp.dAs : nullptr;
    g.AsT = gp.AsT;
    DMI_HIP(err, dmi::launch_grid_basis(g, s));
    if (search) DMI_HIP(err, dmi::launch_transpose(gp.dAs, gp.dAsT, m, n, K, s));
    DMI_HIP(err, hipStreamSynchronize(s));
    return DMI_OK;
}

int Grids::ensure(std::string &err, const dmi_t2map_opts *o, hipStream_t s) {
    const GridKey k = GridKey::of(o);
    if (k == key) return DMI_OK;
    const bool search = std::isnan(o->SetFlipAngle);
    const int m = k.m, n = k.n, K = k.K;
    key = GridKey{};  // invalid until the rebuild completes
    DMI_HIP(err, E2.ensure(sizeof(double) * n));
    DMI_HIP(err, logT2.ensure(sizeof(double) * n));
    DMI_HIP(err, alphas.ensure(sizeof(double) * K));
    const size_t bytes = sizeof(double) * (size_t)m * n * K;
    const size_t bytesT = sizeof(double) * (size_t)dmi::basis_ld(m) * n * K;
    if (search) {
        DMI_HIP(err, As.ensure(bytes));
        DMI_HIP(err, dAs.ensure(bytes));
        DMI_HIP(err, AsT.ensure(bytesT));
        DMI_HIP(err, dAsT.ensure(bytesT));
    } else {
        DMI_HIP(err, Afixed.ensure(bytes));
        DMI_HIP(err, AfixedT.ensure(bytesT));
    }
    int rc = build_grids(err, o, ptrs_for(search), s);
    if (rc) return rc;
    key = k;
    return DMI_OK;
}

GridPtrs Grids::ptrs() const { return ptrs_for(std::isnan(key.setflip)); }

GridPtrs Grids::ptrs_for(bool search) const {
    GridPtrs g;
    g.E2 = E2.as<double>();
    g.logT2 = logT2.as<double>();
    g.alphas = alphas.as<double>();
    if (search) {
        g.As = As.as<double>(); g.dAs = dAs.as<double>(); g.AsT = AsT.as<double>(); g.dAsT = dAsT.as<double>();
    } else {
        g.As = Afixed.as<double>(); g.AsT = AfixedT.as<double>();
    }
    return g;
}

void Grids::release() {
    for (DevBuf *b : {&E2, &logT2, &alphas, &As, &dAs, &Afixed, &AsT, &dAsT, &AfixedT}) b->release();
    key = GridKey{};
}

void LaunchBufs::release() {
    for (DevBuf *b : {&counter, &counters, &status, &scratch}) b->release();
}

dmi::T2mapParams t2map_params(const dmi_t2map_opts *o, const GridPtrs &g, const double *d_image, const double *d_alpha,
                              const dmi_t2map_out *d) {
    const int m = o->nTE, n = o->nT2;
    const bool fixed = !std::isnan(o->SetFlipAngle);
    dmi::T2mapParams p{};
    p.m = m; p.n = n; p.K = fixed ? 1 : o->nRefAngles; p.mineval = o->nRefAnglesMin; p.reg = o->Reg;
    p.mode = d_alpha ? dmi::kModeB1 : (fixed ? dmi::kModeFixed : dmi::kModeSearch);
    p.general = o->RefConAngle != 180.0;
    p.fp32 = o->precision == DMI_PREC_MIXED;
    p.lda = dmi::t2map_lda(m, n);
    p.mw = m + n;
    p.rl = p.lda;  // all rows in LDS unless a schedule says otherwise
    p.gld = 0;
    p.save_resnorm = o->SaveResidualNorm; p.save_decaycurve = o->SaveDecayCurve;
    p.save_regparam = o->SaveRegParam; p.save_basis = o->SaveNNLSBasis;
    p.legacy = o->legacy != 0;
    p.V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    p.E1 = std::exp(-(o->TE / 2) / o->T1);
    p.beta = o->RefConAngle;
    p.chi2factor = o->Chi2Factor;
    p.noiselevel = o->NoiseLevel;
    p.setflip = o->SetFlipAngle;
    p.image = d_image;
    p.E2 = g.E2;
    p.logT2 = g.logT2;
    p.alphas = g.alphas;
    p.AsT = fixed ? nullptr : g.AsT;
    p.dAsT = fixed ? nullptr : g.dAsT;
    p.Afixed = fixed ? g.AsT : nullptr;
    p.alpha_in = d_alpha;
    p.dist = d->dist; p.gdn = d->gdn; p.ggm = d->ggm; p.gva = d->gva; p.fnr = d->fnr; p.snr = d->snr; p.alpha = d->alpha;
    p.resnorm = o->SaveResidualNorm ? d->resnorm : nullptr;
    p.decaycurve = o->SaveDecayCurve ? d->decaycurve : nullptr;
    p.mu = o->SaveRegParam ? d->mu : nullptr;
    p.chi2f = o->SaveRegParam ? d->chi2factor : nullptr;
    p.decaybasis = (o->SaveNNLSBasis && !fixed) ? d->decaybasis : nullptr;
    return p;
}

int t2map_launch_raw(std::string &err, dmi::T2mapParams p, const int64_t *d_vox, int64_t nvox, unsigned *counter,
                     unsigned long long *counters, int *status, double *slots, int64_t slot_doubles, int num_cu,
                     hipStream_t s) {
    if (nvox <= 0) return DMI_OK;
    const size_t lds = dmi::t2map_lds_bytes(p.m, p.n, p.lda, p.K, p.legacy);
    int per_cu = 0;
    DMI_HIP(err, dmi::t2map_occupancy(p.reg, p.legacy, &per_cu, lds, p.rl < p.mw));
    if (per_cu < 1) return errorf(err, DMI_EUNSUPPORTED, "voxel kernel does not fit on a CU (LDS %zu bytes)", lds);
    p.slot_stride = dmi::t2map_slot_doubles(p.m, p.n, p.gld);
    const int64_t fit = slot_doubles / p.slot_stride;
    if (fit < 1) return errorf(err, DMI_ECAPACITY, "scratch of %lld doubles holds no voxel slot of %lld doubles",
                               (long long)slot_doubles, (long long)p.slot_stride);
    const int grid = (int)std::min<int64_t>(std::min<int64_t>((int64_t)num_cu * per_cu, nvox), fit);
    p.scratch = slots;
    p.vox = d_vox;
    p.nvox = nvox;
    p.counter = counter;
    p.counters = counters;
    p.status = status;
    DMI_HIP(err, dmi::launch_fill(counter, sizeof(unsigned), 0u, s));
    DMI_HIP(err, dmi::launch_t2map(p, grid, lds, s));
    return DMI_OK;
}

int t2map_launch(std::string &err, dmi::T2mapParams p, const int64_t *d_vox, int64_t nvox, LaunchBufs &lb,
                 int num_cu, hipStream_t s) {
    if (nvox <= 0) return DMI_OK;
    DMI_HIP(err, lb.counter.ensure(sizeof(unsigned) * 4));
    DMI_HIP(err, lb.counters.ensure(sizeof(unsigned long long) * 4));
    DMI_HIP(err, lb.status.ensure(sizeof(int) * 4));
    const size_t lds = dmi::t2map_lds_bytes(p.m, p.n, p.lda, p.K, p.legacy);
    int per_cu = 0;
    DMI_HIP(err, dmi::t2map_occupancy(p.reg, p.legacy, &per_cu, lds, p.rl < p.mw));
    if (per_cu < 1) return errorf(err, DMI_EUNSUPPORTED, "voxel kernel does not fit on a CU (LDS %zu bytes)", lds);
    const int64_t grid = std::min<int64_t>((int64_t)num_cu * per_cu, nvox);
    const int64_t slot_doubles = dmi::t2map_slot_doubles(p.m, p.n, p.gld) * grid;
    DMI_HIP(err, lb.scratch.ensure(sizeof(double) * (size_t)slot_doubles));
    return t2map_launch_raw(err, p, d_vox, nvox, lb.counter.as<unsigned>(), lb.counters.as<unsigned long long>(),
                            lb.status.as<int>(), lb.scratch.as<double>(), slot_doubles, num_cu, s);
}

int t2map_schedules(std::string &err, const dmi_t2map_opts *o, std::vector<Schedule> &out) {
    out.clear();
    const int m = o->nTE, n = o->nT2, K = std::isnan(o->SetFlipAngle) ? o->nRefAngles : 1;
    Schedule full;
    full.lda = dmi::t2map_lda(m, n);
    full.rl = full.lda;
    full.gld = 0;
    DMI_HIP(err, dmi::t2map_occupancy(o->Reg, o->legacy, &full.per_cu, dmi::t2map_lds_bytes(m, n, full.lda, K, o->legacy)));
    out.push_back(full);
    // Only the Tikhonov-regularized solves append rows beyond m; GCV also keeps its Gram matrix in rows [m, m + n)
    if (o->Reg != DMI_REG_LCURVE && o->Reg != DMI_REG_CHI2 && o->Reg != DMI_REG_MDP) return DMI_OK;
    // the basis rows, and every row a triangular factor or a reflector head can reach (< n + 2), stay in LDS
    int rmin = std::max(dmi::basis_ld(m), n + 2);
    rmin += (rmin % 4 == 0) ? 2 : 0;  // lda = 2 mod 4 (bank spread of 16-byte column loads)
    int best = full.per_cu;
    for (int target = full.per_cu + 1; target <= full.per_cu + 2; ++target) {
        for (int lda = full.lda - 4; lda >= rmin; lda -= 4) {
            int pc = 0;
            DMI_HIP(err, dmi::t2map_occupancy(o->Reg, o->legacy, &pc, dmi::t2map_lds_bytes(m, n, lda, K, o->legacy), true));
            if (pc >= target) {
                if (pc > best) {
                    Schedule s;
                    s.lda = lda;
                    s.rl = lda;
                    s.gld = ((m + n - lda) + 1) & ~1;
                    s.per_cu = pc;
                    out.push_back(s);
                    best = pc;
                }
                break;
            }
        }
    }
    return DMI_OK;
}

namespace {
std::mutex g_sched_mu;
std::map<std::tuple<int, int, int, int, int, int>, Schedule> g_sched;
std::tuple<int, int, int, int, int, int> sched_key(const dmi_t2map_opts *o) {
    const int K = std::isnan(o->SetFlipAngle) ? o->nRefAngles : 1;
    return {o->nTE, o->nT2, K, o->Reg, o->precision, o->RefConAngle != 180.0};
}
}  // namespace

bool schedule_lookup(const dmi_t2map_opts *o, Schedule &s) {
    std::lock_guard<std::mutex> lk(g_sched_mu);
    auto it = g_sched.find(sched_key(o));
    if (it == g_sched.end()) return false;
    s = it->second;
    return true;
}

void schedule_store(const dmi_t2map_opts *o, const Schedule &s) {
    std::lock_guard<std::mutex> lk(g_sched_mu);
    g_sched[sched_key(o)] = s;
}

int t2map_schedule(std::string &err, const dmi_t2map_opts *o, Schedule &s) {
    std::vector<Schedule> c;
    int rc = t2map_schedules(err, o, c);
    if (rc) return rc;
    if (o->schedule == DMI_SCHED_SPLIT) { s = c.back(); return DMI_OK; }  // the most workgroups per CU
    if (o->schedule == DMI_SCHED_AUTO && schedule_lookup(o, s)) return DMI_OK;
    s = c.front();
    return DMI_OK;
}

void apply_schedule(dmi::T2mapParams &p, const Schedule &s) {
    p.lda = s.lda;
    p.rl = s.rl;
    p.gld = s.rl >= p.mw ? 0 : s.gld;
}

int t2part_prepare(std::string &err, const dmi_t2part_opts *o, std::vector<double> &lg, std::vector<double> &wts,
                   dmi::T2partParams &p) {
    if (!o) return errorf(err, DMI_EINVAL, "opts is NULL");
    // legacy changes nothing in T2partSEcorr (src/T2partSEcorr.jl never reads it)
    if (!(o->MatrixSize[0] >= 1 && o->MatrixSize[1] >= 1 && o->MatrixSize[2] >= 1)) return errorf(err, DMI_EINVAL, "all(MatrixSize .>= 1)");
    if (!(o->nT2 >= 2)) return errorf(err, DMI_EINVAL, "nT2 >= 2");
    if (!(0.0 < o->T2Range[0] && o->T2Range[0] < o->T2Range[1])) return errorf(err, DMI_EINVAL, "0.0 < T2Range[1] < T2Range[2]");
    if (!(o->SPWin[0] < o->SPWin[1])) return errorf(err, DMI_EINVAL, "SPWin[1] < SPWin[2]");
    if (!(o->MPWin[0] < o->MPWin[1])) return errorf(err, DMI_EINVAL, "MPWin[1] < MPWin[2]");
    if (!(std::isnan(o->Sigmoid) || o->Sigmoid > 0)) return errorf(err, DMI_EINVAL, "Sigmoid === nothing || Sigmoid > 0");
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
        return errorf(err, DMI_EINVAL, "T2 window outside T2Range: findfirst/findlast returned nothing");
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

int64_t t2map_request_bytes(const dmi_t2map_opts *o, bool b1, bool with_t2part) {
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    const int64_t m = o->nTE, n = o->nT2;
    const bool fixed = !std::isnan(o->SetFlipAngle);
    int64_t d = V * m + V * n + 6 * V;             // image, dist, six maps
    if (b1) d += V;
    if (o->SaveResidualNorm) d += V;
    if (o->SaveDecayCurve) d += V * m;
    if (o->SaveRegParam) d += 2 * V;
    if (o->SaveNNLSBasis) d += fixed ? m * n : V * m * n;
    if (with_t2part) d += 4 * V;
    return 8 * d + 8 * V + 4096;                   // + voxel list
}

}  // namespace dmi_host
