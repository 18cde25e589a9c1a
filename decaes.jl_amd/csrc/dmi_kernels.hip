// gfx950 kernels of the voxelwise T2mapSEcorr / T2partSEcorr path.
//
// Execution model: one 64-lane wavefront per voxel, one wave per workgroup, a persistent grid pulling
// voxel indices from a device counter. Per-voxel dense state (the Lawson-Hanson workspace, right-hand
// side, duals, the Tikhonov cache with its solutions) lives in LDS; the per-voxel EPG basis and the L-curve
// caches live in a per-workgroup global scratch slot that stays resident in L2. Lanes take the
// natural parallel dimension of each step: columns for Householder applications and duals, rows for
// reflector construction, residuals and column shifts, phase states for the EPG recursion, grid angles for
// the flip-angle surrogate. Scalar control flow (active-set loop, Brent, bisection search, L-curve) is
// wave-uniform: every lane computes identical values from identical inputs, and every reduction is a
// butterfly whose partner lanes hold identical partials, finished with a readfirstlane broadcast.
//
// Arithmetic follows the reference statement by statement, with fma() where DECAES writes muladd or sits
// under @muladd (src/NNLS.jl:47, src/EPGdecaycurve.jl:858-1024); built with -ffp-contract=off.
//
// This file is compiled twice. As itself it builds every kernel, the voxel kernel with every workspace row in LDS
// (the slot branches of the split schedule compile away) and the host launchers; dmi_kernels_split.hip includes it
// with DMI_SPLIT_TU = 1 to build only the voxel kernel for the split schedule, whose rows beyond an LDS cutoff live in
// the workgroup's global slot. The instrumented build (DMI_PHASE_PROFILE) is one unit with the split branches kept.
#ifndef DMI_SPLIT_TU
#define DMI_SPLIT_TU 0
#endif
#if DMI_SPLIT_TU || defined(DMI_PHASE_PROFILE)
#define DMI_SLOT_ROWS 1
#else
#define DMI_SLOT_ROWS 0
#endif
#include <hip/hip_runtime.h>
#include <mutex>
#include <type_traits>

#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>

#include "dmi_internal.h"

namespace dmi {
namespace {

// Explicit address spaces: pointers kept in structs otherwise decay to generic (flat) addresses.
typedef __attribute__((address_space(3))) double ldbl;
typedef __attribute__((address_space(1))) double gdbl;
typedef const __attribute__((address_space(1))) double cgdbl;
typedef const __attribute__((address_space(1))) int64_t cgi64;
typedef double dv2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(3))) dv2 cldv2;
typedef __attribute__((address_space(3))) dv2 ldv2;
typedef const __attribute__((address_space(1))) dv2 cgv2;
typedef __attribute__((address_space(1))) dv2 gv2;

constexpr double kDeg2Rad = 3.141592653589793 / 180.0;  // Julia deg2rad: x * (Float64(pi)/180)
constexpr double kPi180Hi = 0.017453292519943295;      // pi/180 as a double-double
constexpr double kPi180Lo = 2.9486522708701687e-19;
constexpr double kGolden = 1.618033988749895;          // Float64(Base.MathConstants.golden)
constexpr double kSqrtEps = 1.4901161193847656e-8;
constexpr double kEps = 2.220446049250313e-16;

// ------------------------------------------------------------------------------------------------
// Wave utilities
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double rfl(double v) {
    unsigned long long u = __double_as_longlong(v);
    unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(u & 0xffffffffull));
    unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double readlane(double v, int l) {
    unsigned long long u = __double_as_longlong(v);
    unsigned lo = __builtin_amdgcn_readlane((unsigned)(u & 0xffffffffull), l);
    unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Reductions: DPP butterflies within each 16-lane row (quad swaps, half-row mirror, row mirror) leave every lane
// of row r holding the row's result s_r; two row broadcasts then combine the rows as (s3 . s2) . (s1 . s0) into
// lane 63, which is read out uniformly. Every combine is commutative, so this equals (s0 . s1) . (s2 . s3) bit for
// bit, and the result is deterministic.
// Full-mask permutations write every lane (a source lane that is switched off reads as 0), so no `old` operand
// is needed.
template <int CTRL> __device__ __forceinline__ int dpp_i(int v) { return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, true); }
template <int CTRL> __device__ __forceinline__ double dpp_d(double v) {
    unsigned long long u = __double_as_longlong(v);
    int lo = dpp_i<CTRL>((int)(u & 0xffffffffull)), hi = dpp_i<CTRL>((int)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// Row broadcasts write only the rows in ROWS; the other rows hold unspecified values, which wave_reduce never
// lets reach lane 63 (row 3 is enabled in both broadcasts and its sources, lanes 47 and 31, are always valid).
template <int CTRL, int ROWS> __device__ __forceinline__ double dpp_rows(double v) {
    unsigned long long u = __double_as_longlong(v);
    int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffull), CTRL, ROWS, 0xf, false);
    int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
constexpr int kQuadSwap1 = 0xB1;  // quad_perm [1,0,3,2]
constexpr int kQuadSwap2 = 0x4E;  // quad_perm [2,3,0,1]
constexpr int kHalfMirror = 0x141;
constexpr int kRowMirror = 0x140;
constexpr int kRowBcast15 = 0x142;  // lane 15 of each row -> the next row (rows 1 and 3 enabled)
constexpr int kRowBcast31 = 0x143;  // lane 31 -> rows 2 and 3
template <class Op> __device__ __forceinline__ double row_reduce(double v, Op op) {
    v = op(v, dpp_d<kQuadSwap1>(v));
    v = op(v, dpp_d<kQuadSwap2>(v));
    v = op(v, dpp_d<kHalfMirror>(v));
    v = op(v, dpp_d<kRowMirror>(v));
    return v;
}
template <class Op> __device__ __forceinline__ double wave_reduce(double v, Op op) {
    v = row_reduce(v, op);
    v = op(v, dpp_rows<kRowBcast15, 0xA>(v));
    v = op(v, dpp_rows<kRowBcast31, 0xC>(v));
    return readlane(v, 63);
}
__device__ __forceinline__ double wsum(double v) {
    return wave_reduce(v, [](double a, double b) { return a + b; });
}
__device__ __forceinline__ double wmaxd(double v) {
    return wave_reduce(v, [](double a, double b) { return fmax(a, b); });
}
__device__ __forceinline__ double wmind(double v) {
    return wave_reduce(v, [](double a, double b) { return fmin(a, b); });
}
// Arg-max / arg-min over lanes whose index i is the lane id (INT_MAX marks lanes that do not take part): the
// extreme value, then the lowest lane holding it, i.e. (value desc | asc, index asc). Values are never NaN.
__device__ __forceinline__ void wargmax(double &v, int &i) {
    const double vm = wmaxd(v);
    const unsigned long long hit = __ballot(v == vm && i != INT_MAX);
    v = vm;
    i = hit ? __ffsll((long long)hit) - 1 : INT_MAX;
}
__device__ __forceinline__ void wargmin(double &v, int &i) {
    const double vm = wmind(v);
    const unsigned long long hit = __ballot(v == vm && i != INT_MAX);
    v = vm;
    i = hit ? __ffsll((long long)hit) - 1 : INT_MAX;
}
__device__ __forceinline__ unsigned long long ballot(bool p) { return __ballot(p); }
// Orders the wave's LDS accesses where a workgroup barrier would: every workgroup here is one wave, whose LDS
// operations complete in issue order, so no wait is needed (a barrier compiles to lgkmcnt(0), a stall until every
// outstanding LDS operation has completed); the wavefront-scope fence only keeps the compiler from moving LDS
// accesses across it. Not for global memory written by one lane and read by another.
__device__ __forceinline__ void lds_order() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }

__device__ __forceinline__ int first_set(unsigned long long m) { return m ? __ffsll((long long)m) - 1 : -1; }
__device__ __forceinline__ int last_set(unsigned long long m) { return m ? 63 - __clzll((long long)m) : -1; }

__device__ __forceinline__ double jl_max(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : (a > b ? a : b); }
__device__ __forceinline__ double jl_min(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : (a < b ? a : b); }
__device__ __forceinline__ double jl_sign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }
__device__ __forceinline__ bool jl_isapprox(double x, double y) {
    if (x == y) return true;
    if (!isfinite(x) || !isfinite(y)) return false;
    return fabs(x - y) <= kSqrtEps * fmax(fabs(x), fabs(y));
}
__device__ __forceinline__ bool jl_isless(double a, double b) {
    if (isnan(a)) return false;
    if (isnan(b)) return true;
    return a < b;
}
__device__ __forceinline__ double norm2(double a, double b) {
    if (isnan(a) || isnan(b)) return NAN;
    if (isinf(a) || isinf(b)) return INFINITY;
    return sqrt(a * a + b * b);
}

// sind / cosd with exact degree reduction (Julia Base.Math) and a double-double pi/180 product.
__device__ __forceinline__ double sin_dd(double x) {
    double hi = x * kPi180Hi, lo = fma(x, kPi180Hi, -hi) + x * kPi180Lo, s, c;
    sincos(hi, &s, &c);
    return fma(c, lo, s);
}
__device__ __forceinline__ double cos_dd(double x) {
    double hi = x * kPi180Hi, lo = fma(x, kPi180Hi, -hi) + x * kPi180Lo, s, c;
    sincos(hi, &s, &c);
    return fma(-s, lo, c);
}
__device__ __forceinline__ double dsind(double x) {
    if (isinf(x) || isnan(x)) return NAN;
    double rx = copysign(fmod(x, 360.0), x), arx = fabs(rx);
    if (rx == 0) return rx;
    if (arx < 45.0) return sin_dd(rx);
    if (arx <= 135.0) return copysign(cos_dd(90.0 - arx), rx);
    if (arx == 180.0) return copysign(0.0, rx);
    if (arx < 225.0) return sin_dd((180.0 - arx) * jl_sign(rx));
    if (arx <= 315.0) return -copysign(cos_dd(270.0 - arx), rx);
    return sin_dd(rx - copysign(360.0, rx));
}
__device__ __forceinline__ double dcosd(double x) {
    if (isinf(x) || isnan(x)) return NAN;
    double rx = fabs(fmod(x, 360.0));
    if (rx <= 45.0) return cos_dd(rx);
    if (rx < 135.0) return sin_dd(90.0 - rx);
    if (rx <= 225.0) return -cos_dd(180.0 - rx);
    if (rx < 315.0) return sin_dd(rx - 270.0);
    return cos_dd(360.0 - rx);
}

// ------------------------------------------------------------------------------------------------
// Numbers for the EPG recursion: double, or a forward-mode dual (value, d/dalpha) as ForwardDiff.Dual
// computes in EPGJacobianFunctor (src/EPGdecaycurve.jl:224-248).
// ------------------------------------------------------------------------------------------------
struct Dual {
    double v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double s) { return {a.v * s, a.d * s}; }
__device__ __forceinline__ Dual operator/(Dual a, double s) { return {a.v / s, a.d / s}; }
__device__ __forceinline__ Dual tfma(Dual a, Dual b, Dual c) { return {fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d}; }
__device__ __forceinline__ double tfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ Dual tabs(Dual a) { return signbit(a.v) ? -a : a; }
__device__ __forceinline__ double tabs(double a) { return fabs(a); }
__device__ __forceinline__ void tsincos(Dual x, Dual *s, Dual *c) {
    double sv, cv;
    sincos(x.v, &sv, &cv);
    *s = {sv, cv * x.d};
    *c = {cv, -sv * x.d};
}
__device__ __forceinline__ void tsincos(double x, double *s, double *c) { sincos(x, s, c); }
// FP32 twins for the mixed-precision path: the recursion runs in FP32 (coefficients are formed in FP64 and
// rounded once), results are stored as FP64.
struct DualF {
    float v, d;
};
__device__ __forceinline__ DualF operator+(DualF a, DualF b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ DualF operator-(DualF a, DualF b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ DualF operator-(DualF a) { return {-a.v, -a.d}; }
__device__ __forceinline__ DualF operator*(DualF a, DualF b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ DualF operator*(DualF a, double s) { return {a.v * (float)s, a.d * (float)s}; }
__device__ __forceinline__ DualF operator/(DualF a, double s) { return {a.v / (float)s, a.d / (float)s}; }
__device__ __forceinline__ DualF tfma(DualF a, DualF b, DualF c) { return {fmaf(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d}; }
__device__ __forceinline__ float tfma(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ DualF tabs(DualF a) { return signbit(a.v) ? -a : a; }
__device__ __forceinline__ float tabs(float a) { return fabsf(a); }
__device__ __forceinline__ void tsincos(DualF x, DualF *s, DualF *c) {
    double sv, cv;
    sincos((double)x.v, &sv, &cv);
    *s = {(float)sv, (float)(cv * x.d)};
    *c = {(float)cv, (float)(-sv * x.d)};
}
__device__ __forceinline__ void tsincos(float x, float *s, float *c) {
    double sv, cv;
    sincos((double)x, &sv, &cv);
    *s = (float)sv;
    *c = (float)cv;
}

template <typename T> struct Num;
template <> struct Num<double> {
    static __device__ __forceinline__ double c(double v) { return v; }
    static __device__ __forceinline__ double var(double v, double) { return v; }
    static __device__ __forceinline__ void put(double x, gdbl *v, gdbl *, int64_t o) { v[o] = x; }
    static __device__ __forceinline__ double val(double x) { return x; }
};
template <> struct Num<Dual> {
    static __device__ __forceinline__ Dual c(double v) { return {v, 0.0}; }
    static __device__ __forceinline__ Dual var(double v, double d) { return {v, d}; }
    static __device__ __forceinline__ void put(Dual x, gdbl *v, gdbl *d, int64_t o) {
        v[o] = x.v;
        if (d) d[o] = x.d;
    }
    static __device__ __forceinline__ double val(Dual x) { return x.v; }
};
template <> struct Num<float> {
    static __device__ __forceinline__ float c(double v) { return (float)v; }
    static __device__ __forceinline__ float var(double v, double) { return (float)v; }
    static __device__ __forceinline__ void put(float x, gdbl *v, gdbl *, int64_t o) { v[o] = (double)x; }
    static __device__ __forceinline__ double val(float x) { return (double)x; }
};
template <> struct Num<DualF> {
    static __device__ __forceinline__ DualF c(double v) { return {(float)v, 0.0f}; }
    static __device__ __forceinline__ DualF var(double v, double d) { return {(float)v, (float)d}; }
    static __device__ __forceinline__ void put(DualF x, gdbl *v, gdbl *d, int64_t o) {
        v[o] = (double)x.v;
        if (d) d[o] = (double)x.d;
    }
    static __device__ __forceinline__ double val(DualF x) { return (double)x.v; }
};
// Phase-state shifts between neighbouring lanes: DPP whole-wave shifts (wave_shr:1, lane i <- lane i - 1, and
// wave_shl:1, lane i <- lane i + 1), a few cycles instead of an LDS permute. Every lane of the wave is active
// here. At a group boundary (LPC = 32: lanes 31 | 32; lane 0 and lane 63 read 0) the value crossing over is
// discarded by the caller: state s = 0 takes its own Fbar and state LPC - 1 takes Fbar = 0.
constexpr int kWaveShr1 = 0x138;
constexpr int kWaveShl1 = 0x130;
template <int CTRL> __device__ __forceinline__ float dpp_f(float a) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(a), CTRL, 0xf, 0xf, true));
}
template <int LPC> __device__ __forceinline__ float shup(float a) { return dpp_f<kWaveShr1>(a); }
template <int LPC> __device__ __forceinline__ float shdn(float a) { return dpp_f<kWaveShl1>(a); }
template <int LPC> __device__ __forceinline__ DualF shup(DualF a) { return {dpp_f<kWaveShr1>(a.v), dpp_f<kWaveShr1>(a.d)}; }
template <int LPC> __device__ __forceinline__ DualF shdn(DualF a) { return {dpp_f<kWaveShl1>(a.v), dpp_f<kWaveShl1>(a.d)}; }
template <int LPC> __device__ __forceinline__ double shup(double a) { return dpp_d<kWaveShr1>(a); }
template <int LPC> __device__ __forceinline__ double shdn(double a) { return dpp_d<kWaveShl1>(a); }
template <int LPC> __device__ __forceinline__ Dual shup(Dual a) { return {dpp_d<kWaveShr1>(a.v), dpp_d<kWaveShr1>(a.d)}; }
template <int LPC> __device__ __forceinline__ Dual shdn(Dual a) { return {dpp_d<kWaveShl1>(a.v), dpp_d<kWaveShl1>(a.d)}; }

// One EPG decay curve per LPC-lane group, lane s = phase state s+1. Each echo maps state k = (F, Fbar, Z)
// to F_{k+1}, Fbar_{k-1} and Z_k; the state-1 Fbar result becomes the new F_1 and is the echo. The
// reference's split loops drop states that cannot reach an echo before the train ends; a window of
// ETL/2 + 2 lanes reproduces them exactly (states outside it never flow back into it in time).
//
// Constant-flip train: EPGConstantFlipAngleOptions, DualFlat recursion (:948-1028) + scaling (:936-946).
template <int LPC, typename T>
__device__ __forceinline__ void epg_const(double alpha_deg, double E1, double E2, int ETL, int s, bool store, gdbl *out_v,
                          gdbl *out_d, int64_t ostride, gdbl *outT) {
    using N = Num<T>;
    const T al = N::var(alpha_deg * kDeg2Rad, kDeg2Rad);
    T sn, cs;
    tsincos(al, &sn, &cs);
    const T m0 = N::var(dsind(alpha_deg / 2), dcosd(alpha_deg / 2) * kDeg2Rad * 0.5);
    const double E2h = (E2 * E2) / 2, E1E2 = E1 * E2, E1sq = E1 * E1;
    const T a = N::c(E2h), b = cs * E2h, c = sn * E1E2, d = cs * E1sq, cp = (-c) / 2.0, mc = -c;
    T F = N::c(s == 0 ? 1.0 : 0.0), Fb = N::c(0.0), Z = N::c(0.0);
    for (int i = 0; i < ETL; ++i) {
        const T C = F + Fb, S = F - Fb;
        T echo;
        if (i < ETL - 1) {
            const T Cp = a * C, Sp = b * S;
            const T Fr = tfma(c, Z, Cp + Sp), Fbr = tfma(mc, Z, Cp - Sp), Zr = tfma(cp, S, d * Z);
            const T Fup = shup<LPC>(Fr), Fbdn = shdn<LPC>(Fbr);
            echo = Fbr;
            F = (s == 0) ? Fbr : Fup;
            Fb = (s == LPC - 1) ? N::c(0.0) : Fbdn;
            Z = Zr;
        } else {
            echo = tfma(mc, Z, tfma(a, C, -(b * S)));
        }
        if (store && s == 0) {
            const T r = tabs(m0 * echo);
            N::put(r, out_v, out_d, i * ostride);
            if (outT) outT[i] = N::val(r);
        }
    }
}

// Two constant-flip curves per lane group at once (FP64 values only: the voxel's final basis), each with exactly
// the operations of epg_const<LPC, double>: the two recursions are independent, so their dependent chains
// interleave and the lane's issue slots are filled by one while the other waits.
template <int LPC>
__device__ __forceinline__ void epg_const_pair(double alpha_deg, double E1, double E2a, double E2b, int ETL, int s,
                                               bool store_a, bool store_b, gdbl *out_a, gdbl *out_b, int64_t ostride) {
    double sn, cs;
    sincos(alpha_deg * kDeg2Rad, &sn, &cs);
    const double m0 = dsind(alpha_deg / 2);
    const double E1sq = E1 * E1, d = cs * E1sq;
    const double E2ha = (E2a * E2a) / 2, E2hb = (E2b * E2b) / 2;
    const double aa = E2ha, ba = cs * E2ha, ca = sn * (E1 * E2a), cpa = (-ca) / 2.0, mca = -ca;
    const double ab = E2hb, bb = cs * E2hb, cb = sn * (E1 * E2b), cpb = (-cb) / 2.0, mcb = -cb;
    double Fa = (s == 0) ? 1.0 : 0.0, Fba = 0.0, Za = 0.0;
    double Fb = Fa, Fbb = 0.0, Zb = 0.0;
    for (int i = 0; i < ETL; ++i) {
        const double Ca = Fa + Fba, Sa = Fa - Fba, Cb = Fb + Fbb, Sb = Fb - Fbb;
        double ea, eb;
        if (i < ETL - 1) {
            const double Cpa = aa * Ca, Spa = ba * Sa, Cpb = ab * Cb, Spb = bb * Sb;
            const double Fra = fma(ca, Za, Cpa + Spa), Fbra = fma(mca, Za, Cpa - Spa), Zra = fma(cpa, Sa, d * Za);
            const double Frb = fma(cb, Zb, Cpb + Spb), Fbrb = fma(mcb, Zb, Cpb - Spb), Zrb = fma(cpb, Sb, d * Zb);
            const double Fupa = shup<LPC>(Fra), Fbdna = shdn<LPC>(Fbra);
            const double Fupb = shup<LPC>(Frb), Fbdnb = shdn<LPC>(Fbrb);
            ea = Fbra;
            eb = Fbrb;
            Fa = (s == 0) ? Fbra : Fupa;
            Fb = (s == 0) ? Fbrb : Fupb;
            Fba = (s == LPC - 1) ? 0.0 : Fbdna;
            Fbb = (s == LPC - 1) ? 0.0 : Fbdnb;
            Za = Zra;
            Zb = Zrb;
        } else {
            ea = fma(mca, Za, fma(aa, Ca, -(ba * Sa)));
            eb = fma(mcb, Zb, fma(ab, Cb, -(bb * Sb)));
        }
        if (s == 0) {
            if (store_a) out_a[i * ostride] = fabs(m0 * ea);
            if (store_b) out_b[i * ostride] = fabs(m0 * eb);
        }
    }
}

// General refocusing control angle: EPGOptions, DualVector recursion (:722-818). Dot products follow
// StaticArrays' order (a1 b1 + a2 b2) + a3 b3 without fma; first pulse A*180, later pulses A*beta.
template <typename T> __device__ __forceinline__ T dot3(T x0, T x1, T x2, T y0, T y1, T y2) {
    return (x0 * y0 + x1 * y1) + x2 * y2;
}
template <int LPC, typename T>
__device__ __forceinline__ void epg_general(double alpha_deg, double beta, double E1, double E2, int ETL, int s, bool store,
                            gdbl *out_v, gdbl *out_d, int64_t ostride, gdbl *outT) {
    using N = Num<T>;
    const T alpha = N::var(alpha_deg, 1.0);
    const T A = alpha / 180.0;
    const T a1r = (A * 180.0) * kDeg2Rad, air = (A * beta) * kDeg2Rad;
    T s_h, c_h, sinai, cosai;
    tsincos(a1r / 2.0, &s_h, &c_h);
    tsincos(air, &sinai, &cosai);
    const T s2 = s_h * s_h, c2 = c_h * c_h, sina1 = (s_h * 2.0) * c_h;
    const T c2i = (N::c(1.0) + cosai) / 2.0, s2i = N::c(1.0) - c2i;
    const double E2sq = E2 * E2, E1E2 = E1 * E2, E1sq = E1 * E1;
    const T a1 = c2 * E2sq, b1 = s2 * E2sq, c1 = sina1 * E1E2;
    const T ai = c2i * E2sq, bi = s2i * E2sq, ci = sinai * E1E2, di = cosai * E1sq;
    const T zc0 = (-ci) / 2.0, zc1 = ci / 2.0, mci = -ci;
    const T m0 = s_h;
    T F = N::c(0.0), Fb = N::c(0.0), Z = N::c(0.0);
    if (s == 0) { F = b1 * m0; Z = ((-c1) * m0) / 2.0; }
    if (s == 1) F = a1 * m0;
    if (store && s == 0) {
        const T r = tabs(b1 * m0);
        N::put(r, out_v, out_d, 0);
        if (outT) outT[0] = N::val(r);
    }
    for (int i = 1; i < ETL; ++i) {
        const T FM = dot3(ai, bi, ci, F, Fb, Z), FbM = dot3(bi, ai, mci, F, Fb, Z), ZM = dot3(zc0, zc1, di, F, Fb, Z);
        if (store && s == 0) {
            const T r = tabs(FbM);
            N::put(r, out_v, out_d, i * ostride);
            if (outT) outT[i] = N::val(r);
        }
        const T Fup = shup<LPC>(FM), Fbdn = shdn<LPC>(FbM);
        F = (s == 0) ? FbM : Fup;
        Fb = (s == LPC - 1) ? N::c(0.0) : Fbdn;
        Z = ZM;
    }
}

template <int LPC, typename T>
__device__ __forceinline__ void epg_curve(bool general, double alpha, double beta, double E1, double E2, int ETL,
                                          int s, bool store, gdbl *out_v, gdbl *out_d, int64_t ostride,
                                          gdbl *outT) {
    if (general) epg_general<LPC, T>(alpha, beta, E1, E2, ETL, s, store, out_v, out_d, ostride, outT);
    else epg_const<LPC, T>(alpha, E1, E2, ETL, s, store, out_v, out_d, ostride, outT);
}

// Grid bases A_k and dA_k/dalpha for every flip angle of the search (EPGBasisSetEnsemble with
// ∇epg_decay_basis!, src/T2mapSEcorr.jl:339-371). Block = (angle k, T2 pair); row-major K x m x n, and the
// column-major copy AsT with column stride basis_ld(m) that the solves and the flip-angle gradient read.
template <int LPC>
__global__ __launch_bounds__(64) void grid_basis_kernel(GridParams p) {
    constexpr int per = kWave / LPC;
    const int k = blockIdx.x, lane = lane_id(), h = lane / LPC, s = lane % LPC;
    const int j = blockIdx.y * per + h;
    const bool valid = j < p.n;
    const int jj = valid ? j : p.n - 1;
    const int64_t base = (int64_t)k * p.m * p.n + jj;
    const int ldg = basis_ld(p.m);
    gdbl *outT = p.AsT ? (gdbl *)p.AsT + ((int64_t)k * p.n + jj) * ldg : nullptr;
    if (p.dAs && p.fp32)
        epg_curve<LPC, DualF>(p.general, p.alphas[k], p.beta, p.E1, p.E2[jj], p.m, s, valid, (gdbl *)p.As + base,
                              (gdbl *)p.dAs + base, p.n, outT);
    else if (p.dAs)
        epg_curve<LPC, Dual>(p.general, p.alphas[k], p.beta, p.E1, p.E2[jj], p.m, s, valid, (gdbl *)p.As + base,
                             (gdbl *)p.dAs + base, p.n, outT);
    else if (p.fp32)
        epg_curve<LPC, float>(p.general, p.alphas[k], p.beta, p.E1, p.E2[jj], p.m, s, valid, (gdbl *)p.As + base,
                              nullptr, p.n, outT);
    else
        epg_curve<LPC, double>(p.general, p.alphas[k], p.beta, p.E1, p.E2[jj], p.m, s, valid, (gdbl *)p.As + base,
                               nullptr, p.n, outT);
}

// dA_k/dalpha with columns contiguous (K x n columns of stride basis_ld(m)), for the coalesced row-lane gradient of
// the search.
__global__ void transpose_kernel(const double *src, double *dst, int m, int n, int K) {
    const int64_t total = (int64_t)K * m * n;
    const int ldg = basis_ld(m);
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        int64_t k = t / ((int64_t)m * n), r = t % ((int64_t)m * n);
        int j = (int)(r / m), i = (int)(r % m);
        dst[(k * n + j) * ldg + i] = src[(k * m + i) * n + j];
    }
}

// ------------------------------------------------------------------------------------------------
// Phase cycle counters (built only with -DDMI_PHASE_PROFILE, into a separate library): per-workgroup LDS
// accumulators of shader-clock cycles and call counts per phase, summed into g_phase at wave exit.
// ------------------------------------------------------------------------------------------------
enum Phase : int {
    kPhConstruct = 1, kPhApply = 2, kPhTriSolve = 3, kPhDelete = 4, kPhDual = 5, kPhInit = 6, kPhNnls = 7,
    kPhGram = 8, kPhTridiag = 9, kPhFlipSearch = 10, kPhVoxelBasis = 11, kPhSturm = 12, kPhGramEigs = 13,
    kPhRegTotal = 14, kPhVoxel = 15, kPhSelect = 16, kPhFeasible = 17, kPhPostApply = 18,
    kPhSwapApply = 19, kPhTriSolveOuter = 20, kPhEigBracket = 21,
    kPhCount = 24
};
#ifdef DMI_PHASE_PROFILE
__device__ unsigned long long g_phase[2 * kPhCount];
__shared__ unsigned long long s_phase[2 * kPhCount];
#define DMI_PT0() unsigned long long _pt0 = clock64()
// DMI_LAP(k): charge the cycles since the last DMI_PT0 / DMI_LAP to phase k and restart the timer
#define DMI_LAP(k)                                                 \
    do {                                                           \
        const unsigned long long _now = clock64();                 \
        if (lane_id() == 0) {                                      \
            s_phase[k] += _now - _pt0;                             \
            s_phase[kPhCount + (k)] += 1;                          \
        }                                                          \
        _pt0 = _now;                                               \
    } while (0)
#define DMI_PT1(k)                                                 \
    do {                                                           \
        if (lane_id() == 0) {                                      \
            s_phase[k] += clock64() - _pt0;                        \
            s_phase[kPhCount + (k)] += 1;                          \
        }                                                          \
    } while (0)
#define DMI_PHASE_BEGIN() \
    do { if (lane_id() == 0) for (int _k = 0; _k < 2 * kPhCount; ++_k) s_phase[_k] = 0; } while (0)
#define DMI_PHASE_END() \
    do { if (lane_id() == 0) for (int _k = 0; _k < 2 * kPhCount; ++_k) atomicAdd(&g_phase[_k], s_phase[_k]); } while (0)
// Tikhonov solves by the number of Tikhonov rows they appended (columns that entered the active set)
__device__ unsigned long long g_tikh_rows[kMaxT2 + 1];
#define DMI_TIKH_ROWS(k) do { if (lane_id() == 0) atomicAdd(&g_tikh_rows[(k)], 1ull); } while (0)
// voxels by the number of Laguerre steps their GCV eigenvalue brackets took
__device__ unsigned long long g_eig_steps[32];
#define DMI_EIG_STEPS(k) do { if (lane_id() == 0) atomicAdd(&g_eig_steps[(k) < 31 ? (k) : 31], 1ull); } while (0)
#else
#define DMI_TIKH_ROWS(k) do {} while (0)
#define DMI_EIG_STEPS(k) do {} while (0)
#define DMI_PT0() do {} while (0)
#define DMI_PT1(k) do {} while (0)
#define DMI_LAP(k) do {} while (0)
#define DMI_PHASE_BEGIN() do {} while (0)
#define DMI_PHASE_END() do {} while (0)
#endif

// ------------------------------------------------------------------------------------------------
// Voxel selection: indices with image[I, 1] > Threshold (src/T2mapSEcorr.jl:177). Order is irrelevant:
// voxels are independent and results are written at their own index.
// ------------------------------------------------------------------------------------------------
__global__ void select_kernel(const double *image, int64_t v0, int64_t v1, double thr, int64_t *vox, unsigned int *count) {
    for (int64_t I = v0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < v1; I += (int64_t)gridDim.x * blockDim.x)
        if (image[I] > thr) vox[atomicAdd(count, 1u)] = I;
}

// 32-bit word fill, so counters and padding are cleared by a kernel node rather than a memset
__global__ void fill_kernel(unsigned int *p, int64_t nwords, unsigned int v) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nwords; t += (int64_t)gridDim.x * blockDim.x)
        p[t] = v;
}

// ------------------------------------------------------------------------------------------------
// Static scratch of the allocation-free entry point (dmi_t2map_f64_device_noalloc): grid bases at the device
// limits, a voxel-list chunk, the per-workgroup slots of the voxel kernel and the launch counters, in one array
// each device holds for the life of the process. Host code never passes `scratch` itself to a kernel: it takes
// the device's own address of the array once with hipGetSymbolAddress (noalloc_scratch below).
// ------------------------------------------------------------------------------------------------
#if !DMI_SPLIT_TU
static __device__ __align__(256) double scratch[kNoAllocDoubles];
#endif

// ------------------------------------------------------------------------------------------------
// Per-voxel wave state
// ------------------------------------------------------------------------------------------------
struct Wave {
    // LDS. A is column-major with column stride lda (even, lda = 2 mod 4, lda >= mw rounded up to even):
    // a lane that owns a column streams it with 16-byte loads, and 16-byte loads of 16 consecutive lanes
    // at stride lda hit distinct banks. Every LDS double array starts 16-byte aligned.
    ldbl *A, *b, *zz, *x, *cmu, *crn, *csemi;  // zz: the warm start's residual, kept in b's storage
    // in the workgroup's global slot; every lane reads and writes only its own entries (rows L and L + 64 of b0,
    // entry L of xf), except the fit of save_results, which reads all of xf after a workgroup barrier
    gdbl *xf, *b0;
    gdbl *cx;         // the cached Tikhonov solutions, n per entry, in the workgroup's global slot (each lane stores
                      // and later loads only its own entries)
    ldbl *spt, *spc;  // legacy = true: spline knots (kSplineKnots) and coefficients (kSplineMaxPts)
    int idxl;         // the column at position L when the last solve ended (the reference's idx[L])
    int lda, mw, m, n, L;
    // Rows [0, rl) of the workspace are in LDS (column stride lda); rows [rl, mw) of column j, when rl < mw, are in
    // this workgroup's global slot at G[j * gld + (r - rl)] (the split schedule, which lets more workgroups share a
    // CU's LDS). rl is even and at least max(basis_ld(m), n + 2): the basis rows and every row of the triangular
    // factor and of a reflector's head stay in LDS; only Tikhonov rows appended late live in G.
    int rl, gld;
    // Every test of a row against rl below is written `!DMI_SLOT_ROWS || ...` (or `DMI_SLOT_ROWS && ...`), so the
    // unit built without slot rows keeps no slot code at all.
    gdbl *G;
    // the dual w_j of position j = L (the reference's w), in a register: only this lane writes or reads it
    double wl;
    // per-workgroup global scratch
    gdbl *A0s, *lc;
    cgdbl *A0;   // basis used by the regularization solves (A0s or the fixed basis), column-major, stride basis_ld(m)
    int ccur;
    unsigned long long flops, flops32, solves, nfailed;
    int status;
    bool failed;  // this voxel gets no outputs (a per-voxel cache overflowed)

    __device__ __forceinline__ ldbl &Aat(int i, int j) { return A[j * lda + i]; }
    // row i of column j wherever it lives (i < rl: LDS, else the global slot); per-lane rows may differ
    __device__ __forceinline__ double Aget(int i, int j) {
        return (!DMI_SLOT_ROWS || i < rl) ? (double)A[j * lda + i] : (double)G[j * gld + (i - rl)];
    }
    __device__ __forceinline__ void Aput(int i, int j, double v) {
        if (!DMI_SLOT_ROWS || i < rl) A[j * lda + i] = v;
        else G[j * gld + (i - rl)] = v;
    }

    // -------------------------------- Lawson-Hanson NNLS (src/NNLS.jl) --------------------------------
    // construct_apply_householder! (:259-336), 0-based: pivot row ip, column jp, rows [0, mr). Lane L holds
    // rows L, L + 64 and L + 128 of column jp (mr <= m + n <= 188) in registers; each element is divided by
    // alpha once and reused for the right-hand-side update and the column write.
    __device__ __forceinline__ double construct_apply_householder(int ip, int jp, int mr) {
        if (ip >= mr) return 0.0;
        const ldbl *cp = A + jp * lda;
        // t: column ip, which moves to column jp when the pivot column is swapped into place (read up front;
        // nothing writes it before the swap)
        double c[3], bv[3], t[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int i = L + k * kWave;
            c[k] = (i < mr) ? Aget(i, jp) : 0.0;
            bv[k] = (i < mr) ? b[i] : 0.0;
            t[k] = (ip != jp && i < mr) ? Aget(i, ip) : 0.0;
        }
        double alpha = rfl(cp[ip]);
        double part = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int i = L + k * kWave;
            if (i >= ip && i < mr) part = fma(c[k], c[k], part);
        }
        double xnorm = sqrt(wsum(part));
        flops += 2ull * (mr - ip) + 8;
        if (xnorm == 0) return -1.0;
        double beta = copysign(xnorm, alpha);
        alpha = alpha + beta;
        double tau = alpha / beta;
        double bip = rfl(b[ip]);
        double u[3];
        part = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int i = L + k * kWave;
            u[k] = 0.0;
            if (i < mr) {  // wave-uniformly false for the third row unless mr > 128
                u[k] = c[k] / alpha;
                if (i > ip) part = fma(bv[k], u[k], part);
            }
        }
        double sm = bip + wsum(part);
        sm *= -tau;
        double A1 = -beta, b1 = bip + sm;
        flops += 3ull * (mr - ip);
        // b1 / A1 > 0 (A1 = -beta is nonzero): a sign test, with the divide kept for the quotient-underflow case
        const bool pos = (b1 > 0 && A1 > 0) || (b1 < 0 && A1 < 0);
        if (!(fabs(b1) >= 1e-290 ? pos : (b1 / A1 > 0))) return -1.0;
        if (ip < mr - 1) {
            // branch-free per lane: column ip takes (c above the pivot, A1, u below), column jp takes the old
            // column ip, and b takes (b unchanged, b1, b + sm u)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int i = L + k * kWave;
                if (i >= mr) continue;
                Aput(i, ip, (i < ip) ? c[k] : ((i == ip) ? A1 : u[k]));
                if (ip != jp) Aput(i, jp, t[k]);
                b[i] = (i < ip) ? bv[k] : ((i == ip) ? b1 : fma(sm, u[k], bv[k]));
            }
            flops += 3ull * (mr - ip);
        } else {
            tau = 0.0;
            if (ip != jp)
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int i = L + k * kWave;
                    if (i < mr) { Aput(i, ip, c[k]); Aput(i, jp, t[k]); }
                }
        }
        lds_order();
        return tau;
    }

    // apply_householder_dual! (:379-436): columns j > j1 (one per lane), rows [j1, m1). The lane streams its
    // own column and the reflector (column j1) with 16-byte loads over even-aligned row pairs; an odd first or
    // last row is taken singly. Dot products use independent partial sums (the reference's @simd loops
    // reassociate the same sums).
    __device__ __forceinline__ void apply_householder_dual(double tau, int j1, int m1) {
        if (j1 >= m1 - 1) return;
        const int j = L;
        if (j > j1 && j < n) {
            ldbl *cj = A + j * lda;
            const ldbl *cu = A + j1 * lda;
            cldv2 *vu = (cldv2 *)cu, *vb = (cldv2 *)b;
            ldv2 *vj = (ldv2 *)cj;
            const int lo = j1 + 1, hi = m1;
            const int h = lo & 1, bs = lo + h, t = (hi - bs) & 1, be = hi - t;
            const int q0 = bs >> 1, qe = be >> 1;
            // pairs [q0, q1) are in LDS, pairs [qg, qe) in the global slot (split schedule; empty otherwise)
            const int qR = DMI_SLOT_ROWS ? (rl >> 1) : qe, q1 = min(qe, qR), qg = max(q0, qR);
            double s0 = cj[j1], s1 = 0.0, s2 = 0.0, s3 = 0.0;
            if (h) s1 = cj[lo] * cu[lo];
            if (t) s2 = Aget(hi - 1, j) * Aget(hi - 1, j1);
            int q = q0;
            // eight row pairs per step (all sixteen loads issued before the first use), then four, two, one;
            // accumulator k always takes the same row of a step, as a plain two-pair loop would
            for (; q + 8 <= q1; q += 8) {
                dv2 a[8], u[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) { a[k] = vj[q + k]; u[k] = vu[q + k]; }
#pragma unroll
                for (int k = 0; k < 8; k += 2) {
                    s0 = fma(a[k].x, u[k].x, s0); s1 = fma(a[k].y, u[k].y, s1);
                    s2 = fma(a[k + 1].x, u[k + 1].x, s2); s3 = fma(a[k + 1].y, u[k + 1].y, s3);
                }
            }
            if (q + 4 <= q1) {
                const dv2 a0 = vj[q], u0 = vu[q], a1 = vj[q + 1], u1 = vu[q + 1];
                const dv2 a2 = vj[q + 2], u2 = vu[q + 2], a3 = vj[q + 3], u3 = vu[q + 3];
                s0 = fma(a0.x, u0.x, s0); s1 = fma(a0.y, u0.y, s1); s2 = fma(a1.x, u1.x, s2); s3 = fma(a1.y, u1.y, s3);
                s0 = fma(a2.x, u2.x, s0); s1 = fma(a2.y, u2.y, s1); s2 = fma(a3.x, u3.x, s2); s3 = fma(a3.y, u3.y, s3);
                q += 4;
            }
            if (q + 2 <= q1) {
                const dv2 a0 = vj[q], u0 = vu[q], a1 = vj[q + 1], u1 = vu[q + 1];
                s0 = fma(a0.x, u0.x, s0); s1 = fma(a0.y, u0.y, s1); s2 = fma(a1.x, u1.x, s2); s3 = fma(a1.y, u1.y, s3);
                q += 2;
            }
            if (q < q1) {
                const dv2 a0 = vj[q], u0 = vu[q];
                s0 = fma(a0.x, u0.x, s0);
                s1 = fma(a0.y, u0.y, s1);
            }
            // global pairs continue the same accumulator pattern: pair q goes to (s0, s1) when q - q0 is even
            if (DMI_SLOT_ROWS && qg < qe) {
                cgv2 *gj = (cgv2 *)(G + (size_t)j * gld), *gu = (cgv2 *)(G + (size_t)j1 * gld);
                int r = qg;
                auto acc = [&](const dv2 &a, const dv2 &u, int pq) {
                    if (((pq - q0) & 1) == 0) { s0 = fma(a.x, u.x, s0); s1 = fma(a.y, u.y, s1); }
                    else { s2 = fma(a.x, u.x, s2); s3 = fma(a.y, u.y, s3); }
                };
                for (; r + 8 <= qe; r += 8) {
                    dv2 a[8], u[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) { a[k] = gj[r + k - qR]; u[k] = gu[r + k - qR]; }
#pragma unroll
                    for (int k = 0; k < 8; ++k) acc(a[k], u[k], r + k);
                }
                for (; r < qe; ++r) {
                    const dv2 a = gj[r - qR], u = gu[r - qR];
                    acc(a, u, r);
                }
            }
            const double sm = ((s0 + s1) + (s2 + s3)) * -tau;
            cj[j1] = cj[j1] + sm;
            double w0 = 0.0, w1 = 0.0;
            if (h) {
                const double a = fma(sm, cu[lo], cj[lo]);
                w0 = a * b[lo];
                cj[lo] = a;
            }
            q = q0;
            // NB row pairs per step: all loads are issued before the stores of the step; eight, then four, then
            // single pairs (w0 and w1 take the rows in order whatever the step)
            auto update = [&](auto nb, int qs) {
                constexpr int NB = decltype(nb)::value;
                dv2 a[NB], u[NB], bb[NB];
#pragma unroll
                for (int k = 0; k < NB; ++k) { a[k] = vj[qs + k]; u[k] = vu[qs + k]; bb[k] = vb[qs + k]; }
#pragma unroll
                for (int k = 0; k < NB; ++k) {
                    a[k].x = fma(sm, u[k].x, a[k].x);
                    a[k].y = fma(sm, u[k].y, a[k].y);
                    w0 = fma(a[k].x, bb[k].x, w0);
                    w1 = fma(a[k].y, bb[k].y, w1);
                }
#pragma unroll
                for (int k = 0; k < NB; ++k) vj[qs + k] = a[k];
            };
            for (; q + 8 <= q1; q += 8) update(std::integral_constant<int, 8>(), q);
            if (q + 4 <= q1) { update(std::integral_constant<int, 4>(), q); q += 4; }
            for (; q < q1; ++q) {
                dv2 a = vj[q];
                const dv2 u = vu[q], bb = vb[q];
                a.x = fma(sm, u.x, a.x);
                a.y = fma(sm, u.y, a.y);
                w0 = fma(a.x, bb.x, w0);
                w1 = fma(a.y, bb.y, w1);
                vj[q] = a;
            }
            if (DMI_SLOT_ROWS && qg < qe) {
                gv2 *gj = (gv2 *)(G + (size_t)j * gld);
                cgv2 *gu = (cgv2 *)(G + (size_t)j1 * gld);
                int r = qg;
                for (; r + 8 <= qe; r += 8) {
                    dv2 a[8], u[8], bb[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) { a[k] = gj[r + k - qR]; u[k] = gu[r + k - qR]; bb[k] = vb[r + k]; }
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        a[k].x = fma(sm, u[k].x, a[k].x);
                        a[k].y = fma(sm, u[k].y, a[k].y);
                        w0 = fma(a[k].x, bb[k].x, w0);
                        w1 = fma(a[k].y, bb[k].y, w1);
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k) gj[r + k - qR] = a[k];
                }
                for (; r < qe; ++r) {
                    dv2 a = gj[r - qR];
                    const dv2 u = gu[r - qR], bb = vb[r];
                    a.x = fma(sm, u.x, a.x);
                    a.y = fma(sm, u.y, a.y);
                    w0 = fma(a.x, bb.x, w0);
                    w1 = fma(a.y, bb.y, w1);
                    gj[r - qR] = a;
                }
            }
            if (t) {
                const double a = fma(sm, Aget(hi - 1, j1), Aget(hi - 1, j));
                w1 = fma(a, b[hi - 1], w1);
                Aput(hi - 1, j, a);
            }
            wl = w0 + w1;
        }
        flops += 6ull * (m1 - j1) * (n - j1 - 1);
        lds_order();
    }

    // compute_dual! (:441-470): w_j = sum_{i >= j1} A_ij b_i for j >= j1, rows [j1, m1).
    __device__ __forceinline__ void compute_dual(int j1, int m1) {
        const int j = L;
        if (j >= j1 && j < n && j1 < m1) {
            const ldbl *cj = A + j * lda;
            cldv2 *vj = (cldv2 *)cj, *vb = (cldv2 *)b;
            const int lo = j1, hi = m1;
            const int h = lo & 1, bs = lo + h, t = (hi - bs) & 1, be = hi - t;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            if (h) s0 = cj[lo] * b[lo];
            if (t) s1 = Aget(hi - 1, j) * b[hi - 1];
            int q = bs >> 1;
            const int q0 = q, qe = be >> 1, qR = DMI_SLOT_ROWS ? (rl >> 1) : qe, q1 = min(qe, qR), qg = max(q0, qR);
            // four row pairs per step (all eight loads issued before the first use), then two, then one;
            // accumulator k always takes the same row of a step, as a plain two-pair loop would
            for (; q + 4 <= q1; q += 4) {
                const dv2 a0 = vj[q], u0 = vb[q], a1 = vj[q + 1], u1 = vb[q + 1];
                const dv2 a2 = vj[q + 2], u2 = vb[q + 2], a3 = vj[q + 3], u3 = vb[q + 3];
                s0 = fma(a0.x, u0.x, s0); s1 = fma(a0.y, u0.y, s1); s2 = fma(a1.x, u1.x, s2); s3 = fma(a1.y, u1.y, s3);
                s0 = fma(a2.x, u2.x, s0); s1 = fma(a2.y, u2.y, s1); s2 = fma(a3.x, u3.x, s2); s3 = fma(a3.y, u3.y, s3);
            }
            if (q + 2 <= q1) {
                const dv2 a0 = vj[q], u0 = vb[q], a1 = vj[q + 1], u1 = vb[q + 1];
                s0 = fma(a0.x, u0.x, s0); s1 = fma(a0.y, u0.y, s1); s2 = fma(a1.x, u1.x, s2); s3 = fma(a1.y, u1.y, s3);
                q += 2;
            }
            if (q < q1) {
                const dv2 a0 = vj[q], u0 = vb[q];
                s0 = fma(a0.x, u0.x, s0);
                s1 = fma(a0.y, u0.y, s1);
            }
            // global pairs (split schedule) continue the pattern: pair q goes to (s0, s1) when q - q0 is even
            if (DMI_SLOT_ROWS && qg < qe) {
                cgv2 *gj = (cgv2 *)(G + (size_t)j * gld);
                auto acc = [&](const dv2 &a, const dv2 &u, int pq) {
                    if (((pq - q0) & 1) == 0) { s0 = fma(a.x, u.x, s0); s1 = fma(a.y, u.y, s1); }
                    else { s2 = fma(a.x, u.x, s2); s3 = fma(a.y, u.y, s3); }
                };
                int r = qg;
                for (; r + 8 <= qe; r += 8) {
                    dv2 a[8], u[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) { a[k] = gj[r + k - qR]; u[k] = vb[r + k]; }
#pragma unroll
                    for (int k = 0; k < 8; ++k) acc(a[k], u[k], r + k);
                }
                for (; r < qe; ++r) {
                    const dv2 a = gj[r - qR], u = vb[r];
                    acc(a, u, r);
                }
            }
            wl = (s0 + s1) + (s2 + s3);
        } else if (j >= j1 && j < n) {
            wl = 0.0;
        }
        flops += 2ull * (m1 - j1) * (n - j1);
        lds_order();
    }

    // solve_triangular_system! (:506-524) on R = A[0:k, 0:k]; lane i holds z_i in a register. The reference's
    // zz[j] / A[j,j] is formed in lane j from its own reciprocal y = 1/A[j,j] as q = q0 + (z - q0 A[j,j]) y with
    // q0 = z y: Markstein's correction, the correctly rounded quotient when y is the correctly rounded
    // reciprocal. Every lane evaluates it on its own z (only lane j's is the final z_j at step j), so the
    // sequential chain per column is fma, mul, fma, fma, readlane; column j - 3 is loaded during step j.
    // Lane i's z is final from step i on (later steps only update lanes below j), so its quotient is formed once
    // more after the loop instead of being captured at step i.
    __device__ __forceinline__ double tri_solve(int k, double zr) {
        if (k <= 0) return zr;
        const double aL = (L < k) ? Aat(L, L) : 1.0;
        const double rL = 1.0 / aL;
        // rows L < 64 of any column j < n stay inside the LDS allocation (b and x follow the matrix, and the
        // allocation is padded to (n - 1) lda + 64 doubles at least)
        // volatile: keeps each load at its place in the pipeline instead of being moved next to its use
        volatile ldbl *Av = A;
        // three column registers in rotation, unrolled so that no register is copied: the column of step j - 3
        // is loaded as soon as step j has used its register, two steps before it is needed
        double ca = Av[(k - 1) * lda + L];
        double cb = Av[max(k - 2, 0) * lda + L];
        double cc = Av[max(k - 3, 0) * lda + L];
        auto step = [&](int j, double c) {
            const double q0 = zr * rL;
            const double qL = fma(fma(-q0, aL, zr), rL, q0);
            const double qj = readlane(qL, j);
            zr = fma((L < j) ? c : 0.0, -qj, zr);
        };
        // whole blocks of three steps (one exit, so every pass around the loop has the same loads in flight),
        // then the last one or two steps with the columns already in ca and cb
        int j = k - 1;
        for (; j >= 2; j -= 3) {
            step(j, ca);
            ca = Av[max(j - 3, 0) * lda + L];
            __builtin_amdgcn_sched_barrier(0);
            step(j - 1, cb);
            cb = Av[max(j - 4, 0) * lda + L];
            __builtin_amdgcn_sched_barrier(0);
            step(j - 2, cc);
            cc = Av[max(j - 5, 0) * lda + L];
        }
        if (j >= 0) step(j, ca);
        if (j >= 1) step(j - 1, cb);
        flops += (unsigned long long)k * k;
        const double q0 = zr * rL;
        return (L < k) ? fma(fma(-q0, aL, zr), rL, q0) : zr;
    }

    // Remove coefficient imv from set P by Givens rotations and column shifts, then drop any further
    // nonpositive coefficients (:729-779 / :965-1015). pidx / xp: the column and coefficient at position L.
    // The rows below the diagonal of an active position hold zeros in the reference; here they are left as they
    // are (the NNLS loop does not clear the reflector it leaves there) and stand for zeros: the removed column is
    // cleared below its diagonal before the rotations fill it in, and the rotation at rows (i - 1, i) skips the
    // other active positions j < i, whose two entries are zero before and after it in the reference (below the
    // diagonal, or, for j = i - 1 > imv, the zero the previous rotation wrote).
    __device__ __forceinline__ void nnls_delete(int &imv, int &nsetp, int mrows, int &pidx, double &xp) {
        for (;;) {
            if (L == imv) xp = 0.0;
            for (int r = imv + 1 + L; r < mrows; r += kWave) Aput(r, imv, 0.0);
            lds_order();
            if (imv != nsetp - 1) {
                for (int i = imv + 1; i <= nsetp - 1; ++i) {
                    double aa = rfl(Aat(i - 1, i)), bb = rfl(Aat(i, i));
                    double sig = hypot(aa, bb), cc = aa / sig, ss = bb / sig;
                    lds_order();
                    const int j = L;
                    if (j < n && (j >= i || j == imv)) {
                        if (j == i) { Aat(i - 1, i) = sig; Aat(i, i) = 0.0; }
                        else {
                            double u = Aat(i - 1, j), v = Aat(i, j);
                            Aat(i - 1, j) = fma(cc, u, ss * v);
                            Aat(i, j) = fma(-ss, u, cc * v);
                        }
                    }
                    if (L == 0) {
                        double u = b[i - 1], v = b[i];
                        b[i - 1] = fma(cc, u, ss * v);
                        b[i] = fma(-ss, u, cc * v);
                    }
                    flops += 6ull * (n + 1);
                    lds_order();
                }
                for (int r = L; r < mrows; r += kWave) {
                    if (!DMI_SLOT_ROWS || r < rl) {
                        double t = Aat(r, imv);
                        for (int j = imv; j < nsetp - 1; ++j) Aat(r, j) = Aat(r, j + 1);
                        Aat(r, nsetp - 1) = t;
                    } else {
                        gdbl *gr = G + (r - rl);
                        double t = gr[(size_t)imv * gld];
                        for (int j = imv; j < nsetp - 1; ++j) gr[(size_t)j * gld] = gr[(size_t)(j + 1) * gld];
                        gr[(size_t)(nsetp - 1) * gld] = t;
                    }
                }
                // positions imv + 1 .. nsetp - 1 move down by one; position imv goes last
                const int pim = readlane(pidx, imv);
                const double xim = readlane(xp, imv);
                const int pdn = __shfl_down(pidx, 1);
                const double xdn = __shfl_down(xp, 1);
                if (L >= imv && L < nsetp - 1) { pidx = pdn; xp = xdn; }
                else if (L == nsetp - 1) { pidx = pim; xp = xim; }
                lds_order();
            }
            nsetp -= 1;
            bool bad = (L < nsetp) && (xp <= 0);
            int f = first_set(ballot(bad));
            if (f < 0) break;
            imv = f;
        }
    }

    // unsafe_nnls! (:605-825) and its Tikhonov twin unsafe_nnls!(work, λ) (:827-1061). Workspace rows
    // [0, M) with M = m (plain) or m + n (Tikhonov rows appended lazily when a column first enters).
    // The permutation (position -> column), the coefficients of the active positions and the set of columns
    // whose Tikhonov row is present live in registers (lane = position) and a wave-uniform bit mask; x, idx
    // and diag are written back to LDS on exit.
    __device__ __forceinline__ void nnls_main(bool tikh, double lam, int &nsetp_out, double &rnorm_out) {
        const int M = tikh ? m + n : m;
        int mr = m;  // current number of active rows (the reference's m)
        int nsetp = 0, iter = 0;
        const int max_iter = 3 * n;
        bool terminated = false;
        int pidx = L;           // column at position L (the reference's idx)
        double xp = 0.0;        // its coefficient (zero outside the active set)
        unsigned long long dmask = 0ull;  // columns whose Tikhonov row has been appended
        for (;;) {
            if (tikh ? (nsetp >= n) : (nsetp >= n || nsetp >= mr)) break;
            int jmax = -1, cmax = -1;
            double tau = 0.0;
            bool term = false;
            DMI_PT0();
            double wv = (L >= nsetp && L < n) ? wl : 0.0;
            for (;;) {
                double v = (wv > 0) ? wv : 0.0;
                int wi = (v > 0) ? L : INT_MAX;
                wargmax(v, wi);
                if (!(v > 0)) { term = true; break; }
                jmax = wi;
                cmax = readlane(pidx, jmax);
                const bool newdiag = tikh && !((dmask >> cmax) & 1ull);
                if (newdiag) {
                    if (L == 0) Aput(mr, jmax, lam);
                    lds_order();
                }
                const int mrh = tikh ? min(mr + 1, M) : mr;
                {
                    DMI_PT0();
                    tau = construct_apply_householder(nsetp, jmax, mrh);
                    DMI_PT1(kPhConstruct);
                }
                if (tau >= 0) break;
                if (L == jmax) { wv = 0.0; wl = 0.0; }
                if (L == 0 && tikh && mr < M) Aput(mr, jmax, 0.0);
                lds_order();
            }
            DMI_LAP(kPhSelect);
            if (term) break;
            if (tikh && !((dmask >> cmax) & 1ull)) {
                mr = min(mr + 1, M);
                dmask |= 1ull << cmax;
            }
            nsetp += 1;
            {
                // swap positions nsetp - 1 and jmax (both inactive, so both coefficients are zero)
                const int pa = readlane(pidx, nsetp - 1);
                if (L == nsetp - 1) pidx = cmax;
                else if (L == jmax) pidx = pa;
            }
            {
                DMI_PT0();
                if (nsetp < n) apply_householder_dual(tau, nsetp - 1, mr);
                DMI_PT1(kPhApply);
            }
            DMI_LAP(kPhSwapApply);
            // The reference now zeroes the new pivot column below its diagonal and its dual; neither is read
            // again while the position is active (nnls_delete treats those rows as zeros, and duals are read
            // and recomputed only for inactive positions), so both are left as they are.
            double zr = (L < nsetp) ? b[L] : 0.0;
            DMI_LAP(kPhPostApply);
            {
                DMI_PT0();
                zr = tri_solve(nsetp, zr);
                DMI_PT1(kPhTriSolve);
            }
            bool dual_flag = false;
            DMI_LAP(kPhTriSolveOuter);
            for (;;) {
                iter += 1;
                if (iter > max_iter) { status |= kStatusNnlsMaxIter; terminated = true; break; }
                if (ballot(L < nsetp && zr <= 0) == 0) break;  // alpha == 2 without a candidate: all feasible
                double t = INFINITY;
                int ti = INT_MAX;
                if (L < nsetp && zr <= 0) {
                    double tt = -xp / (zr - xp);
                    if (tt < 2.0) { t = tt; ti = L; }
                }
                wargmin(t, ti);
                if (ti == INT_MAX) break;  // alpha == 2: all feasible
                dual_flag = true;
                const double alpha = t;
                int imv = ti;
                if (L < nsetp) xp = fma(alpha, zr - xp, xp);
                {
                    DMI_PT0();
                    nnls_delete(imv, nsetp, mr, pidx, xp);
                    DMI_PT1(kPhDelete);
                }
                zr = (L < nsetp) ? b[L] : 0.0;
                {
                    DMI_PT0();
                    zr = tri_solve(nsetp, zr);
                    DMI_PT1(kPhTriSolve);
                }
            }
            DMI_LAP(kPhFeasible);
            if (terminated) break;
            if (dual_flag) {
                DMI_PT0();
                compute_dual(nsetp, mr);
                DMI_PT1(kPhDual);
            }
            if (L < nsetp) xp = zr;
        }
        if (L < n) {
            x[pidx] = (L < nsetp) ? xp : 0.0;
            idxl = pidx;
        }
        lds_order();
        const int Mlim = tikh ? M : mr;
        double part = 0.0;
        for (int i = nsetp + L; i < Mlim; i += kWave) { double bi = b[i]; part = fma(bi, bi, part); }
        rnorm_out = sqrt(wsum(part));
        nsetp_out = nsetp;
        if (tikh) DMI_TIKH_ROWS(__popcll(dmask));
    }

    // Rows [0, basis_ld(m)) of the n columns of a column-major global basis (column stride basis_ld(m), 16-byte
    // aligned) straight into the LDS matrix: one 16-byte global-to-LDS load per lane and column from lanes
    // [0, basis_ld(m) / 2), which lands lane l's bytes at the column start + 16 l. No register holds the data, so
    // all n loads stay in flight together whatever the caller keeps live; basis_wait() completes them.
    __device__ __forceinline__ void load_basis(cgdbl *__restrict__ src) {
        const int ldg = basis_ld(m);
        if (L < ldg / 2)
            for (int j = 0; j < n; ++j)
                __builtin_amdgcn_global_load_lds((__attribute__((address_space(1))) void *)(src + (size_t)j * ldg + 2 * L),
                                                 (__attribute__((address_space(3))) void *)(A + j * lda), 16, 0, 0);
    }
    // Callers follow it with __syncthreads(), whose LDS fence keeps every read of the copied rows after the wait.
    __device__ __forceinline__ static void basis_wait() { asm volatile("s_waitcnt vmcnt(0)"); }

    // NNLSProblem solve! with the reference warm start (src/lsqnonneg.jl:30-84), or the Tikhonov form
    // (:86-164) when tikh. A0g: global column-major m x n basis (column stride basis_ld(m)); right-hand side b0
    // in LDS.
    __device__ __forceinline__ void solve(cgdbl *__restrict__ A0g, bool tikh, double mu, int &nsetp, double &rnorm) {
        DMI_PT0();
        const int M = tikh ? m + n : m;
        const int ldg = basis_ld(m);
        // rows L and L + 64 of the last column (m <= 124) in registers first, so that no global read waits behind
        // the basis copy, which then runs while the last column is fitted
        cgdbl *an = A0g + (size_t)(n - 1) * ldg;
        const double a0 = (L < m) ? an[L] : 0.0, a1 = (L + kWave < m) ? an[L + kWave] : 0.0;
        double pd = 0.0;
        if (L < m) pd += a0 * a0;
        if (L + kWave < m) pd += a1 * a1;
        // pd (hence both loads) is formed before the copy is issued: otherwise its wait would cover the copy too
        asm volatile("" ::"v"(pd));
        load_basis(A0g);
        double den = wsum(pd);
        if (tikh) den += mu * mu;
        double px = 0.0;
        if (L < m) px += (a0 / den) * b0[L];
        if (L + kWave < m) px += (a1 / den) * b0[L + kWave];
        double xj = wsum(px);
        if (L < m) zz[L] = b0[L] - a0 * xj;
        if (L + kWave < m) zz[L + kWave] = b0[L + kWave] - a1 * xj;
        // Tikhonov rows below the copied ones are cleared while the copy is in flight; a pad row (odd m) after it
        if (tikh && L < n) {
            for (int i = ldg; i < (DMI_SLOT_ROWS ? min(M, rl) : M); ++i) Aat(i, L) = 0.0;
            if (DMI_SLOT_ROWS)
                for (int i = max(ldg, rl); i < M; ++i) G[(size_t)L * gld + (i - rl)] = 0.0;
        }
        basis_wait();
        if (tikh && L < n && ldg > m) Aat(m, L) = 0.0;
        __syncthreads();
        bool pos = false;
        if (L < n) {
            // w_j = sum_i A_ij zz_i over the rows in order, two rows per 16-byte load
            const ldbl *cj = A + L * lda;
            double wj = 0.0;
            int i = 0;
#pragma unroll 4
            for (; i + 2 <= m; i += 2) {
                const dv2 a = *(cldv2 *)(cj + i), z = *(cldv2 *)(zz + i);
                wj += a.x * z.x;
                wj += a.y * z.y;
            }
            if (i < m) wj += cj[i] * zz[i];
            if (L < n - 1) { wl = wj; pos = !(wj <= 0); }
            x[L] = 0.0;
            idxl = L;
        }
        unsigned long long bal = ballot(pos);
        if (L == n - 1) wl = (bal == 0) ? 1.0 : 0.0;
        lds_order();  // every lane has read zz before b (the same storage) takes the right-hand side
        for (int i = L; i < m; i += kWave) b[i] = b0[i];
        if (tikh)
            for (int i = m + L; i < M; i += kWave) b[i] = 0.0;
        flops += 2ull * m * n + 6ull * m;
        solves += 1;
        __syncthreads();
        DMI_PT1(kPhInit);
        {
            DMI_PT0();
            nnls_main(tikh, mu, nsetp, rnorm);
            DMI_PT1(kPhNnls);
        }
    }

    // ------------------------- Tikhonov cache (src/lsqnonneg.jl:392-444) -------------------------
    // A hit needs exactly equal mu; re-solving the same mu reproduces the same bits, so the cache is a pure
    // memo. Entries keep mu, rnorm and ||x||^2 in LDS and x in the global slot.
    __device__ __forceinline__ void cache_reset() {
        if (L < kTikhCache) cmu[L] = NAN;
        __syncthreads();
    }
    __device__ __forceinline__ int cache_solve(double mu) {
        int hit = -1, emp = -1;
        for (int i = 0; i < kTikhCache; ++i) {
            double c = rfl(cmu[i]);
            if (hit < 0 && !isnan(c) && c == mu) hit = i;
            if (emp < 0 && isnan(c)) emp = i;
        }
        if (hit >= 0) { ccur = hit; return hit; }
        int slot = emp >= 0 ? emp : ((ccur + 1) % kTikhCache);
        ccur = slot;
        int ns;
        double rn;
        solve(A0, true, mu, ns, rn);
        double v = (L < ns) ? x[idxl] : 0.0;
        double semi = wsum(v * v);
        if (L < n) cx[slot * n + L] = x[L];
        if (L == 0) { cmu[slot] = mu; crn[slot] = rn; csemi[slot] = semi; }
        __syncthreads();
        return slot;
    }
    __device__ __forceinline__ double cur_mu() { return rfl(cmu[ccur]); }
    __device__ __forceinline__ double cur_semi() { return rfl(csemi[ccur]); }
    __device__ __forceinline__ double cur_resnorm_sq() {
        double mu = cur_mu(), rn = rfl(crn[ccur]);
        return jl_max(rn * rn - (mu * mu) * cur_semi(), 0.0);
    }
    __device__ __forceinline__ void take_cur_solution() {
        if (L < n) xf[L] = cx[ccur * n + L];
        __syncthreads();
    }
    __device__ __forceinline__ double plain_solve_to_xf() {
        int ns;
        double rn;
        solve(A0, false, 0.0, ns, rn);
        if (L < n) xf[L] = x[L];
        __syncthreads();
        return rn;
    }
    __device__ __forceinline__ double plain_solve_res2(int &ns) {
        double rn;
        solve(A0, false, 0.0, ns, rn);
        return rn * rn;
    }
};

// --------------------------------- scalar optimizers (src/optimization.jl) ---------------------------------
// bracket_root_monotonic (:191-219)
template <class F> __device__ __forceinline__ void bracket_root_monotonic(F &&f, double a, double delta, double dilate, double mono,
                                                          int maxiters, double out[4]) {
    double fa = f(a);
    if (!isfinite(fa)) { out[0] = a; out[1] = a; out[2] = NAN; out[3] = NAN; return; }
    if (fa == 0) { out[0] = a; out[1] = a; out[2] = fa; out[3] = fa; return; }
    double sgn = jl_sign(mono) * jl_sign(fa);
    double b = a - sgn * delta, fb = f(b);
    if (!isfinite(fb)) { out[0] = a; out[1] = a; out[2] = fa; out[3] = fa; return; }
    if (fb == 0) { out[0] = b; out[1] = b; out[2] = fb; out[3] = fb; return; }
    delta *= dilate;
    int cnt = 0;
    while (fa * fb > 0 && cnt < maxiters) {
        a = b; fa = fb;
        b = a - sgn * delta;
        fb = f(b);
        if (!isfinite(fb)) { out[0] = a; out[1] = a; out[2] = fa; out[3] = fa; return; }
        if (fb == 0) { out[0] = b; out[1] = b; out[2] = fb; out[3] = fb; return; }
        delta *= dilate;
        cnt += 1;
    }
    if (a < b) { out[0] = a; out[1] = b; out[2] = fa; out[3] = fb; }
    else { out[0] = b; out[1] = a; out[2] = fb; out[3] = fa; }
}

// brent_root (:71-128)
template <class F> __device__ __forceinline__ void brent_root(F &&f, double x0, double x1, double fx0, double fx1, double xatol,
                                              double xrtol, double ftol, int maxiters, double &xo, double &fo) {
    if (fx0 == 0) { xo = x0; fo = fx0; return; }
    if (fx1 == 0) { xo = x1; fo = fx1; return; }
    double a = x0, b = x1, fa = fx0, fb = fx1;
    if (fabs(fa) < fabs(fb)) { double t = a; a = b; b = t; t = fa; fa = fb; fb = t; }
    double c = x0, d = x0, fc = fx0;
    bool mflag = true;
    for (int iter = 1; iter <= maxiters; ++iter) {
        if (fabs(b - a) <= 2 * (xatol + xrtol * fabs(b))) { xo = b; fo = fb; return; }
        double s = 0.0;
        s += a * fb * fc / (fa - fb) / (fa - fc);
        s += b * fa * fc / (fb - fa) / (fb - fc);
        s += c * fa * fb / (fc - fa) / (fc - fb);
        if (isnan(s) || isinf(s)) s = a - fa * (b - a) / (fb - fa);
        double u = (3 * a + b) / 4, v = b;
        if (u > v) { double t = u; u = v; v = t; }
        double tol = jl_max(xatol, xrtol * jl_max(jl_max(fabs(b), fabs(c)), fabs(d)));
        if (!(u < s && s < v) || (mflag && fabs(s - b) >= fabs(b - c) / 2) || (!mflag && fabs(s - b) >= fabs(b - c) / 2) ||
            (mflag && fabs(b - c) <= tol) || (!mflag && fabs(c - d) <= tol)) {
            s = (a + b) / 2;
            mflag = true;
        } else {
            mflag = false;
        }
        double fs = f(s);
        if (fs == 0) { xo = s; fo = fs; return; }
        if (isnan(fs) || isinf(fs)) { xo = b; fo = fb; return; }
        if (fabs(fs) <= ftol) { xo = s; fo = fs; return; }
        d = c; c = b; fc = fb;
        if (jl_sign(fa) * jl_sign(fs) < 0) { b = s; fb = fs; }
        else { a = s; fa = fs; }
        if (fabs(fa) < fabs(fb)) { double t = a; a = b; b = t; t = fa; fa = fb; fb = t; }
    }
    xo = b; fo = fb;
}

// brent_minimize (:319-413)
template <class F> __device__ __forceinline__ void brent_minimize(F &&f, double x1, double x2, double xrtol, double xatol, int maxiters,
                                                  double &xo, double &yo) {
    const double alpha = 2 - kGolden;
    double x = x1 + alpha * (x2 - x1);
    double y = f(x);
    double dx_old = 0.0, dx = 0.0, x_older = x, x_old = x, y_older = y, y_old = y;
    int iter = 0;
    while (iter < maxiters) {
        double p = 0.0, q = 0.0, xm = (x2 + x1) / 2, dx_tol = xatol + xrtol * fabs(x);
        if (fabs(x - xm) + (x2 - x1) / 2 <= 2 * dx_tol) break;
        iter += 1;
        if (fabs(dx_old) > dx_tol) {
            double r = (x - x_old) * (y - y_older);
            q = (x - x_older) * (y - y_old);
            p = (x - x_older) * q - (x - x_old) * r;
            q = 2 * (q - r);
            if (q > 0) p = -p;
            else q = -q;
        }
        if (fabs(p) < fabs(q * dx_old / 2) && p < q * (x2 - x) && p < q * (x - x1)) {
            dx_old = dx;
            dx = p / q;
            double x_tmp = x + dx;
            if ((x_tmp - x1) < 2 * dx_tol || (x2 - x_tmp) < 2 * dx_tol) dx = (x < xm) ? dx_tol : -dx_tol;
        } else {
            dx_old = (x < xm) ? x2 - x : x1 - x;
            dx = alpha * dx_old;
        }
        double x_new = (fabs(dx) >= dx_tol) ? x + dx : x + ((dx > 0) ? dx_tol : -dx_tol);
        double y_new = f(x_new);
        if (y_new < y) {
            if (x_new < x) x2 = x;
            else x1 = x;
            x_older = x_old; x_old = x; x = x_new;
            y_older = y_old; y_old = y; y = y_new;
        } else {
            if (x_new < x) x1 = x_new;
            else x2 = x_new;
            if (y_new <= y_old || x_old == x) {
                x_older = x_old; x_old = x_new;
                y_older = y_old; y_old = y_new;
            } else if (y_new <= y_older || x_older == x || x_older == x_old) {
                x_older = x_new;
                y_older = y_new;
            }
        }
    }
    xo = x;
    yo = y;
}

// ------------------------------------ legacy path (legacy = true) ------------------------------------
// Dierckx interpolating splines sampled on a 0.001 grid (src/splines.jl:312-315, 414-446, 494-500) and the chi2
// doubling search (src/lsqnonneg.jl:521-532, 595-636). Dierckx.jl wraps P. Dierckx's FITPACK: curfit with s = 0
// (the knots of fpcurf's interpolation case, Givens rotations fpgivs / fprota of each collocation row into a band
// triangle, back substitution fpback) and splev (the de Boor-Cox values of fpbspl). Each step keeps FITPACK's order
// of plain multiplies and adds (its Fortran is built for generic x86-64, without fma), so values match FITPACK's.

// Julia's Float64 range start:step:stop (base/twiceprecision.jl). With small exact rationals for start, step and
// stop (rat: continued fractions bounded by maxintfloat(Float32)), element i is (sn + i stn) / den to twice
// precision, i.e. the correctly rounded quotient of exact integers; otherwise start + i*step, product and sum each
// rounded.
struct JlRange {
    bool rational;
    long long sn, stn, den, len;
    double start, step;
    __device__ __forceinline__ double at(long long i) const {
        return rational ? (double)(sn + i * stn) / (double)den : start + (double)i * step;
    }
};
__device__ __forceinline__ void jl_rat(double x, long long &num, long long &den) {
    const double mf = 16777216.0;
    double y = x;
    long long a = 1, d = 1, b = 0, c = 0;
    while (fabs(y) <= mf) {
        const long long f = (long long)trunc(y);
        y -= (double)f;
        const long long a2 = f * a + c, b2 = f * b + d;
        c = a; a = a2;
        d = b; b = b2;
        if (!(llabs(a) <= 16777216ll && llabs(b) <= 16777216ll)) { num = c; den = d; return; }
        if ((double)a / (double)b == x) break;
        y = 1.0 / y;
    }
    num = a; den = b;
}
__device__ __forceinline__ bool jl_isbetween(double a, double x, double b) {
    return (a <= x && x <= b) || (b <= x && x <= a);
}
__device__ JlRange jl_range(double start, double step, double stop) {
    JlRange r{false, 0, 0, 1, 0, start, step};
    long long step_n, step_d;
    jl_rat(step, step_n, step_d);
    if (step_d != 0 && (double)step_n / (double)step_d == step) {
        long long start_n, start_d, stop_n, stop_d;
        jl_rat(start, start_n, start_d);
        jl_rat(stop, stop_n, stop_d);
        if (start_d != 0 && stop_d != 0 && (double)start_n / (double)start_d == start &&
            (double)stop_n / (double)stop_d == stop) {
            long long g = llabs(start_d), h = llabs(step_d);
            while (h) { const long long t = g % h; g = h; h = t; }
            const long long den = llabs(start_d / g * step_d);
            const double mi = 9007199254740992.0;
            if (den != 0 && fabs(start * (double)den) <= mi && fabs(step * (double)den) <= mi && den % start_d == 0 &&
                den % step_d == 0) {
                const long long sn = llrint(start * (double)den), stn = llrint(step * (double)den);
                // Int64 arithmetic wraps in Julia; unsigned arithmetic wraps the same way here
                typedef unsigned long long u64;
                const long long num = (long long)((u64)den * (u64)stop_n - (u64)stop_d * (u64)sn + (u64)stn * (u64)stop_d);
                long long len = num / (long long)((u64)stn * (u64)stop_d);
                if (len < 0) len = 0;
                if (jl_isbetween(start, start + (double)(len - 1) * step, stop + step / 2) &&
                    !jl_isbetween(start, start + (double)len * step, stop)) {
                    r.rational = true; r.sn = sn; r.stn = stn; r.den = den; r.len = len;
                    return r;
                }
            }
        }
    }
    const double lf = (stop - start) / step;
    long long len;
    if (lf < 0) len = 0;
    else if (lf == 0) len = 1;
    else {
        len = llrint(lf) + 1;
        const double stop2 = start + (double)(len - 1) * step;
        len -= (long long)(start < stop && stop < stop2) + (long long)(start > stop && stop > stop2);
    }
    r.len = len;
    return r;
}

// fpbspl: h[1..k+1] = the nonzero B-splines of degree k <= 3 at x, t(l) <= x < t(l + 1); t1 is 1-based
__device__ __forceinline__ void fpbspl(const ldbl *t1, int k, double x, int l, double h[5]) {
    double hh[3];
    h[1] = 1.0;
#pragma unroll
    for (int j = 1; j <= 3; ++j) {
        if (j > k) break;
#pragma unroll
        for (int i = 1; i <= j; ++i) hh[i - 1] = h[i];
        h[1] = 0.0;
#pragma unroll
        for (int i = 1; i <= j; ++i) {
            const int li = l + i, lj = li - j;
            const double tli = t1[li], tlj = t1[lj];
            if (tli == tlj) { h[i + 1] = 0.0; continue; }
            const double f = hh[i - 1] / (tli - tlj);
            h[i] = h[i] + f * (tli - x);
            h[i + 1] = f * (x - tlj);
        }
    }
}
__device__ __forceinline__ void fpgivs(double piv, double &ww, double &cs, double &sn) {
    const double store = fabs(piv);
    double dd;
    if (store >= ww) dd = store * sqrt(1.0 + (ww / piv) * (ww / piv));
    else dd = ww * sqrt(1.0 + (piv / ww) * (piv / ww));
    cs = ww / dd;
    sn = piv / dd;
    ww = dd;
}
__device__ __forceinline__ void fprota(double cs, double sn, double &a, double &b) {
    const double s1 = a, s2 = b;
    b = cs * s2 + sn * s1;
    a = cs * s1 - sn * s2;
}

// Interpolating spline of degree k = min(3, np - 1) through the np (2 <= np <= 64) points held one per lane (lane i:
// point i + 1, x increasing). Knots go to W.spt, coefficients to W.spc; returns the knot count n = np + k + 1.
// Lane j - 1 holds row j of the band triangle (4 entries) and of the rotated right-hand side; every rotation reads
// its row with readlane, computes in all lanes alike, and the owner keeps the result.
__device__ int spline_fit(Wave &W, double xv, double yv, int np, int &k_out) {
    const int L = W.L;
    const int k = min(3, np - 1), k1 = k + 1, n = np + k1, nk1 = np, k3 = k / 2;
    ldbl *t1 = W.spt - 1;
    // knots: t(1..k1) = x(1), t(n-k..n) = x(np); t(i) for k2 <= i <= n - k1: x(j) (odd k) or (x(j) + x(j-1)) / 2
    // (even k), j = i - k + k3
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int i = L + 1 + r * kWave;
        const int j = min(max(i - k + k3, 2), np);
        const double xj = __shfl(xv, j - 1, kWave), xj1 = __shfl(xv, j - 2, kWave);
        const double x1 = readlane(xv, 0), xe = readlane(xv, np - 1);
        double v = (k3 * 2 == k) ? (xj + xj1) * 0.5 : xj;
        if (i <= k1) v = x1;
        if (i > n - k1) v = xe;
        if (i <= n) t1[i] = v;
    }
    lds_order();
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0, zr = 0.0;
    int l = k1;
    for (int it = 1; it <= np; ++it) {
        const double xi = readlane(xv, it - 1);
        double yi = readlane(yv, it - 1);
        while (!(xi < rfl(t1[l + 1]) || l == nk1)) ++l;
        double h[5];
        fpbspl(t1, k, xi, l, h);
        int jj = l - k1;
#pragma unroll
        for (int ii = 1; ii <= 4; ++ii) {
            if (ii > k1) break;
            ++jj;
            const double piv = h[ii];
            if (piv == 0.0) continue;
            double r[4] = {readlane(a0, jj - 1), readlane(a1, jj - 1), readlane(a2, jj - 1), readlane(a3, jj - 1)};
            double zj = readlane(zr, jj - 1), cs, sn;
            fpgivs(piv, r[0], cs, sn);
            fprota(cs, sn, yi, zj);
#pragma unroll
            for (int i1 = ii + 1; i1 <= 4; ++i1) {
                if (i1 > k1) break;
                fprota(cs, sn, h[i1], r[i1 - ii]);
            }
            if (L == jj - 1) { a0 = r[0]; a1 = r[1]; a2 = r[2]; a3 = r[3]; zr = zj; }
        }
    }
    // fpback: c(i) = (z(i) - sum_{q=1..min(k, nk1-i)} c(i+q) a(i, q+1)) / a(i, 1), from i = nk1 down
    ldbl *c1p = W.spc - 1;
    double c1 = 0.0, c2 = 0.0, c3 = 0.0;
    for (int i = nk1; i >= 1; --i) {
        double store = readlane(zr, i - 1);
        const double r0 = readlane(a0, i - 1), r1 = readlane(a1, i - 1), r2 = readlane(a2, i - 1), r3 = readlane(a3, i - 1);
        const int i1 = min(k, nk1 - i);
        if (i1 >= 1) store = store - c1 * r1;
        if (i1 >= 2) store = store - c2 * r2;
        if (i1 >= 3) store = store - c3 * r3;
        const double ci = store / r0;
        if (L == 0) c1p[i] = ci;
        c3 = c2; c2 = c1; c1 = ci;
    }
    lds_order();
    W.flops += 40ull * np * k1 + 8ull * np;
    k_out = k;
    return n;
}

// spline_opt_legacy (src/splines.jl:419-431; root = false) / spline_root_legacy (:435-446; root = true, the value
// |s(x) - v|) over knots[1]:0.001:knots[end] (knots = t(k+1..n-k)): the first sample of least value wins, as the
// reference's scan with strict '<' keeps it (a NaN first value is never replaced). Lane L scans a contiguous block
// of samples; the interval search starts where the lane's previous sample left it (the samples increase, so this
// is the interval splev finds from t(k+1)).
__device__ void spline_scan(Wave &W, int k, int n, bool root, double v, double &xo, double &yo) {
    const int L = W.L, k1 = k + 1, nk1 = n - k1;
    const ldbl *t1 = W.spt - 1, *c1p = W.spc - 1;
    const JlRange r = jl_range(rfl(t1[k1]), 0.001, rfl(t1[n - k]));
    const long long len = r.len;
    if (!(len >= 1 && len <= kSplineMaxSamples)) {  // only non-finite knots get here; the voxel gets NaN
        W.failed = true;
        xo = NAN; yo = NAN;
        return;
    }
    const long long per = (len + kWave - 1) / kWave, i0 = (long long)L * per, i1 = min(len, i0 + per);
    double best = INFINITY, first = NAN;
    long long bi = -1;
    int l = k1;
    for (long long i = i0; i < i1; ++i) {
        const double xi = r.at(i);
        while (!(xi < t1[l + 1] || l == nk1)) ++l;
        double h[5];
        fpbspl(t1, k, xi, l, h);
        double sp = 0.0;
        int ll = l - k1;
#pragma unroll
        for (int j = 1; j <= 4; ++j) {
            if (j > k1) break;
            ++ll;
            sp = sp + c1p[ll] * h[j];
        }
        const double y = root ? fabs(sp - v) : sp;
        if (i == 0) first = y;
        if (!isnan(y) && (bi < 0 || y < best)) { best = y; bi = i; }
    }
    W.flops += (unsigned long long)len * (unsigned long long)(6 * k * (k + 1) / 2 + 2 * k1 + 2);
    const double y0 = readlane(first, 0);
    int ord = bi >= 0 ? L : INT_MAX;
    if (bi < 0) best = INFINITY;
    wargmin(best, ord);
    if (isnan(y0) || ord == INT_MAX) { xo = r.at(0); yo = y0; return; }
    const long long b = (long long)__shfl((long long)bi, ord, kWave);
    xo = r.at(b);
    yo = best;
}

// ------------------------------------ regularization (src/lsqnonneg.jl) ------------------------------------
// lsqnonneg_chi2! with method :brent (:504-593)
// (legacy: method :legacy, :521-532, with chi2_search_from_minimum)
__device__ __forceinline__ void reg_chi2(Wave &W, double chi2_target, bool legacy, double &mu_out, double &chi2_out) {
    int ns;
    double res2_min = W.plain_solve_res2(ns);
    if (res2_min == 0 || ns == 0) {
        if (W.L < W.n) W.xf[W.L] = W.x[W.L];
        __syncthreads();
        mu_out = 0.0; chi2_out = 1.0;
        return;
    }
    if (W.L < W.n) W.xf[W.L] = W.x[W.L];  // x_unreg, kept for the no-root branch
    __syncthreads();
    const double target = chi2_target * res2_min;
    W.cache_reset();
    if (legacy) {
        // chi2_search_from_minimum (:595-636): lane i holds the (mu, res2) pair i, starting from (0, res2_min)
        double mu_l = (W.L == 0) ? 0.0 : NAN, r2_l = (W.L == 0) ? res2_min : NAN;
        int np = 1;
        double munew = 1e-3;
        for (int it = 0;; ++it) {
            W.cache_solve(munew);
            const double r2 = W.cur_resnorm_sq();
            if (W.L == np) { mu_l = munew; r2_l = r2; }
            np += 1;
            if (r2 >= chi2_target * res2_min) break;
            if (it == kLegacyMaxDoublings) { W.failed = true; return; }  // the reference keeps doubling without end
            munew *= 2.0;
        }
        int k;
        const int nk = spline_fit(W, mu_l, r2_l, np, k);
        double mu, unused;
        spline_scan(W, k, nk, true, chi2_target * res2_min, mu, unused);
        double res2;
        if (mu == 0) {
            res2 = res2_min;  // x_final = x_unreg, already in xf
        } else {
            W.cache_solve(mu);
            res2 = W.cur_resnorm_sq();
            W.take_cur_solution();
        }
        mu_out = mu;
        chi2_out = res2 / res2_min;
        return;
    }
    auto f = [&](double logmu) {
        W.cache_solve(exp(logmu));
        return (W.cur_resnorm_sq() - target) / target;
    };
    double br[4];
    bracket_root_monotonic(f, -4.0, 1.0, 1.5, 1.0, 6, br);
    double a = br[0], b = br[1], fa = br[2], fb = br[3], xf, ff;
    if (fa * fb < 0) brent_root(f, a, b, fa, fb, 0.0, 0.0, 1e-3 * (chi2_target - 1), 100, xf, ff);
    else if (!isfinite(fa)) { xf = b; ff = fb; }
    else if (!isfinite(fb)) { xf = a; ff = fa; }
    else if (fabs(fa) < fabs(fb)) { xf = a; ff = fa; }
    else { xf = b; ff = fb; }
    double mu_final, res2_final;
    if (isfinite(ff)) {
        mu_final = exp(xf);
        res2_final = target * (1 + ff);
        W.cache_solve(mu_final);
        W.take_cur_solution();
    } else {
        mu_final = 0.0;
        res2_final = 1.0;
    }
    mu_out = mu_final;
    chi2_out = res2_final / res2_min;
}

// lsqnonneg_mdp! (:700-747)
__device__ __forceinline__ void reg_mdp(Wave &W, double delta, double &mu_out, double &chi2_out) {
    int ns;
    double res2_min = W.plain_solve_res2(ns);
    if (W.L < W.n) W.xf[W.L] = W.x[W.L];
    __syncthreads();
    if (delta <= sqrt(res2_min)) { mu_out = 0.0; chi2_out = 1.0; return; }
    double p = 0.0;
    for (int i = W.L; i < W.m; i += kWave) p += W.b0[i] * W.b0[i];
    double res2_max = wsum(p);
    if (delta >= sqrt(res2_max)) {
        if (W.L < W.n) W.xf[W.L] = 0.0;
        __syncthreads();
        mu_out = INFINITY;
        chi2_out = res2_max / res2_min;
        return;
    }
    W.cache_reset();
    const double d2 = delta * delta;
    auto f = [&](double logmu) {
        W.cache_solve(exp(logmu));
        return W.cur_resnorm_sq() - d2;
    };
    double br[4];
    bracket_root_monotonic(f, -4.0, 1.0, 1.5, 1.0, 6, br);
    double a = br[0], b = br[1], fa = br[2], fb = br[3], xf, ff;
    if (fa * fb < 0) brent_root(f, a, b, fa, fb, 0.0, 0.0, 1e-3 * d2, 100, xf, ff);
    else if (!isfinite(fa)) { xf = b; ff = fb; }
    else if (!isfinite(fb)) { xf = a; ff = fa; }
    else if (fabs(fa) < fabs(fb)) { xf = a; ff = fa; }
    else { xf = b; ff = fb; }
    double mu_final, res2_final;
    if (isfinite(ff)) {
        mu_final = exp(xf);
        res2_final = d2 + ff;
        W.cache_solve(mu_final);
        W.take_cur_solution();
    } else {
        mu_final = 0.0;
        res2_final = 1.0;
    }
    mu_out = mu_final;
    chi2_out = res2_final / res2_min;
}

// Squared singular values of the m x n basis for gcv_dof (src/lsqnonneg.jl:1321-1329). The reference
// calls LAPACK dgesdd (src/utils.jl:96-134); here: Gram matrix of the smaller side, Householder
// tridiagonalization, then Sturm-sequence bisection, one eigenvalue per lane. Absolute accuracy is
// eps * ||A||^2, far below the mu^2 >= exp(-16) scale at which gcv_dof uses them.
__device__ __forceinline__ void gram_eigs(Wave &W, double &lam_lane, int &N) {
    const int m = W.m, n = W.n, L = W.L, lda = W.lda;
    DMI_PT0();
    W.load_basis(W.A0);
    W.basis_wait();
    __syncthreads();
    N = min(m, n);
    // N x N Gram matrix at rows [m, m + N) of columns [0, N) (lda >= m + n leaves the room), built with FP64
    // MFMA 16x16x4 blocks: G = A^T A when m >= n, A A^T otherwise. For block (R, C) and k-step s, lane l supplies
    // X[l % 16][s + l / 16] and Y[s + l / 16][l % 16]; the result lane l, register r is D[4r + l / 16][l % 16].
    // Only blocks with R <= C are formed and each value is stored at (row, col) and (col, row), so G is exactly
    // symmetric.
    auto G = [&](int r, int c) -> ldbl & { return W.A[c * lda + m + r]; };
    {
        typedef double d4 __attribute__((ext_vector_type(4)));
        const int NB = (N + 15) / 16, K = (m >= n) ? m : n;
        const int lr = L & 15, lk = L >> 4;
        for (int R = 0; R < NB; ++R)
            for (int C = R; C < NB; ++C) {
                d4 acc = {0.0, 0.0, 0.0, 0.0};
                const int ra = R * 16 + lr, ca = C * 16 + lr;
                for (int s0 = 0; s0 < K; s0 += 4) {
                    const int k = s0 + lk;
                    double a = 0.0, bb = 0.0;
                    if (k < K) {
                        if (m >= n) {
                            if (ra < N) a = W.Aat(k, ra);
                            if (ca < N) bb = W.Aat(k, ca);
                        } else {
                            if (ra < N) a = W.Aat(ra, k);
                            if (ca < N) bb = W.Aat(ca, k);
                        }
                    }
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc, 0, 0, 0);
                }
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    const int row = R * 16 + 4 * rr + lk, col = C * 16 + lr;
                    if (row < N && col < N && (R < C || row <= col)) {
                        G(row, col) = acc[rr];
                        G(col, row) = acc[rr];
                    }
                }
            }
    }
    W.flops += 2ull * N * N * (m >= n ? m : n);
    lds_order();
    DMI_LAP(kPhGram);
    // the tridiagonal in b (2N <= m + n entries), the reflector in x (N <= n); the vector q of the rank-2 update
    // stays in the lanes' registers and is broadcast with readlane
    ldbl *dv = W.b, *ev = W.b + N, *vv = W.x;
    for (int k = 0; k < N - 2; ++k) {
        double xr = (L > k && L < N) ? G(L, k) : 0.0;
        double nrm = sqrt(wsum(xr * xr));
        double x1 = rfl(G(k + 1, k));
        if (nrm == 0.0) {
            if (L == 0) ev[k] = x1;
            lds_order();
            continue;
        }
        double ah = -copysign(nrm, x1);
        double vr = xr;
        if (L == k + 1) vr = xr - ah;
        double vn = sqrt(wsum(vr * vr));
        if (vn == 0.0) {
            if (L == 0) ev[k] = x1;
            lds_order();
            continue;
        }
        vr /= vn;
        if (L < N) vv[L] = vr;
        if (L == 0) ev[k] = ah;
        lds_order();
        // p = G v over the trailing rows: four partial sums, so the FMA chain is a quarter as long
        double pr = 0.0;
        if (L > k && L < N) {
            double p0 = 0.0, p1 = 0.0, p2 = 0.0, p3 = 0.0;
            int c = k + 1;
            // the step's loads are issued together (the scheduler would otherwise trade them for registers)
            for (; c + 4 <= N; c += 4) {
                double g[4], v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) { g[r] = G(L, c + r); v[r] = vv[c + r]; }
                __builtin_amdgcn_sched_barrier(0);
                p0 = fma(g[0], v[0], p0);
                p1 = fma(g[1], v[1], p1);
                p2 = fma(g[2], v[2], p2);
                p3 = fma(g[3], v[3], p3);
            }
            for (; c < N; ++c) p0 = fma(G(L, c), vv[c], p0);
            pr = (p0 + p1) + (p2 + p3);
        }
        double Kq = wsum(vr * pr);
        double qr = pr - Kq * vr;
        lds_order();
        // rank-2 update, four columns per step with every load of the step ahead of its stores (the loads of vv
        // could alias the stores to G as far as the compiler knows, so a plain loop would serialize)
        if (L > k && L < N) {
            int c = k + 1;
            for (; c + 4 <= N; c += 4) {
                double g[4], q[4], v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) { g[r] = G(L, c + r); q[r] = readlane(qr, c + r); v[r] = vv[c + r]; }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int r = 0; r < 4; ++r) G(L, c + r) = g[r] - 2.0 * (vr * q[r] + qr * v[r]);
            }
            for (; c < N; ++c) G(L, c) -= 2.0 * (vr * readlane(qr, c) + qr * vv[c]);
        }
        W.flops += 8ull * (N - k) * (N - k);
        lds_order();
    }
    if (L < N) dv[L] = G(L, L);
    if (N >= 2 && L == 0) ev[N - 2] = G(N - 1, N - 2);
    lds_order();
    DMI_LAP(kPhTridiag);
    // Gershgorin interval
    double lo = INFINITY, hi = -INFINITY;
    if (L < N) {
        double r = 0.0;
        if (L > 0) r += fabs(ev[L - 1]);
        if (L < N - 1) r += fabs(ev[L]);
        lo = dv[L] - r;
        hi = dv[L] + r;
    }
    lo = wmind(lo);
    hi = wmaxd(hi);
    double gnorm = fmax(fabs(lo), fabs(hi));
    const double eg = kEps * gnorm;
    const int t = L;
    // One Sturm sequence at x (the same pivots as the sweeps below): count(x) = #eigenvalues < x, and
    // g(x) = sum_i q_i' / q_i = sum_j 1 / (x - lambda_j), the logarithmic derivative of det(T - x I), from the
    // pivots' derivatives q_i' = -1 + e_{i-1}^2 q_{i-1}' / q_{i-1}^2.
    // The diagonal, e^2 and the zero-pivot guard of step i come from lane i's registers by readlane (uniform
    // values, no LDS traffic in the chain).
    const double dL = (L < N) ? dv[L] : 0.0, eL = (L < N - 1) ? ev[L] : 0.0;
    const double e2L = eL * eL, gL = kEps * (fabs(eL) + kEps);
    auto chain = [&](double x, int &count, double &g) {
        double q = readlane(dL, 0) - x, qp = -1.0, gs = 0.0;
        int c = q < 0;
        for (int i = 1; i < N; ++i) {
            const double d = readlane(dL, i), e2 = readlane(e2L, i - 1), guard = readlane(gL, i - 1);
            q = (q == 0.0) ? guard : q;
            double r = __builtin_amdgcn_rcp(q);
            r = fma(r, fma(-q, r, 1.0), r);
            r = fma(r, fma(-q, r, 1.0), r);
            gs = fma(qp, r, gs);
            qp = fma(e2 * qp, r * r, -1.0);
            q = fma(-e2, r, d - x);
            c += q < 0;
        }
        count = c;
        g = gs + qp / ((q == 0.0) ? eg : q);
    };
    // Initial brackets: lane l counts the eigenvalues below threshold x_l (x_0 above the spectrum, x_1..x_61
    // log-spaced from gnorm down to eps * gnorm, x_62 = 0, x_63 = -4 eps * gnorm); eigenvalue t lies between the two
    // thresholds where the count passes t, a relative width of 0.8 however small it is.
    const double hi1 = hi + 4.0 * eg + 4.0 * kEps * fabs(hi), lo1 = lo - 4.0 * eg - 4.0 * kEps * fabs(lo);
    double xl;
    if (L == 0) xl = hi1;
    else if (L <= 61) xl = gnorm * exp2(-52.0 * (double)L / 61.0);
    else if (L == 62) xl = 0.0;
    else xl = -4.0 * eg;
    int cl;
    double gl;
    chain(xl, cl, gl);
    unsigned long long mymask = 0ull;
    for (int tt = 0; tt < N; ++tt) {
        const unsigned long long msk = ballot(cl > tt);
        if (L == tt) mymask = msk;
    }
    const int Lt = __popcll(mymask);
    const bool contiguous = mymask == ((Lt == 64) ? ~0ull : ((1ull << Lt) - 1ull));
    const double xb = __shfl(xl, max(Lt - 1, 0), kWave), xa = __shfl(xl, min(Lt, 63), kWave);
    double a = lo1, b = hi1;
    if (contiguous && Lt > 0) {
        b = xb;
        a = (Lt < 64) ? xa : lo1;
    }
    // Two Sturm sequences per step: at x1 with the pivots' first and second derivatives (g = sum 1 / (x - lambda_j)
    // and g' = -sum 1 / (x - lambda_j)^2), and at x2 for its count only.
    auto chain2 = [&](double x1, double x2, int &c1, double &g, double &gd, int &c2) {
        const double d0 = readlane(dL, 0);
        double q1 = d0 - x1, q2 = d0 - x2, qp = -1.0, qpp = 0.0, gs = 0.0, hs = 0.0;
        int k1 = q1 < 0, k2 = q2 < 0;
        for (int i = 1; i < N; ++i) {
            const double d = readlane(dL, i), e2 = readlane(e2L, i - 1), guard = readlane(gL, i - 1);
            q1 = (q1 == 0.0) ? guard : q1;
            q2 = (q2 == 0.0) ? guard : q2;
            double r1 = __builtin_amdgcn_rcp(q1), r2 = __builtin_amdgcn_rcp(q2);
            r1 = fma(r1, fma(-q1, r1, 1.0), r1);
            r2 = fma(r2, fma(-q2, r2, 1.0), r2);
            r1 = fma(r1, fma(-q1, r1, 1.0), r1);
            r2 = fma(r2, fma(-q2, r2, 1.0), r2);
            const double r1s = r1 * r1, qp2 = qp * qp;
            gs = fma(qp, r1, gs);                                // q'/q
            hs = fma(fma(qpp, q1, -qp2), r1s, hs);               // (q'' q - q'^2) / q^2
            qpp = e2 * fma(qpp, q1, -2.0 * qp2) * (r1s * r1);   // q_i'' = e^2 (q'' q - 2 q'^2) / q^3
            qp = fma(e2 * qp, r1s, -1.0);                        // q_i' = -1 + e^2 q' / q^2
            q1 = fma(-e2, r1, d - x1);
            q2 = fma(-e2, r2, d - x2);
            k1 += q1 < 0;
            k2 += q2 < 0;
        }
        const double qz = (q1 == 0.0) ? eg : q1, rz = 1.0 / qz;
        c1 = k1;
        c2 = k2;
        g = gs + qp * rz;
        gd = hs + fma(qpp, qz, -(qp * qp)) * (rz * rz);
    };
    auto mid = [](double lo_, double hi_) { return (lo_ > 0.0 && hi_ > 2.0 * lo_) ? sqrt(lo_ * hi_) : 0.5 * (lo_ + hi_); };
    // Laguerre's iteration for the real-rooted det(T - x I) (cubic convergence to the root next to x, on either
    // side) inside the bracket, with a midpoint count in the same step: each step moves the bracket ends by both
    // counts, and a Laguerre point that leaves the bracket is replaced by the midpoint (geometric for a positive
    // bracket wider than 2x). Once a step is below the tolerance, the next step counts at the root +- tol, which
    // closes the bracket around it. Lanes still open after 30 steps finish in the four-way sweeps below.
    bool conv = false;
    double x = mid(a, b), xn = x;
    int itn = 0;
    for (; itn < 30; ++itn) {
        bool open = t < N;
        double p1 = 0.0, p2 = 0.0;
        if (open) {
            const double tol = fmax(2.0 * kEps * fmax(fabs(a), fabs(b)), eg), mm = mid(a, b);
            open = (b - a) > tol && mm > a && mm < b;
            p1 = conv ? xn - tol : x;
            p2 = conv ? xn + tol : mm;
            if (!(p1 > a && p1 < b)) p1 = mm;
            if (!(p2 > a && p2 < b)) p2 = mm;
        }
        if (ballot(open) == 0) break;
        int c1, c2;
        double g, gd;
        chain2(p1, p2, c1, g, gd, c2);
        if (open) {
            if (c1 > t) b = fmin(b, p1);
            else a = fmax(a, p1);
            if (c2 > t) b = fmin(b, p2);
            else a = fmax(a, p2);
            const double tol = fmax(2.0 * kEps * fmax(fabs(a), fabs(b)), eg);
            const double disc = (double)(N - 1) * ((double)N * (-gd) - g * g);
            const double sq = sqrt(fmax(disc, 0.0)), den = (g >= 0.0) ? g + sq : g - sq;
            xn = p1 - (double)N / den;
            conv = isfinite(xn) && fabs(xn - p1) <= tol;
            x = (isfinite(xn) && xn > a && xn < b) ? xn : mid(a, b);
        }
    }
    W.flops += 5ull * N * N * (1 + 2 * itn);  // the threshold sequence and two sequences per step
    DMI_LAP(kPhEigBracket);
    DMI_EIG_STEPS(itn);
    // Each lane brackets eigenvalue t with count(a) <= t < count(b), count(x) = #eigenvalues < x, and splits the
    // bracket in four per sweep: three Sturm sequences (LDL^T pivots q_i = (d_i - x) - e_{i-1}^2 / q_{i-1}) run
    // interleaved. Each quotient uses the hardware reciprocal refined by two Newton steps (within an ulp or two of
    // 1/q; the reciprocal alone is too coarse, since a pivot near zero amplifies a quotient's error into later
    // pivots' signs). The diagonal and off-diagonal are loaded one step ahead.
    int it = 0;
    for (; it < 100; ++it) {
        const double h = b - a;
        const double x1 = a + 0.25 * h, x2 = a + 0.5 * h, x3 = a + 0.75 * h;
        bool done = !(t < N) || h <= fmax(2.0 * kEps * fmax(fabs(a), fabs(b)), kEps * gnorm) || x2 <= a || x2 >= b;
        if (ballot(!done) == 0) break;
        if (!done) {
            volatile ldbl *dvv = dv, *evv = ev;
            const double d0 = dvv[0];
            double q1 = d0 - x1, q2 = d0 - x2, q3 = d0 - x3;
            int c1 = q1 < 0, c2 = q2 < 0, c3 = q3 < 0;
            double dn = dvv[N > 1 ? 1 : 0], en = evv[0];
            for (int i = 1; i < N; ++i) {
                const double d = dn, e = en;
                dn = dvv[i + 1 < N ? i + 1 : i];
                en = evv[i];
                const double e2 = e * e, guard = kEps * (fabs(e) + kEps);
                q1 = (q1 == 0.0) ? guard : q1;
                q2 = (q2 == 0.0) ? guard : q2;
                q3 = (q3 == 0.0) ? guard : q3;
                double r1 = __builtin_amdgcn_rcp(q1), r2 = __builtin_amdgcn_rcp(q2), r3 = __builtin_amdgcn_rcp(q3);
                r1 = fma(r1, fma(-q1, r1, 1.0), r1);
                r2 = fma(r2, fma(-q2, r2, 1.0), r2);
                r3 = fma(r3, fma(-q3, r3, 1.0), r3);
                r1 = fma(r1, fma(-q1, r1, 1.0), r1);
                r2 = fma(r2, fma(-q2, r2, 1.0), r2);
                r3 = fma(r3, fma(-q3, r3, 1.0), r3);
                q1 = fma(-e2, r1, d - x1);
                q2 = fma(-e2, r2, d - x2);
                q3 = fma(-e2, r3, d - x3);
                c1 += q1 < 0;
                c2 += q2 < 0;
                c3 += q3 < 0;
            }
            if (c1 > t) b = x1;
            else if (c2 > t) { a = x1; b = x2; }
            else if (c3 > t) { a = x2; b = x3; }
            else a = x3;
        }
    }
    W.flops += 10ull * N * N * it;  // two bisection steps per sweep, 5N^2 each
    lam_lane = (t < N) ? fmax(0.5 * (a + b), 0.0) : 0.0;
    __syncthreads();
    DMI_LAP(kPhSturm);
}

// lsqnonneg_gcv! with method :brent (:1136-1205)
__device__ __forceinline__ void reg_gcv(Wave &W, double &mu_out, double &chi2_out) {
    double lam;
    int N;
    {
        DMI_PT0();
        gram_eigs(W, lam, N);
        DMI_PT1(kPhGramEigs);
    }
    const int m = W.m, n = W.n;
    const double low = (kEps * kEps) / m;
    W.cache_reset();
    auto f = [&](double logmu) {
        double mu = exp(logmu);
        W.cache_solve(mu);
        double res2 = W.cur_resnorm_sq();
        double l2 = mu * mu;
        double term = (W.L < N) ? l2 / (lam + l2) : 0.0;
        double dof = (double)(m - n > 0 ? m - n : 0) + wsum(term);
        double g = res2 / (dof * dof);
        return log(jl_max(g, low));
    };
    double xf, yf;
    brent_minimize(f, -8.0, 2.0, 0.0, 1e-4, 20, xf, yf);
    double mu_final = exp(xf);
    W.cache_solve(mu_final);
    W.take_cur_solution();
    double res2 = W.cur_resnorm_sq();
    int ns;
    double res2u = W.plain_solve_res2(ns);
    mu_out = mu_final;
    chi2_out = res2 / res2u;
}

// lsqnonneg_lcurve! (:812-840) with lcurve_corner (:872-972) and the isapprox-keyed GrowableCache /
// CachedFunction / mapfindmax semantics of src/utils.jl:140-254. Caches live in global scratch; scans are
// lane-parallel reductions that reproduce the reference's sequential first-match / first-max rules.
//
// Sizes. Every move (move_left :933-939, move_right :941-946) shrinks the bracket from x4 - x1 to x3 - x1 = x4 - x2,
// a factor 1/phi, and the loop stops once the bracket is narrower than xtol: from [-8, 2] with xtol = 1e-4 that is
// at most ceil(log(1e5) / log(phi)) = 24 moves, each evaluating one new point, so the function and point caches hold
// at most 4 + 24 = 28 entries (kLcurveCap = 32). The same shrinking makes the backtracking of :892-900 keep the
// current state: it may only switch to a cached state no wider than the current one, and every earlier state is
// wider, so the state cache is not kept. A voxel whose caches would still overflow gets no outputs (they stay NaN)
// and is counted; the rest of the volume is unaffected.
struct LCurve {
    Wave &W;
    gdbl *fcx, *fcP0, *fcP1;
    gdbl *pcx, *pcP0, *pcP1, *pcC;
    int nfc, npc;
    bool overflow;

    __device__ __forceinline__ LCurve(Wave &w) : W(w), nfc(0), npc(0), overflow(false) {
        gdbl *p = w.lc;
        fcx = p; fcP0 = p + kLcurveCap; fcP1 = p + 2 * kLcurveCap;
        p += 3 * kLcurveCap;
        pcx = p; pcP0 = p + kLcurveCap; pcP1 = p + 2 * kLcurveCap; pcC = p + 3 * kLcurveCap;
    }
    __device__ __forceinline__ int find(const gdbl *keys, int cnt, double x) {
        int first = INT_MAX;
        for (int base = 0; base < cnt; base += kWave) {
            int k = base + W.L;
            bool hit = k < cnt && jl_isapprox(x, keys[k]);
            int f = first_set(ballot(hit));
            if (f >= 0) { first = base + f; break; }
        }
        return first == INT_MAX ? -1 : first;
    }
    __device__ __forceinline__ void eval(double x, double &P0, double &P1) {
        int i = find(fcx, nfc, x);
        if (i >= 0) { P0 = rfl(fcP0[i]); P1 = rfl(fcP1[i]); return; }
        W.cache_solve(exp(x));
        P0 = log(W.cur_resnorm_sq());
        P1 = log(W.cur_semi());
        if (nfc >= kLcurveCap) { overflow = true; return; }
        if (W.L == 0) { fcx[nfc] = x; fcP0[nfc] = P0; fcP1[nfc] = P1; }
        nfc++;
        __syncthreads();
    }
    __device__ __forceinline__ void pc_set(double x, double P0, double P1, double C) {
        int i = find(pcx, npc, x);
        if (i < 0) {
            if (npc >= kLcurveCap) { overflow = true; return; }
            i = npc++;
            if (W.L == 0) pcx[i] = x;
        }
        if (W.L == 0) { pcP0[i] = P0; pcP1[i] = P1; pcC[i] = C; }
        __syncthreads();
    }
    __device__ __forceinline__ int pc_argmax() {
        // findmax with isless (NaN largest), first occurrence
        double bc = -INFINITY;
        int bi = INT_MAX;
        bool bnan = false;
        for (int base = 0; base < npc; base += kWave) {
            int k = base + W.L;
            if (k < npc) {
                double c = pcC[k];
                bool cn = isnan(c);
                bool better = (bi == INT_MAX) || (cn && !bnan) || (!cn && !bnan && c > bc);
                if (better) { bc = c; bi = k; bnan = cn; }
            }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            double oc = __shfl_xor(bc, o, kWave);
            int oi = __shfl_xor(bi, o, kWave);
            bool on = __shfl_xor((int)bnan, o, kWave) != 0;
            bool take;
            if (oi == INT_MAX) take = false;
            else if (bi == INT_MAX) take = true;
            else if (on != bnan) take = on;
            else if (on) take = oi < bi;
            else take = jl_isless(bc, oc) || (oc == bc && oi < bi);
            if (take) { bc = oc; bi = oi; bnan = on; }
        }
        return rfl(bi);
    }
    __device__ __forceinline__ void update_curvature(const double xs[4], const double Ps[4][2], const double Ptl[2], const double Pbr[2],
                                     double Ctol) {
        for (int i = 0; i < 4; ++i) {
            double x = xs[i], C = -INFINITY;
            const double *P = Ps[i];
            double d1 = norm2(P[0] - Ptl[0], P[1] - Ptl[1]), d2 = norm2(P[0] - Pbr[0], P[1] - Pbr[1]);
            if (jl_min(d1, d2) > Ctol) {
                // x- = largest key < x (first such entry), x+ = smallest key > x (first such entry)
                double lv = -INFINITY, rv = INFINITY;
                int li = INT_MAX, ri = INT_MAX;
                for (int base = 0; base < npc; base += kWave) {
                    int k = base + W.L;
                    if (k < npc) {
                        double xk = pcx[k];
                        if (xk < x && (li == INT_MAX || xk > lv)) { lv = xk; li = k; }
                        if (xk > x && (ri == INT_MAX || xk < rv)) { rv = xk; ri = k; }
                    }
                }
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) {
                    double olv = __shfl_xor(lv, o, kWave), orv = __shfl_xor(rv, o, kWave);
                    int oli = __shfl_xor(li, o, kWave), ori = __shfl_xor(ri, o, kWave);
                    if (oli != INT_MAX && (li == INT_MAX || olv > lv || (olv == lv && oli < li))) { lv = olv; li = oli; }
                    if (ori != INT_MAX && (ri == INT_MAX || orv < rv || (orv == rv && ori < ri))) { rv = orv; ri = ori; }
                }
                li = rfl(li);
                ri = rfl(ri);
                double Pm0 = P[0], Pm1 = P[1], Pp0 = P[0], Pp1 = P[1];
                if (li != INT_MAX) { Pm0 = rfl(pcP0[li]); Pm1 = rfl(pcP1[li]); }
                if (ri != INT_MAX) { Pp0 = rfl(pcP0[ri]); Pp1 = rfl(pcP1[ri]); }
                // menger (:967-972)
                double djk0 = Pm0 - P[0], djk1 = Pm1 - P[1];
                double dkl0 = P[0] - Pp0, dkl1 = P[1] - Pp1;
                double dlj0 = Pp0 - Pm0, dlj1 = Pp1 - Pm1;
                double a = djk0 * djk0 + djk1 * djk1, b = dkl0 * dkl0 + dkl1 * dkl1, c = dlj0 * dlj0 + dlj1 * dlj1;
                C = 2 * (djk0 * dkl1 - djk1 * dkl0) / sqrt(a * b * c);
            }
            pc_set(x, P[0], P[1], C);
        }
    }
    __device__ __forceinline__ double corner(double xlow, double xhigh, double xtol, double Ptol, double Ctol) {
        double xs[4], Ps[4][2];
        xs[0] = xlow;
        xs[3] = xhigh;
        xs[1] = (kGolden * xs[0] + xs[3]) / (kGolden + 1);
        xs[2] = xs[0] + (xs[3] - xs[1]);
        for (int i = 0; i < 4; ++i) eval(xs[i], Ps[i][0], Ps[i][1]);
        for (int i = 0; i < 4; ++i) {
            if (npc >= kLcurveCap) { overflow = true; break; }
            if (W.L == 0) { pcx[npc] = xs[i]; pcP0[npc] = Ps[i][0]; pcP1[npc] = Ps[i][1]; pcC[npc] = -INFINITY; }
            npc++;
        }
        __syncthreads();
        const double Ptl[2] = {Ps[0][0], Ps[0][1]}, Pbr[2] = {Ps[3][0], Ps[3][1]};
        update_curvature(xs, Ps, Ptl, Pbr, Ctol);
        while (!overflow && !(fabs(xs[3] - xs[0]) < xtol || norm2(Ps[0][0] - Ps[3][0], Ps[0][1] - Ps[3][1]) < Ptol)) {
            // backtracking keeps the current state (see above)
            int i2 = find(pcx, npc, xs[1]), i3 = find(pcx, npc, xs[2]);
            double C2 = rfl(pcC[i2]), C3 = rfl(pcC[i3]);
            if (C2 > C3) {
                // move_left (:933-939)
                double nx[4] = {xs[0], (kGolden * xs[0] + xs[2]) / (kGolden + 1), xs[1], xs[2]};
                double nP[4][2];
                nP[0][0] = Ps[0][0]; nP[0][1] = Ps[0][1];
                eval(nx[1], nP[1][0], nP[1][1]);
                nP[2][0] = Ps[1][0]; nP[2][1] = Ps[1][1];
                nP[3][0] = Ps[2][0]; nP[3][1] = Ps[2][1];
                for (int i = 0; i < 4; ++i) { xs[i] = nx[i]; Ps[i][0] = nP[i][0]; Ps[i][1] = nP[i][1]; }
            } else {
                // move_right (:941-946)
                double nx[4] = {xs[1], xs[2], xs[1] + (xs[3] - xs[2]), xs[3]};
                double nP[4][2];
                nP[0][0] = Ps[1][0]; nP[0][1] = Ps[1][1];
                nP[1][0] = Ps[2][0]; nP[1][1] = Ps[2][1];
                eval(nx[2], nP[2][0], nP[2][1]);
                nP[3][0] = Ps[3][0]; nP[3][1] = Ps[3][1];
                for (int i = 0; i < 4; ++i) { xs[i] = nx[i]; Ps[i][0] = nP[i][0]; Ps[i][1] = nP[i][1]; }
            }
            update_curvature(xs, Ps, Ptl, Pbr, Ctol);
        }
        return rfl(pcx[pc_argmax()]);
    }
};

__device__ __forceinline__ void reg_lcurve(Wave &W, double &mu_out, double &chi2_out) {
    W.cache_reset();
    LCurve lc(W);
    double logmu = lc.corner(-8.0, 2.0, 1e-4, 1e-4, 1e-4);
    if (lc.overflow) W.failed = true;
    double mu_final = exp(logmu);
    W.cache_solve(mu_final);
    W.take_cur_solution();
    double res2 = W.cur_resnorm_sq();
    int ns;
    double res2u = W.plain_solve_res2(ns);
    mu_out = mu_final;
    chi2_out = res2 / res2u;
}

// ------------------------------------ flip-angle search ------------------------------------
// optimize_flip_angle! (src/T2mapSEcorr.jl:409-423): cubic Hermite surrogate (src/splines.jl:504-566)
// driven by the discrete bisection searcher (:705-850) over the angle grid; each evaluation is an NNLS
// solve on the grid basis plus the surrogate gradient (:1010-1041). Lane k owns grid angle k.
__device__ __forceinline__ double evalpoly4(double t, double c0, double c1, double c2, double c3) {
    return fma(t, fma(t, fma(t, c3, c2), c1), c0);
}
// CubicHermiteInterpolator + minimize (src/splines.jl:62-110)
__device__ __forceinline__ void hermite_minimize(double a, double b, double u0, double u1, double m0, double m1, double &xo, double &uo) {
    double r = (b - a) / 2;
    m0 = r * m0;
    m1 = r * m1;
    double du = u1 - u0, dm = m1 - m0, su = u1 + u0, sm = m1 + m0;
    double c0 = su / 2 - dm / 4, c1 = (3 * du - sm) / 4, c2 = dm / 4, c3 = (sm - du) / 4;
    double xend = u0 < u1 ? a : b, uend = u0 < u1 ? u0 : u1;
    double D = 3 * (u0 - u1), th = D / 2 + (m0 + m1), g = th * th - m0 * m1;
    g = g > 0 ? -sqrt(g) : 0.0;
    double p = -(D + (m0 + m1)), q = 2 * g + (m0 - m1);
    if (fabs(p) < fabs(q)) {
        double t = p / q, y = evalpoly4(t, c0, c1, c2, c3);
        if (y < uend) {
            double c = (a + b) / 2, rr = (b - a) / 2;
            double x = fma(rr, t, c);
            xo = x < a ? a : (x > b ? b : x);
            uo = y;
            return;
        }
    }
    xo = xend;
    uo = uend;
}

struct FlipSearch {
    Wave &W;
    cgdbl *AsT, *dAsT;  // K column-major bases and their alpha-derivatives, column stride basis_ld(m)
    int K, numeval, mineval;
    unsigned long long seen;
    double al, u, du;  // lane k: grid angle, loss, gradient
    bool legacy;       // CubicSplineSurrogate (loss only, spline_opt_legacy) instead of the Hermite surrogate
    unsigned long long lseen;  // legacy: the evaluated set of the last suggestion, and that suggestion
    double lx, lu;

    __device__ __forceinline__ void eval(int k) {
        const int m = W.m, n = W.n;
        const int ldg = basis_ld(m);
        cgdbl *AT = AsT + (size_t)k * n * ldg, *dAT = dAsT + (size_t)k * n * ldg;
        int ns;
        double rn;
        W.solve(AT, false, 0.0, ns, rn);
        double ell = rn * rn;
        if (legacy) {  // loss! (src/splines.jl:1010-1014): no gradient
            if (W.L == k) { u = ell; du = 0.0; }
            return;
        }
        double part = 0.0;
        for (int i = W.L; i < m; i += kWave) {
            double ax = 0.0, dax = 0.0;
            // every column, without a branch on x_j > 0, eight columns per step with all sixteen loads issued
            // together: an inactive column has x_j = 0 exactly and fma(0, a, s) = s for finite a, which also covers
            // the columns past n - 1 that a last partial step re-reads
            for (int j0 = 0; j0 < n; j0 += 8) {
                double a[8], da[8], xv[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const int j = min(j0 + k, n - 1);
                    a[k] = AT[j * ldg + i];
                    da[k] = dAT[j * ldg + i];
                    xv[k] = (j0 + k < n) ? W.x[j] : 0.0;
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    ax = fma(xv[k], a[k], ax);
                    dax = fma(xv[k], da[k], dax);
                }
            }
            ax -= W.b0[i];
            part += dax * ax;
        }
        double dt = wsum(part);
        W.flops += 4ull * m * n + 2ull * m;
        if (W.L == k) { u = ell; du = 2 * dt; }
    }
    // update!(surr, state, I; maxeval) (:736-744)
    __device__ __forceinline__ bool update(int k) {
        if (numeval >= K) return true;
        if ((seen >> k) & 1ull) return false;
        eval(k);
        seen |= 1ull << k;
        numeval += 1;
        return numeval >= K;
    }
    __device__ __forceinline__ bool box_evaluated(int lo, int hi) { return ((seen >> lo) & 1ull) && ((seen >> hi) & 1ull); }
    // evaluate_box! (:802-815); natural corner order, or sorted by distance to x
    __device__ __forceinline__ void evaluate_box(int lo, int hi, bool sorted, double x) {
        int c0 = lo, c1 = hi;
        if (sorted) {
            double alo = readlane(al, lo), ahi = readlane(al, hi);
            double d0 = (alo - x) * (alo - x), d1 = (ahi - x) * (ahi - x);
            if (d1 < d0) { c0 = hi; c1 = lo; }
        }
        if (box_evaluated(lo, hi)) return;
        if (update(c0)) return;
        if (box_evaluated(lo, hi)) return;
        update(c1);
    }
    // suggest_point of CubicSplineSurrogate (:494-500): spline_opt_legacy through the evaluated points in grid
    // order. The result depends only on the evaluated set, so an unchanged set reuses the last one.
    __device__ void suggest_legacy(double &xo, double &uo) {
        if (seen == lseen) { xo = lx; uo = lu; return; }
        const int L = W.L;
        const bool mine = (seen >> L) & 1ull;
        const int rank = __popcll(seen & ((1ull << L) - 1ull)), np = __popcll(seen);
        if (mine) { W.spt[rank] = al; W.spc[rank] = u; }
        lds_order();
        const double xv = (L < np) ? (double)W.spt[L] : 0.0, yv = (L < np) ? (double)W.spc[L] : 0.0;
        lds_order();
        int k;
        const int nk = spline_fit(W, xv, yv, np, k);
        spline_scan(W, k, nk, false, 0.0, xo, uo);
        lseen = seen; lx = xo; lu = uo;
    }
    // suggest_point (:544-566): one Hermite segment per evaluated lane, earliest minimum wins.
    __device__ __forceinline__ void suggest(double &xo, double &uo) {
        if (legacy) { suggest_legacy(xo, uo); return; }
        const int k = W.L;
        int f = first_set(seen);
        double uf = readlane(u, f), xf = readlane(al, f);
        unsigned long long below = (k == 0) ? 0ull : (seen & ((1ull << k) - 1ull));
        int p = last_set(below);
        int src = p >= 0 ? p : k;
        double ap = __shfl(al, src, kWave), up = __shfl(u, src, kWave), dp = __shfl(du, src, kWave);
        double ys = INFINITY;
        int ord = INT_MAX;
        double xs = 0.0;
        if (((seen >> k) & 1ull) && p >= 0) {
            double xx, yy;
            hermite_minimize(ap, al, up, u, dp, du, xx, yy);
            if (yy < uf) { ys = yy; ord = k; xs = xx; }
        }
        W.flops += 30ull;
        wargmin(ys, ord);
        if (isnan(uf) || ord == INT_MAX) { xo = xf; uo = uf; return; }
        xo = readlane(xs, ord);
        uo = ys;
    }
    __device__ __forceinline__ double run() {
        seen = 0ull;
        numeval = 0;
        // initialize! (:716-734): depth-limited pre-order bisection, walked over the paths of the box tree (bit k of
        // path: the half taken at level k + 1, 0 left, 1 right) instead of a stack. A box at level len is evaluated,
        // and its children (len + 1 < depth) are visited only if fewer than mineval points have been evaluated;
        // the boxes already due (right siblings of the path) are evaluated either way, as the recursion does.
        for (int depth = 1; depth <= mineval; ++depth) {
            unsigned long long path = 0ull;
            int len = 0;
            for (;;) {
                int lo = 0, hi = K - 1;
                for (int b = 0; b < len; ++b) {
                    const int mid = (lo + hi) / 2;
                    if ((path >> b) & 1ull) lo = mid;
                    else hi = mid;
                }
                evaluate_box(lo, hi, false, 0.0);
                if (numeval < mineval && len + 1 < depth) {  // left child
                    path &= ~(1ull << len);
                    ++len;
                    continue;
                }
                while (len > 0 && ((path >> (len - 1)) & 1ull)) --len;  // leave the finished right halves
                if (len == 0) break;
                path |= 1ull << (len - 1);  // the right half after the left one
            }
            if (numeval >= mineval) break;
        }
        // bisection_search (:750-775) with minimal_bounding_box (:778-800)
        double x, uu;
        suggest(x, uu);
        for (;;) {
            int lo = 0, hi = K - 1, blo, bhi;
            for (;;) {
                int mid = (lo + hi) / 2;
                double alo = readlane(al, lo), amid = readlane(al, mid);
                if (alo <= x && x <= amid) {
                    if (!box_evaluated(lo, mid) || !(mid - lo > 1)) { blo = lo; bhi = mid; break; }
                    hi = mid;
                } else {
                    if (!box_evaluated(mid, hi) || !(hi - mid > 1)) { blo = mid; bhi = hi; break; }
                    lo = mid;
                }
            }
            evaluate_box(blo, bhi, true, x);
            suggest(x, uu);
            if (numeval >= K || (bhi - blo) <= 1) break;
        }
        return x;
    }
};

// Basis at one flip angle: n EPG curves, two per wave (32-lane groups) when the echo window fits.
template <typename T>
__device__ __forceinline__ void voxel_basis_t(Wave &W, bool general, double alpha, double beta, double E1, cgdbl *E2,
                                              gdbl *A0out) {
    const int m = W.m, n = W.n, L = W.L;
    const int ldg = basis_ld(m);  // column-major output: column j at A0out + j * ldg
    if constexpr (std::is_same<T, double>::value) {
        if (!general) {
            // two curves per lane group and pass (epg_const_pair)
            if (m / 2 + 2 <= 32) {
                const int h = L >> 5, s = L & 31;
                for (int jp = 0; jp < n; jp += 4) {
                    const int j1 = jp + h, j2 = jp + 2 + h;
                    const int jj1 = j1 < n ? j1 : n - 1, jj2 = j2 < n ? j2 : n - 1;
                    epg_const_pair<32>(alpha, E1, E2[jj1], E2[jj2], m, s, j1 < n, j2 < n, A0out + jj1 * ldg,
                                       A0out + jj2 * ldg, 1);
                }
            } else {
                for (int j = 0; j < n; j += 2) {
                    const int j2 = j + 1 < n ? j + 1 : n - 1;
                    epg_const_pair<64>(alpha, E1, E2[j], E2[j2], m, L, true, j + 1 < n, A0out + j * ldg, A0out + j2 * ldg, 1);
                }
            }
            return;
        }
    }
    if (m / 2 + 2 <= 32) {
        const int h = L >> 5, s = L & 31;
        for (int jp = 0; jp < n; jp += 2) {
            int j = jp + h;
            bool valid = j < n;
            int jj = valid ? j : n - 1;
            epg_curve<32, T>(general, alpha, beta, E1, E2[jj], m, s, valid, A0out + jj * ldg, nullptr, 1, nullptr);
        }
    } else {
        for (int j = 0; j < n; ++j)
            epg_curve<64, T>(general, alpha, beta, E1, E2[j], m, L, true, A0out + j * ldg, nullptr, 1, nullptr);
    }
}
__device__ __forceinline__ void voxel_basis(Wave &W, bool fp32, bool general, double alpha, double beta, double E1,
                                            cgdbl *E2, gdbl *A0out) {
    if (fp32) voxel_basis_t<float>(W, general, alpha, beta, E1, E2, A0out);
    else voxel_basis_t<double>(W, general, alpha, beta, E1, E2, A0out);
    // EPG op count of the reference's split recursion: ~13 flops per retained state per echo
    const unsigned long long f = (unsigned long long)W.n * 13ull * ((unsigned long long)W.m * W.m / 4 + W.m);
    // value selects, not a branch on which counter to bump: a pointer select between the two members
    // would keep both counters in scratch memory instead of registers
    W.flops32 += fp32 ? f : 0ull;
    W.flops += fp32 ? 0ull : f;
    __syncthreads();
}

// ------------------------------------ the voxel kernel ------------------------------------
// RK = Reg (0-4) + 8 for legacy = true: the legacy code is compiled only into the legacy instantiations, so the
// others keep their register allocation.
template <int RK>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void t2map_kernel(T2mapParams p) {
    constexpr int REG = RK & 7;
    constexpr bool LEGACY = (RK & 8) != 0;
    extern __shared__ double smem[];
    Wave W;
    W.L = lane_id();
    W.m = p.m;
    W.n = p.n;
    W.lda = p.lda;
    W.mw = p.mw;
    W.rl = p.rl;
    W.gld = p.gld;
    ldbl *s = (ldbl *)smem;
    // every double array has an even length, so each starts 16-byte aligned
    const int bz = (p.mw + 1) & ~1;
    W.A = s; s += (size_t)p.n * p.lda;
    W.b = s; s += bz;
    W.zz = W.b;  // the warm start forms its residual there before b takes the right-hand side
    const int ne = (p.n + 1) & ~1;
    W.x = s; s += ne;
    W.cmu = s; s += kTikhCache;
    W.crn = s; s += kTikhCache;
    W.csemi = s; s += kTikhCache;
    W.spt = s; s += LEGACY ? kSplineKnots : 0;
    W.spc = s; s += LEGACY ? kSplineMaxPts : 0;
    gdbl *slot = (gdbl *)p.scratch + (int64_t)blockIdx.x * p.slot_stride;
    // slot: the voxel's column-major basis (one copy), the L-curve caches, the Tikhonov cache solutions, the
    // normalized signal, the final distribution and the workspace rows of the split schedule
    W.A0s = slot;
    W.lc = slot + (int64_t)basis_ld(p.m) * p.n;
    W.cx = W.lc + 7 * kLcurveCap;
    W.b0 = W.cx + (size_t)kTikhCache * p.n;  // the voxel's normalized signal
    W.xf = W.b0 + ((p.m + 1) & ~1);           // the voxel's final distribution
    W.G = W.xf + ne;  // workspace rows [rl, mw) of the split schedule, n columns of stride gld
    W.flops = 0;
    W.flops32 = 0;
    W.solves = 0;
    W.nfailed = 0;
    W.status = 0;
    W.ccur = 0;
    DMI_PHASE_BEGIN();
    const int m = p.m, n = p.n, L = W.L;
    const int64_t V = p.V;
    // the count of selected voxels: a launch parameter, or (for launches recorded into a graph) the selection
    // kernel's counter, which never exceeds the list length p.nvox
    const int64_t nvox = p.nvox_dev ? (int64_t)min((unsigned long long)*p.nvox_dev, (unsigned long long)p.nvox) : p.nvox;
    for (;;) {
        unsigned int vi = 0;
        if (L == 0) vi = atomicAdd(p.counter, 1u);
        vi = (unsigned)rfl((int)__shfl((int)vi, 0, kWave));
        if ((int64_t)vi >= nvox) break;
        const int64_t I = ((cgi64 *)p.vox)[vi];
        W.failed = false;
        // voxelwise_T2_distribution! (:201-238): normalize by the max echo
        double mx = 0.0;
        for (int i = L; i < m; i += kWave) {
            double bi = ((cgdbl *)p.image)[I + V * i];
            mx = bi > mx ? bi : mx;
            W.b0[i] = bi;
        }
        double max_signal = wmaxd(mx);
        lds_order();
        if (max_signal > 0)
            for (int i = L; i < m; i += kWave) W.b0[i] /= max_signal;
        lds_order();
        double alpha_v;
        if (p.mode == kModeFixed) {
            alpha_v = p.setflip;
            W.A0 = (cgdbl *)p.Afixed;
        } else {
            if (p.mode == kModeB1) {
                alpha_v = p.alpha_in[I];
            } else {
                FlipSearch fs{W, (cgdbl *)p.AsT, (cgdbl *)p.dAsT, p.K, 0, p.mineval, 0ull, 0.0, 0.0, 0.0,
                              LEGACY, 0ull, 0.0, 0.0};
                fs.al = (L < p.K) ? p.alphas[L] : p.alphas[p.K - 1];
                fs.u = NAN;
                fs.du = NAN;
                DMI_PT0();
                alpha_v = fs.run();
                DMI_PT1(kPhFlipSearch);
            }
            DMI_PT0();
            voxel_basis(W, p.fp32 != 0, p.general != 0, alpha_v, p.beta, p.E1, (cgdbl *)p.E2, W.A0s);
            DMI_PT1(kPhVoxelBasis);
            W.A0 = W.A0s;
        }
        // T2_distribution! (src/T2mapSEcorr.jl:475-505)
        double muv = 0.0, chi2v = 1.0;
        DMI_PT0();
        if constexpr (REG == 0) W.plain_solve_to_xf();
        else if constexpr (REG == 1) reg_lcurve(W, muv, chi2v);
        else if constexpr (REG == 2) reg_gcv(W, muv, chi2v);
        else if constexpr (REG == 3) reg_chi2(W, p.chi2factor, LEGACY, muv, chi2v);
        else {
            double sigma = p.noiselevel / max_signal;
            double delta = sqrt((double)m) * sigma;
            reg_mdp(W, delta, muv, chi2v);
        }
        DMI_PT1(kPhRegTotal);
        if (W.failed) {  // outputs stay as the caller filled them (NaN)
            W.nfailed += 1;
            __syncthreads();
            continue;
        }
        // save_results! (:512-591)
        double xs = (L < n) ? W.xf[L] * max_signal : 0.0;
        if (L < n) W.xf[L] = xs;
        __syncthreads();
        // residuals of rows L and L + 64 (m <= 124 < 128)
        double res0 = 0.0, res1 = 0.0;
        double pr2 = 0.0, prs = 0.0;
        for (int i = L; i < m; i += kWave) {
            double fit = 0.0;
            for (int j = 0; j < n; ++j) fit += W.A0[j * basis_ld(m) + i] * W.xf[j];
            double bi = W.b0[i] * max_signal;
            double res = fit - bi;
            if (i == L) res0 = res; else res1 = res;
            pr2 += res * res;
            prs += res;
            if (p.decaycurve) p.decaycurve[I + V * i] = fit;
        }
        double sum_dist = wsum(xs);
        double sum_res2 = wsum(pr2);
        double mean_res = wsum(prs) / m;
        double pv = 0.0;
        if (L < m) pv += (res0 - mean_res) * (res0 - mean_res);
        if (L + kWave < m) pv += (res1 - mean_res) * (res1 - mean_res);
        double sd_res = sqrt(wsum(pv) / (m - 1));
        double lt = (L < n) ? ((cgdbl *)p.logT2)[L] : 0.0;
        double log_ggm = wsum(xs * lt) / sum_dist;
        double dv = (L < n) ? (lt - log_ggm) * (lt - log_ggm) * xs : 0.0;
        double log1p_gva = wsum(dv) / sum_dist;
        W.flops += 2ull * m * n + 12ull * m + 8ull * n;
        if (L == 0) {
            p.gdn[I] = sum_dist;
            p.ggm[I] = exp(log_ggm);
            p.gva[I] = expm1(log1p_gva);
            p.fnr[I] = sum_dist / sqrt(sum_res2 / (m - 1));
            p.snr[I] = max_signal / sd_res;
            p.alpha[I] = alpha_v;
            if (p.mu) p.mu[I] = muv;
            if (p.chi2f) p.chi2f[I] = chi2v;
            if (p.resnorm) p.resnorm[I] = sqrt(sum_res2);
        }
        if (L < n) p.dist[I + V * L] = xs;
        if (p.decaybasis && p.mode != kModeFixed && L < n)
            for (int i = 0; i < m; ++i) p.decaybasis[I + V * (i + (int64_t)m * L)] = W.A0[L * basis_ld(m) + i];
        __syncthreads();
        DMI_PT1(kPhVoxel);
    }
    DMI_PHASE_END();
    if (L == 0) {
        atomicAdd(&p.counters[0], W.flops);
        atomicAdd(&p.counters[1], W.solves);
        atomicAdd(&p.counters[2], W.flops32);
        if (W.nfailed) atomicAdd(&p.counters[3], W.nfailed);
        if (W.status) atomicOr(p.status, W.status);
    }
}

// ------------------------------------ T2partSEcorr ------------------------------------
// voxelwise_T2_parts! (src/T2partSEcorr.jl:95-138): one thread per voxel, coalesced over voxels.
__global__ void t2part_kernel(T2partParams p) {
    for (int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < p.V; I += (int64_t)gridDim.x * blockDim.x) {
        bool bad = false;
        for (int j = 0; j < p.n; ++j) bad |= isnan(p.dist[I + p.V * j]);
        if (bad) continue;
        double S = 0, Ssp = 0, Smp = 0, dsp = 0, dmp = 0, dw = 0;
        for (int j = 0; j < p.n; ++j) {
            double d = p.dist[I + p.V * j];
            S += d;
            if (p.weights) dw += d * p.weights[j];
        }
        for (int j = p.sp0; j <= p.sp1; ++j) { double d = p.dist[I + p.V * j]; dsp += d * p.logT2[j]; Ssp += d; }
        for (int j = p.mp0; j <= p.mp1; ++j) { double d = p.dist[I + p.V * j]; dmp += d * p.logT2[j]; Smp += d; }
        if (S > 0) {
            p.sfr[I] = p.weights ? dw / S : Ssp / S;
            p.mfr[I] = Smp / S;
        }
        if (Ssp > 0) p.sgm[I] = exp(dsp / Ssp);
        if (Smp > 0) p.mgm[I] = exp(dmp / Smp);
    }
}

}  // namespace

// ------------------------------------ host-side launchers ------------------------------------
#if !DMI_SPLIT_TU
#ifdef DMI_PHASE_PROFILE
extern "C" int dmi_phase_read(unsigned long long *out) {
    return (int)hipMemcpyFromSymbol(out, HIP_SYMBOL(g_phase), sizeof(g_phase));
}
extern "C" int dmi_phase_reset() {
    unsigned long long z[2 * kPhCount] = {0};
    unsigned long long zt[kMaxT2 + 1] = {0};
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(g_tikh_rows), zt, sizeof zt);
    if (e != hipSuccess) return (int)e;
    unsigned long long ze[32] = {0};
    e = hipMemcpyToSymbol(HIP_SYMBOL(g_eig_steps), ze, sizeof ze);
    if (e != hipSuccess) return (int)e;
    return (int)hipMemcpyToSymbol(HIP_SYMBOL(g_phase), z, sizeof z);
}
extern "C" int dmi_tikh_rows_read(unsigned long long *out) {
    return (int)hipMemcpyFromSymbol(out, HIP_SYMBOL(g_tikh_rows), sizeof(g_tikh_rows));
}
extern "C" int dmi_eig_steps_read(unsigned long long *out) {
    return (int)hipMemcpyFromSymbol(out, HIP_SYMBOL(g_eig_steps), sizeof(g_eig_steps));
}
#endif
int t2map_lda(int m, int n) {
    int ld = (m + n + 1) & ~1;
    return (ld % 4 == 0) ? ld + 2 : ld;
}

size_t t2map_lds_bytes(int m, int n, int lda, int K, int legacy) {
    (void)K;
    const size_t ne = (size_t)((n + 1) & ~1), bz = (size_t)((m + n + 1) & ~1);
    size_t d = (size_t)n * lda + bz + ne + 3 * kTikhCache + (legacy ? (size_t)kSplineLds : 0);
    // tri_solve reads rows L < 64 of every column: at least (n - 1) lda + 64 doubles
    d = std::max(d, (size_t)(n - 1) * lda + 64);
    return d * sizeof(double);
}

int64_t t2map_slot_doubles(int m, int n, int gld) {
    // basis, L-curve function cache (3) and point cache (4), Tikhonov cache solutions, normalized signal, final
    // distribution, the workspace rows kept outside LDS (split schedule)
    int64_t s = (int64_t)basis_ld(m) * n + 7 * (int64_t)kLcurveCap + (int64_t)kTikhCache * n + ((m + 1) & ~1) +
                ((n + 1) & ~1) + (int64_t)n * gld;
    return (s + 31) / 32 * 32;
}

hipError_t launch_grid_basis(const GridParams &p, hipStream_t s) {
    if (p.m / 2 + 2 <= 32) {
        dim3 grid(p.K, (p.n + 1) / 2);
        hipLaunchKernelGGL(grid_basis_kernel<32>, grid, dim3(64), 0, s, p);
    } else {
        dim3 grid(p.K, p.n);
        hipLaunchKernelGGL(grid_basis_kernel<64>, grid, dim3(64), 0, s, p);
    }
    return hipGetLastError();
}

hipError_t launch_transpose(const double *src, double *dst, int m, int n, int K, hipStream_t s) {
    int64_t total = (int64_t)K * m * n;
    int blocks = (int)std::min<int64_t>((total + 255) / 256, 2048);
    hipLaunchKernelGGL(transpose_kernel, dim3(blocks), dim3(256), 0, s, src, dst, m, n, K);
    return hipGetLastError();
}

template <typename From, typename To>
__global__ void convert_kernel(const From *src, To *dst, int64_t n) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        dst[t] = (To)src[t];
}

hipError_t launch_widen(const float *src, double *dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(blocks), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_narrow(const double *src, float *dst, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(blocks), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_select(const double *image, int64_t v0, int64_t v1, double thr, int64_t *vox, unsigned int *count,
                         hipStream_t s) {
    int blocks = (int)std::min<int64_t>((v1 - v0 + 255) / 256, 4096);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(select_kernel, dim3(blocks), dim3(256), 0, s, image, v0, v1, thr, vox, count);
    return hipGetLastError();
}

#endif  // !DMI_SPLIT_TU
using T2mapKernelFn = void (*)(T2mapParams);
#if DMI_SPLIT_TU
// split schedules exist for the Tikhonov-regularized methods only (lcurve, chi2, mdp)
static T2mapKernelFn t2map_kernel_for(int reg, int legacy) {
    switch (reg + (legacy ? 8 : 0)) {
    case 1: return t2map_kernel<1>;
    case 3: return t2map_kernel<3>;
    case 4: return t2map_kernel<4>;
    case 9: return t2map_kernel<9>;
    case 11: return t2map_kernel<11>;
    case 12: return t2map_kernel<12>;
    default: return nullptr;
    }
}
#else
static T2mapKernelFn t2map_kernel_for(int reg, int legacy) {
    if (legacy) {
        switch (reg) {
        case 0: return t2map_kernel<8>;
        case 1: return t2map_kernel<9>;
        case 2: return t2map_kernel<10>;
        case 3: return t2map_kernel<11>;
        default: return t2map_kernel<12>;
        }
    }
    switch (reg) {
    case 0: return t2map_kernel<0>;
    case 1: return t2map_kernel<1>;
    case 2: return t2map_kernel<2>;
    case 3: return t2map_kernel<3>;
    default: return t2map_kernel<4>;
    }
}
#endif

#if DMI_SPLIT_TU
hipError_t launch_t2map_split(const T2mapParams &p, int grid, size_t lds, hipStream_t s) {
    const T2mapKernelFn k = t2map_kernel_for(p.reg, p.legacy);
    if (!k) return hipErrorInvalidDeviceFunction;
    hipLaunchKernelGGL(k, dim3(grid), dim3(64), lds, s, p);
    return hipGetLastError();
}

hipError_t t2map_occupancy_split(int reg, int legacy, int *blocks_per_cu, size_t lds) {
    const T2mapKernelFn k = t2map_kernel_for(reg, legacy);
    if (!k) return hipErrorInvalidDeviceFunction;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k, 64, lds);
}
#else
hipError_t launch_t2map(const T2mapParams &p, int grid, size_t lds, hipStream_t s) {
#if !DMI_SLOT_ROWS
    if (p.rl < p.mw) return launch_t2map_split(p, grid, lds, s);
#endif
    hipLaunchKernelGGL(t2map_kernel_for(p.reg, p.legacy), dim3(grid), dim3(64), lds, s, p);
    return hipGetLastError();
}

hipError_t t2map_occupancy(int reg, int legacy, int *blocks_per_cu, size_t lds, bool split) {
#if !DMI_SLOT_ROWS
    if (split) return t2map_occupancy_split(reg, legacy, blocks_per_cu, lds);
#else
    (void)split;
#endif
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, t2map_kernel_for(reg, legacy), 64, lds);
}

hipError_t launch_fill(void *p, size_t bytes, unsigned int v, hipStream_t s) {
    const int64_t nw = (int64_t)(bytes / 4);
    if (nw <= 0) return hipSuccess;
    const int blocks = (int)std::min<int64_t>((nw + 255) / 256, 2048);
    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, (unsigned int *)p, nw, v);
    return hipGetLastError();
}

hipError_t noalloc_scratch(int device, double **ptr) {
    static std::mutex mu;
    static double *addr[kNoAllocMaxDevices] = {};
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void *a = nullptr;
        e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        addr[device] = static_cast<double *>(a);
    }
    *ptr = addr[device];
    return hipSuccess;
}

hipError_t launch_t2part(const T2partParams &p, hipStream_t s) {
    int blocks = (int)std::min<int64_t>((p.V + 255) / 256, 8192);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(t2part_kernel, dim3(blocks), dim3(256), 0, s, p);
    return hipGetLastError();
}

#endif  // DMI_SPLIT_TU
}  // namespace dmi
