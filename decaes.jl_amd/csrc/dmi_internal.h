// Internal host/device interface between the C ABI (dmi_api.cpp) and the gfx950 kernels (dmi_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace dmi {

constexpr int kWave = 64;
constexpr int kMaxT2 = 64;        // one lane per T2 column
constexpr int kMaxEcho = 124;     // lane-per-state EPG window (ETL/2 + 2 states)
constexpr int kMaxAngles = 64;    // one lane per flip-angle grid point
constexpr int kTikhCache = 8;     // NNLSTikhonovRegProblemCache size (src/lsqnonneg.jl:396)
constexpr int kLcurveCap = 32;    // L-curve cache entries per voxel: at most 28 are used (4 + 24 bracket moves)
// legacy = true: interpolating splines of at most 64 points (the flip-angle grid; the chi2 doubling search has at
// most 2 + kLegacyMaxDoublings), knots and coefficients in LDS
constexpr int kSplineMaxPts = 64;
constexpr int kSplineKnots = kSplineMaxPts + 8;                 // m + k + 1 <= 68, rounded up to keep 16-byte alignment
constexpr int kSplineLds = kSplineKnots + kSplineMaxPts;        // doubles
constexpr int kLegacyMaxDoublings = 24;                         // chi2 legacy: mu <= 1e-3 * 2^24 (2^24 + 1 samples)
constexpr long long kSplineMaxSamples = (1ll << 24) + 1;        // longest sample grid a legacy scan walks

// Column stride of the column-major bases in device memory (grid bases, fixed basis, the voxel's basis): m rounded
// up to even, so every column starts 16-byte aligned and is copied into LDS with 16-byte global-to-LDS loads.
__host__ __device__ inline int basis_ld(int m) { return (m + 1) & ~1; }

// Static scratch of the allocation-free entry point, in doubles: a header (launch counters), the T2 and angle grids,
// four grid-basis arrays at the device limits (row-major As, dAs; column-major AsT, dAsT), a voxel-list chunk
// (int64) and the per-workgroup slots of the voxel kernel.
constexpr int64_t kNoAllocHeader = 32;
constexpr int64_t kNoAllocGrid = 3 * 64;
constexpr int64_t kNoAllocBasis = (int64_t)kMaxAngles * kMaxT2 * kMaxEcho;   // kMaxEcho is even: basis_ld = m
constexpr int64_t kNoAllocVox = 262144;                                      // voxels selected per chunk
constexpr int64_t kNoAllocSlots = 5242880;                                   // 40 MiB of slots
constexpr int64_t kNoAllocOffGrid = kNoAllocHeader;
constexpr int64_t kNoAllocOffBasis = kNoAllocOffGrid + kNoAllocGrid;
constexpr int64_t kNoAllocOffVox = kNoAllocOffBasis + 4 * kNoAllocBasis;
constexpr int64_t kNoAllocOffSlots = kNoAllocOffVox + kNoAllocVox;
constexpr int64_t kNoAllocDoubles = kNoAllocOffSlots + kNoAllocSlots;
constexpr int kNoAllocMaxDevices = 64;

enum StatusBits : int {
    kStatusNnlsMaxIter = 2,
};

enum VoxelMode : int { kModeSearch = 0, kModeB1 = 1, kModeFixed = 2 };

struct GridParams {
    int m, n, K;
    int fp32;             // 1: FP32 recursion (mixed precision)
    int general;          // 1: RefConAngle != 180 (general refocusing train)
    double beta;          // RefConAngle
    double E1;            // exp(-(TE/2)/T1)
    const double *E2;     // n: exp(-(TE/2)/T2_j)
    const double *alphas; // K
    double *As;           // K x m x n, row-major per angle
    double *dAs;          // K x m x n or nullptr
    double *AsT;          // K x n columns of stride basis_ld(m) (column-major per angle)
};

struct T2mapParams {
    int m, n, K, mineval, reg, mode, general;
    int fp32;             // 1: FP32 EPG recursion for the voxel basis (mixed precision)
    int lda, mw;          // workspace: mw = m + n rows, column-major; rows [0, rl) in LDS with column stride lda
    int rl, gld;          // rows [rl, mw) in the workgroup's global slot, column stride gld (rl >= mw: all in LDS)
    int save_resnorm, save_decaycurve, save_regparam, save_basis;
    int legacy;           // legacy = true: spline surrogate for the flip angle, chi2 doubling search
    int64_t V;            // voxels in the volume
    int64_t nvox;         // voxels above threshold (with nvox_dev: an upper bound, the length of the voxel list)
    const unsigned int *nvox_dev;  // or nullptr; else the count of voxels above threshold, read on the device
    double E1, beta, chi2factor, noiselevel, setflip;
    const double *image;  // (V, m) column-major
    const int64_t *vox;   // nvox linear indices
    unsigned int *counter;
    const double *E2;     // n
    const double *logT2;  // n
    const double *alphas; // K
    const double *AsT;    // K grid bases, n columns of stride basis_ld(m) each
    const double *dAsT;   // d/dalpha, same layout
    const double *Afixed; // the SetFlipAngle basis, n columns of stride basis_ld(m)
    const double *alpha_in;
    double *scratch;      // per-workgroup slots
    int64_t slot_stride;  // doubles per slot
    double *dist, *gdn, *ggm, *gva, *fnr, *snr, *alpha, *resnorm, *decaycurve, *mu, *chi2f, *decaybasis;
    unsigned long long *counters; // [0] fp64 flops, [1] nnls solves, [2] fp32 flops
    int *status;
};

struct T2partParams {
    int n;
    int64_t V;
    int sp0, sp1, mp0, mp1;   // inclusive 0-based window ranges
    const double *logT2;      // n
    const double *weights;    // n or nullptr (sigmoid)
    const double *dist;
    double *sfr, *sgm, *mfr, *mgm;
};

int t2map_lda(int m, int n);
// LDS of one voxel workgroup; K (the flip-angle grid size of the launch) takes no LDS; legacy adds the spline knots
// and coefficients
size_t t2map_lds_bytes(int m, int n, int lda, int K, int legacy);
int64_t t2map_slot_doubles(int m, int n, int gld);

hipError_t launch_grid_basis(const GridParams &p, hipStream_t s);
hipError_t launch_transpose(const double *src, double *dst, int m, int n, int K, hipStream_t s);
// voxels I in [v0, v1) with image[I] > thr (image[I] is echo 1 of voxel I)
hipError_t launch_select(const double *image, int64_t v0, int64_t v1, double thr, int64_t *vox, unsigned int *count,
                         hipStream_t s);
// fills bytes / 4 32-bit words with v
hipError_t launch_fill(void *p, size_t bytes, unsigned int v, hipStream_t s);
// This device's address of the static scratch (looked up once per device with the device made current).
hipError_t noalloc_scratch(int device, double **ptr);
hipError_t launch_t2map(const T2mapParams &p, int grid, size_t lds, hipStream_t s);
hipError_t launch_t2part(const T2partParams &p, hipStream_t s);
// Float32 <-> Float64 element-wise conversion (Float32 entry points)
hipError_t launch_widen(const float *src, double *dst, int64_t n, hipStream_t s);
hipError_t launch_narrow(const double *src, float *dst, int64_t n, hipStream_t s);
// split: of the voxel kernel built for the split schedule (rows beyond an LDS cutoff in the global slot)
hipError_t t2map_occupancy(int reg, int legacy, int *blocks_per_cu, size_t lds, bool split = false);
// the voxel kernel of the split schedule, from dmi_kernels_split.hip (launch_t2map calls it when p.rl < p.mw)
hipError_t launch_t2map_split(const T2mapParams &p, int grid, size_t lds, hipStream_t s);
hipError_t t2map_occupancy_split(int reg, int legacy, int *blocks_per_cu, size_t lds);

}  // namespace dmi
