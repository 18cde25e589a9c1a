// Host-side building blocks shared by the synchronous C ABI (dmi_api.cpp) and the request loop
// (dmi_server.cpp): device buffers, cached grid bases, option validation, launch parameter assembly.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "decaes_amd.h"
#include "dmi_internal.h"

namespace dmi_host {

// printf-style error message into `err`; returns `code`
int errorf(std::string &err, int code, const char *fmt, ...);

#define DMI_HIP(err, expr)                                                                                         \
    do {                                                                                                           \
        hipError_t _e = (expr);                                                                                    \
        if (_e != hipSuccess) return ::dmi_host::errorf((err), DMI_EHIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

// Makes `device` the calling thread's current device for the guard's lifetime and restores the thread's previous
// device at scope exit: an entry point leaves its caller's current device as it found it.
struct DeviceGuard {
    int prev = -1;
    hipError_t status;
    explicit DeviceGuard(int device) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        status = hipSetDevice(device);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    DeviceGuard(const DeviceGuard &) = delete;
    DeviceGuard &operator=(const DeviceGuard &) = delete;
};

// Declares the guard of an entry point on `device`; returns DMI_EHIP with the message in `err` if the device cannot be
// made current.
#define DMI_ON_DEVICE(err, device)                                                                                 \
    ::dmi_host::DeviceGuard dmi_device_guard_(device);                                                             \
    if (dmi_device_guard_.status != hipSuccess)                                                                    \
        return ::dmi_host::errorf((err), DMI_EHIP, "hipSetDevice(%d) failed: %s", (int)(device),                   \
                                  hipGetErrorString(dmi_device_guard_.status))

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n);
    void release();
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Everything the grid bases depend on (the reference builds them per task in thread_buffer_maker,
// src/T2mapSEcorr.jl:596-614).
struct GridKey {
    int m = -1, n = -1, K = -1, general = -1, fp32 = -1;
    double TE = 0, T1 = 0, beta = 0, T2lo = 0, T2hi = 0, minref = 0, setflip = 0;
    bool operator==(const GridKey &o) const;
    static GridKey of(const dmi_t2map_opts *o);
};

// Device arrays of one option set's grids. With SetFlipAngle, As / AsT hold the fixed basis (row-major and
// column-major) and dAs / dAsT are unused.
struct GridPtrs {
    double *E2 = nullptr, *logT2 = nullptr, *alphas = nullptr;
    double *As = nullptr, *dAs = nullptr, *AsT = nullptr, *dAsT = nullptr;
};

// Fills the grids of `o` into `g` (T2 grid, flip angles, EPG bases and their alpha-derivatives); allocates nothing,
// asynchronous on `s` except for the small host-to-device copies of the grids.
int build_grids(std::string &err, const dmi_t2map_opts *o, const GridPtrs &g, hipStream_t s);

// T2 grid, flip-angle grid and the EPG basis sets over them, on one device.
struct Grids {
    GridKey key;
    DevBuf E2, logT2, alphas, As, dAs, Afixed, AsT, dAsT, AfixedT;
    // (Re)builds the bases for `o` unless they are already current; synchronizes `s`.
    int ensure(std::string &err, const dmi_t2map_opts *o, hipStream_t s);
    GridPtrs ptrs() const;
    GridPtrs ptrs_for(bool search) const;
    void release();
};

// Per-launch device state of the voxel kernel: the work counter, tallies, status flags, scratch slots.
struct LaunchBufs {
    DevBuf counter, counters, status, scratch;
    void release();
};

void linrange(double a, double b, int n, double *out);
void logrange(double a, double b, int n, double *out);
double erfinv(double y);

// T2mapOptions @assert checks (src/types.jl:28-84) and the device limits.
int validate_t2map(std::string &err, const dmi_t2map_opts *o);
// Output-pointer checks for the options' Save* flags.
int validate_outputs(std::string &err, const dmi_t2map_opts *o, const dmi_t2map_out *d);

// Launch schedule of the voxel kernel: the LDS column stride lda of the NNLS workspace, the rows rl held in LDS
// (rows [rl, m + n) live in the workgroup's global slot with column stride gld) and the workgroups that fit a CU.
struct Schedule {
    int lda = 0, rl = 0, gld = 0, per_cu = 0;
};
// Candidates for these options: the all-LDS schedule first, then, for the Tikhonov-regularized methods (whose
// workspace has m + n rows), split schedules with the largest LDS part that fits one and two more workgroups per CU.
int t2map_schedules(std::string &err, const dmi_t2map_opts *o, std::vector<Schedule> &out);
// The process-wide choice measured for these sizes and method (false if none yet).
bool schedule_lookup(const dmi_t2map_opts *o, Schedule &s);
void schedule_store(const dmi_t2map_opts *o, const Schedule &s);
// The schedule a launch of these options uses: the caller's explicit choice, else the measured one, else all-LDS.
int t2map_schedule(std::string &err, const dmi_t2map_opts *o, Schedule &s);
void apply_schedule(dmi::T2mapParams &p, const Schedule &s);

// Kernel parameters for one request whose image, B1 map and outputs are device arrays; the voxel list and
// count are set per launch.
dmi::T2mapParams t2map_params(const dmi_t2map_opts *o, const GridPtrs &g, const double *d_image, const double *d_alpha,
                              const dmi_t2map_out *d);

// Launches the voxel kernel on voxels d_vox[0, nvox) of the request in `p` (asynchronous on `s`). Zeroes the
// work counter; tallies and status accumulate into lb.counters / lb.status until the caller clears them.
int t2map_launch(std::string &err, dmi::T2mapParams p, const int64_t *d_vox, int64_t nvox, LaunchBufs &lb,
                 int num_cu, hipStream_t s);
// The same launch on caller-provided device memory: the work counter (zeroed here by a fill kernel), the tallies and
// status words, and `slot_doubles` doubles of per-workgroup slots, which cap the persistent grid.
int t2map_launch_raw(std::string &err, dmi::T2mapParams p, const int64_t *d_vox, int64_t nvox, unsigned *counter,
                     unsigned long long *counters, int *status, double *slots, int64_t slot_doubles, int num_cu,
                     hipStream_t s);

// T2partSEcorr window ranges, log-T2 values and sigmoid weights (src/T2partSEcorr.jl:143-164).
int t2part_prepare(std::string &err, const dmi_t2part_opts *o, std::vector<double> &lg, std::vector<double> &wts,
                   dmi::T2partParams &p);

// Bytes of device memory a request of these options holds while it is in flight (image, B1 map, outputs,
// voxel list); the grid bases and scratch belong to the stream that runs it.
int64_t t2map_request_bytes(const dmi_t2map_opts *o, bool b1, bool with_t2part);

}  // namespace dmi_host
