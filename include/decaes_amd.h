/*
 * decaes_amd.h -- C ABI of the MI355X (gfx950) T2mapSEcorr / T2partSEcorr path.
 *
 * Drop-in boundary for DECAES.jl v0.6.1's voxelwise path. The reference has no FFI of its own: its boundary
 * is the Julia API T2mapSEcorr!/T2partSEcorr, so each entry point below names the reference function it
 * replaces (file:line in the DECAES.jl tree). All arrays are Julia column-major; the library never keeps a
 * caller pointer after a call returns. Every function returns DMI_OK (0) or a negative DMI_E* code;
 * dmi_last_error() then describes the failure.
 */
#ifndef DECAES_AMD_H
#define DECAES_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dmi_ctx dmi_ctx;

enum {
    DMI_OK = 0,
    DMI_EINVAL = -1,       /* an option fails the reference's @assert (message says which) */
    DMI_EUNSUPPORTED = -2, /* nT2 > 64 or nTE > 124: beyond the device limits */
    DMI_ENOMEM = -3,
    DMI_EHIP = -4,         /* HIP runtime error */
    DMI_ENODEV = -5,       /* no usable gfx950 device */
    DMI_ECAPACITY = -6     /* a request does not fit a fixed device capacity (server HBM budget, static scratch) */
};

/* T2mapOptions.Reg (src/types.jl:66-68) */
enum { DMI_REG_NONE = 0, DMI_REG_LCURVE = 1, DMI_REG_GCV = 2, DMI_REG_CHI2 = 3, DMI_REG_MDP = 4 };

/* Arithmetic of the voxel path: FP64 everywhere, or FP32 EPG bases with FP64 NNLS/regularization. */
enum { DMI_PREC_FP64 = 0, DMI_PREC_MIXED = 1 };

/* Launch schedule of the voxel kernel. LDS: the whole NNLS workspace of a voxel in LDS. SPLIT: the rows beyond a
 * cutoff (Tikhonov rows appended late) in the workgroup's L2-resident global slot, so that more workgroups share a
 * CU's LDS (Tikhonov-regularized methods, when it raises the workgroups per CU). AUTO: the choice measured on the
 * first large call for these sizes and kept for the process (the all-LDS schedule until then). Every schedule
 * performs the same arithmetic: results are bit-identical. */
enum { DMI_SCHED_AUTO = 0, DMI_SCHED_LDS = 1, DMI_SCHED_SPLIT = 2 };

/* T2mapOptions (src/types.jl:19-100). Optional reals use NaN for `nothing`. */
typedef struct {
    int64_t MatrixSize[3];
    int32_t nTE;
    int32_t nT2;
    double TE;
    double T2Range[2];
    double T1;
    double Threshold; /* -Inf processes every voxel */
    double MinRefAngle;
    int32_t nRefAngles;
    int32_t nRefAnglesMin;
    int32_t Reg;
    int32_t legacy;   /* spline flip-angle surrogate and chi2 doubling search (src/lsqnonneg.jl:595-636) */
    double Chi2Factor;
    double NoiseLevel;
    double RefConAngle;
    double SetFlipAngle;
    int32_t SaveResidualNorm;
    int32_t SaveDecayCurve;
    int32_t SaveRegParam;
    int32_t SaveNNLSBasis;
    int32_t precision; /* DMI_PREC_* */
    int32_t schedule;  /* DMI_SCHED_*: launch schedule of the voxel kernel (results do not depend on it) */
} dmi_t2map_opts;

/* T2Maps + T2Distributions fields (src/T2mapSEcorr.jl:2-19,64-72). Caller-allocated, column-major, and
 * pre-filled with NaN: voxels at or below Threshold are left untouched. NULL disables an optional map.
 * decaybasis is (MatrixSize..., nTE, nT2); with SetFlipAngle it is the single nTE x nT2 basis. */
typedef struct {
    double *dist;                     /* (X,Y,Z,nT2) */
    double *gdn, *ggm, *gva, *fnr, *snr, *alpha;
    double *resnorm, *decaycurve, *mu, *chi2factor, *decaybasis;
} dmi_t2map_out;

/* T2partOptions (src/types.jl:139-172). Sigmoid = NaN for `nothing`. */
typedef struct {
    int64_t MatrixSize[3];
    int32_t nT2;
    int32_t legacy;
    double T2Range[2];
    double SPWin[2];
    double MPWin[2];
    double Sigmoid;
} dmi_t2part_opts;

/* Work counters of the last dmi_t2map_* call on this context (device-side tallies). */
typedef struct {
    int64_t voxels;        /* voxels above Threshold that were processed */
    int64_t nnls_solves;   /* Lawson-Hanson solves (flip search + regularization) */
    int64_t fp64_flops;    /* algorithmic FP64 flops: per-step counts of EPG, Householder, Givens, solves, SVD */
    int64_t fp32_flops;    /* algorithmic FP32 flops (mixed precision EPG) */
    double kernel_ms;      /* device time of the voxel kernel, HIP events */
    double total_ms;       /* device time of the whole call */
    int64_t voxels_failed; /* voxels left unwritten (NaN) because a per-voxel search cache overflowed; the L-curve
                              caches hold every point the search can visit, so this stays 0 */
    int32_t lds_rows;      /* schedule of the last launch: workspace rows held in LDS (the rest in the global slot) */
    int32_t workgroups_per_cu; /* and its resident workgroups (one voxel each) per CU */
} dmi_stats;

/* Context: one HIP stream on `device`, reusable device buffers and cached grid bases.
 * Replaces the per-task thread_buffer_maker (src/T2mapSEcorr.jl:596-614). flags: reserved, pass 0.
 * Every call on a context runs with its device current and restores the calling thread's current device on return. */
int dmi_create(dmi_ctx **out, int device, uint32_t flags);
int dmi_destroy(dmi_ctx *ctx);
/* Runs ctx's later calls on `stream` (a hipStream_t of ctx's device), or again on ctx's own stream when stream is
 * NULL. No counterpart in the reference (a Julia task has no stream): it puts this library's kernels in the caller's
 * stream order, and lets a caller capture them into a HIP graph with its own kernels: while `stream` is being
 * captured, dmi_t2map_f64_device_noalloc and dmi_t2part_f64_device record their launches and return at once (see
 * there); the other compute entry points return DMI_EINVAL. ctx never destroys a caller's stream; the stream must
 * outlive its use by ctx, and the work ctx queued on it must have finished before dmi_destroy(ctx). */
int dmi_set_stream(dmi_ctx *ctx, void *stream);
/* Error string of the last failed call on ctx (or of dmi_create when ctx is NULL). */
const char *dmi_last_error(const dmi_ctx *ctx);

/* T2mapSEcorr!(maps, dist, image, opts) (src/T2mapSEcorr.jl:151-196) with host arrays.
 * image: (X,Y,Z,nTE). alpha_in: the B1 map passed through load_B1map! (src/T2mapSEcorr.jl:56-61), or NULL. */
int dmi_t2map_f64(dmi_ctx *ctx, const double *image, const dmi_t2map_opts *opts, const double *alpha_in,
                  const dmi_t2map_out *out);
/* Same with every array already in device memory (hipMalloc'd on ctx's device); runs on ctx's stream and
 * returns after the stream has finished. Work the caller queued on other streams for these arrays (an input still
 * being written) must be complete before the call. */
int dmi_t2map_f64_device(dmi_ctx *ctx, const double *d_image, const dmi_t2map_opts *opts, const double *d_alpha_in,
                         const dmi_t2map_out *d_out);

/* dmi_t2map_f64_device without any device allocation. No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Intermediate buffers (grid bases, voxel list, per-workgroup slots, counters) live in a static device array of
 * about 60 MB that each device holds for the life of the process; the caller passes inputs and outputs as for
 * dmi_t2map_f64_device. Limits: nTE <= 124, nT2 <= 64, nRefAngles <= 64 (DMI_ECAPACITY beyond); the volume is
 * processed in chunks of 262144 voxels, so its size is not limited. Every call on one device shares that array:
 * calls that overlap on one device (from several contexts or threads) race and must be serialized by the caller.
 * Under HIP graph capture of ctx's stream (dmi_set_stream) the call records its launches (per chunk: voxel
 * selection and the voxel kernel, the count of selected voxels staying on the device) and returns without waiting;
 * every replay of the graph recomputes the outputs from the image's current contents. The grids of the options must
 * already be in the static array (an earlier call with the same options outside capture; DMI_EINVAL otherwise), and
 * a replay reads them there: a call with other options on the device in between rebuilds them, and the graph's
 * results are wrong until a call with its options has rebuilt them. dmi_last_stats is not filled by a captured
 * call. */
int dmi_t2map_f64_device_noalloc(dmi_ctx *ctx, const double *d_image, const dmi_t2map_opts *opts,
                                 const double *d_alpha_in, const dmi_t2map_out *d_out);

/* Float32 arrays: the reference is generic in the element type (T2mapSEcorr(image::Array{T,4}, opts::
 * T2mapOptions{T}), src/T2mapSEcorr.jl:148-149, and T2partSEcorr(dist::Array{T,4}, ...), src/T2partSEcorr.jl:40).
 * Here a Float32 image is widened to FP64 on the device, the FP64 path runs, and every output is rounded to
 * Float32 once, so the results are the FP64 path's rounded to Float32 (the reference instead computes in Float32). */
typedef struct {
    float *dist;
    float *gdn, *ggm, *gva, *fnr, *snr, *alpha;
    float *resnorm, *decaycurve, *mu, *chi2factor, *decaybasis;
} dmi_t2map_out_f32;
int dmi_t2map_f32(dmi_ctx *ctx, const float *image, const dmi_t2map_opts *opts, const float *alpha_in,
                  const dmi_t2map_out_f32 *out);
int dmi_t2part_f32(dmi_ctx *ctx, const float *dist, const dmi_t2part_opts *opts, float *sfr, float *sgm, float *mfr,
                   float *mgm);

/* maps["decaybasisset"] (src/T2mapSEcorr.jl:31-33): nTE x nT2 x nRefAngles, or nTE x nT2 with SetFlipAngle. */
int dmi_epg_basis_set(dmi_ctx *ctx, const dmi_t2map_opts *opts, double *out);

/* T2partSEcorr(dist, opts) (src/T2partSEcorr.jl:40-71): dist (X,Y,Z,nT2) -> sfr, sgm, mfr, mgm (X,Y,Z),
 * caller-allocated and NaN-filled. Host and device-pointer forms. Under HIP graph capture of ctx's stream the
 * device form records its one launch, reading the T2 tables that the last call outside capture on ctx uploaded;
 * they must be those of the same options (DMI_EINVAL otherwise). */
int dmi_t2part_f64(dmi_ctx *ctx, const double *dist, const dmi_t2part_opts *opts, double *sfr, double *sgm,
                   double *mfr, double *mgm);
int dmi_t2part_f64_device(dmi_ctx *ctx, const double *d_dist, const dmi_t2part_opts *opts, double *d_sfr,
                          double *d_sgm, double *d_mfr, double *d_mgm);

int dmi_last_stats(const dmi_ctx *ctx, dmi_stats *out);

/* ---- Request loop ----
 * A device-resident counterpart of the reference's warm server (api/decaes_server.jl:95-131, which keeps a Julia
 * process alive and runs one CLI call at a time). Requests (a T2mapSEcorr call, optionally followed by
 * T2partSEcorr on its distribution) are admitted while their device footprint fits the HBM budget, split into
 * voxel chunks, and interleaved on the GPU: requests with the same grid options share a stream and its cached
 * grid bases, whose next chunk goes to the request with the fewest voxels left (or to each in turn,
 * dmi_server_config.order); different option sets run on separate streams concurrently. A request
 * that arrives while others run starts at the next chunk boundary. Each voxel's result is the same as in a
 * dmi_t2map_f64 call (voxels are independent and the per-voxel computation does not depend on the batch).
 * Host arrays passed to dmi_server_submit must stay valid until dmi_server_wait returns for that request. */
typedef struct dmi_server dmi_server;

typedef struct {
    int64_t hbm_budget_bytes; /* device memory for admitted requests; <= 0 selects half of the free memory */
    int32_t max_streams;      /* option sets in flight at once, one HIP stream each; <= 0 selects 4 */
    int32_t chunk_voxels;     /* voxels per launch; <= 0 selects 65536 */
    int32_t order;            /* which admitted request of a stream gets the next chunk: DMI_ORDER_* (else DMI_EINVAL) */
    int32_t reserved;         /* pass 0 */
} dmi_server_config;

/* dmi_server_config.order. SHORTEST (the default, 0): the request with the fewest voxels left to launch, the earliest
 * admitted on a tie, so short requests overtake long ones. TURNS: the stream's requests take turns, one chunk each. */
enum { DMI_ORDER_SHORTEST = 0, DMI_ORDER_TURNS = 1 };

typedef struct {
    int32_t status;           /* DMI_OK or the request's error code (dmi_server_request_error has the text) */
    int32_t workgroups_per_cu; /* launch schedule of the request's chunks (dmi_stats.workgroups_per_cu) */
    int64_t voxels;           /* voxels above Threshold */
    double queue_ms;          /* submit -> admitted */
    double latency_ms;        /* submit -> outputs in host memory */
    int64_t trial_voxels;     /* voxels run as schedule trials (the first request of sizes with none measured) */
} dmi_request_info;

typedef struct {
    int64_t submitted, completed, failed;
    int64_t voxels;           /* voxels processed by completed requests */
    double busy_ms;           /* wall time with at least one request admitted */
    double p50_ms, p99_ms;    /* latency percentiles over completed requests */
} dmi_server_stats;

int dmi_server_create(dmi_server **out, int device, const dmi_server_config *cfg);
/* Waits for every submitted request, then frees the server. */
int dmi_server_destroy(dmi_server *s);
/* Queues T2mapSEcorr (+ T2partSEcorr when popts is not NULL) on host arrays; returns at once with *id.
 * Invalid options fail here with the reference's @assert message and nothing is queued. */
int dmi_server_submit(dmi_server *s, const double *image, const dmi_t2map_opts *opts, const double *alpha_in,
                      const dmi_t2map_out *out, const dmi_t2part_opts *popts, double *sfr, double *sgm, double *mfr,
                      double *mgm, int64_t *id);
/* Blocks until request id has finished (successfully or not), then forgets it: a request is waited for once.
 * p50/p99 in dmi_server_get_stats cover the most recent 65536 completed requests. */
int dmi_server_wait(dmi_server *s, int64_t id, dmi_request_info *info);
/* *done = 1 once request id has finished. */
int dmi_server_poll(dmi_server *s, int64_t id, int32_t *done);
int dmi_server_get_stats(dmi_server *s, dmi_server_stats *out);
const char *dmi_server_last_error(const dmi_server *s);
/* Error text of request id (before its wait, or among the 256 most recent failures already waited for). */
const char *dmi_server_request_error(dmi_server *s, int64_t id);

/* Seeded synthetic multi-echo spin-echo volume (host memory, no device work): per voxel a sum of
 * constant-flip EPG compartments plus Rician noise, zero outside an ellipsoid mask. Counter-based RNG keyed
 * by (seed, voxel), so any thread count gives the same bits. ncomp = 3 or 4 (the fourth is T2 ~ U(35,50) ms).
 * Replaces nothing in the reference's hot path; it is the input generator modelled on mock_image
 * (src/utils.jl:623-658). Returns the number of voxels inside the mask. */
int64_t dmi_synth_mse(double *image, int64_t nx, int64_t ny, int64_t nz, int32_t nTE, double TE, double T1,
                      int32_t ncomp, double mask_fraction, uint64_t seed, int32_t nthreads);

#ifdef __cplusplus
}
#endif
#endif
