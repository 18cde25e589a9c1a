// C ABI of the gfx950 T2mapSEcorr / T2partSEcorr path (declared in include/decaes_amd.h): the synchronous
// entry points. All voxel work runs on the device; there is no CPU fallback path.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "decaes_amd.h"
#include "dmi_host.h"
#include "dmi_internal.h"

using dmi_host::DevBuf;

namespace {
thread_local std::string g_create_error;
}  // namespace

struct dmi_ctx {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;      // the stream every call runs on: own_stream, or the caller's (dmi_set_stream)
    hipStream_t own_stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    std::string err;
    std::mutex mu;
    dmi_stats stats{};
    dmi_host::Grids grids;
    dmi_host::LaunchBufs lb;
    double *noalloc = nullptr;  // this device's address of the static scratch (dmi_t2map_f64_device_noalloc)
    DevBuf vox;
    // host-array staging (FP64 working copies, FP32 transfer buffers)
    DevBuf img, ain, out[12], t2p[5];
    DevBuf f32;
    DevBuf weights, tlog;
    std::vector<double> tlog_host, weights_host;  // the T2part tables last uploaded to tlog / weights
    // dmi_t2map_f64 on volumes of several chunks: a copy stream and two staging sets (image, B1 map, outputs)
    hipStream_t copy_stream = nullptr;
    DevBuf pimg[2], pain[2], pout[2][12], pcnt;
};

namespace {

int fail(dmi_ctx *c, int code, const char *msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPCHECK(c, expr) DMI_HIP((c)->err, expr)

// Whether ctx's stream is being captured into a HIP graph: calls then record their launches without host reads,
// waits, events or allocations.
int capturing(dmi_ctx *c, bool &rec) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIPCHECK(c, hipStreamIsCapturing(c->stream, &cs));
    rec = cs != hipStreamCaptureStatusNone;
    return DMI_OK;
}

// Entry points on host arrays copy through pageable memory and wait: refused while the stream is captured.
int no_capture(dmi_ctx *c, const char *name) {
    bool rec = false;
    int rc = capturing(c, rec);
    if (rc) return rc;
    if (rec)
        return dmi_host::errorf(c->err, DMI_EINVAL, "%s copies host arrays and waits, which graph capture does not "
                                "allow: use the device-array entry points", name);
    return DMI_OK;
}

// Device memory one T2mapSEcorr call works in: the grids of its options, a voxel list of vox_cap entries, the
// launch counters and the per-workgroup slots of the voxel kernel.
struct RunBufs {
    dmi_host::GridPtrs g;
    int64_t *vox;
    int64_t vox_cap;
    unsigned *counter;
    unsigned long long *counters;
    int *status;
    double *slots;
    int64_t slot_doubles;
};

// Voxel selection and the voxel kernel over [0, V) in chunks of vox_cap voxels (one chunk when the list holds the
// whole volume); allocates nothing.
int run_t2map_on(dmi_ctx *c, const double *d_image, const dmi_t2map_opts *o, const double *d_alpha_in,
                 const dmi_t2map_out *d, const RunBufs &rb) {
    const int m = o->nTE, n = o->nT2;
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    const bool fixed = !std::isnan(o->SetFlipAngle);
    // the shared basis of SetFlipAngle, nTE x nT2 column-major (the column-major copy without its column padding)
    if (fixed && o->SaveNNLSBasis)
        HIPCHECK(c, hipMemcpy2DAsync(d->decaybasis, sizeof(double) * m, rb.g.AsT, sizeof(double) * dmi::basis_ld(m),
                                     sizeof(double) * m, n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(c, dmi::launch_fill(rb.counters, sizeof(unsigned long long) * 4, 0u, c->stream));
    HIPCHECK(c, dmi::launch_fill(rb.status, sizeof(int) * 4, 0u, c->stream));
    c->stats = dmi_stats{};
    dmi::T2mapParams p = dmi_host::t2map_params(o, rb.g, d_image, d_alpha_in, d);
    std::vector<dmi_host::Schedule> cands;
    int rc0 = dmi_host::t2map_schedules(c->err, o, cands);
    if (rc0) return rc0;
    dmi_host::Schedule sched;
    rc0 = dmi_host::t2map_schedule(c->err, o, sched);
    if (rc0) return rc0;
    // the autotuner: with several candidate schedules and none measured yet for these sizes, the first call large
    // enough runs consecutive slices of its voxel list with each candidate, keeps the fastest for the process and
    // finishes with it (every schedule gives the same bits, so no work is repeated)
    bool tune = o->schedule == DMI_SCHED_AUTO && cands.size() > 1 && !dmi_host::schedule_lookup(o, sched);
    bool launched = false;
    double tune_ms = 0.0;
    for (int64_t v0 = 0; v0 < V; v0 += rb.vox_cap) {
        const int64_t v1 = std::min<int64_t>(V, v0 + rb.vox_cap);
        HIPCHECK(c, dmi::launch_fill(rb.counter, sizeof(unsigned), 0u, c->stream));
        HIPCHECK(c, dmi::launch_select(d_image, v0, v1, o->Threshold, rb.vox, rb.counter, c->stream));
        unsigned nvox = 0;
        HIPCHECK(c, hipMemcpyAsync(&nvox, rb.counter, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        c->stats.voxels += nvox;
        if (nvox == 0) continue;
        int64_t done = 0;
        if (tune) {
            int64_t need = 0;
            for (const auto &s : cands) need += 8ll * c->num_cu * s.per_cu;
            if ((int64_t)nvox >= 2 * need) {
                double best_rate = -1.0;
                for (const auto &s : cands) {
                    const int64_t cnt = 8ll * c->num_cu * s.per_cu;
                    dmi::T2mapParams ps = p;
                    dmi_host::apply_schedule(ps, s);
                    HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
                    int rc = dmi_host::t2map_launch_raw(c->err, ps, rb.vox + done, cnt, rb.counter, rb.counters,
                                                        rb.status, rb.slots, rb.slot_doubles, c->num_cu, c->stream);
                    if (rc) return rc;
                    HIPCHECK(c, hipEventRecord(c->ev[2], c->stream));
                    HIPCHECK(c, hipEventSynchronize(c->ev[2]));
                    float ms = 0;
                    HIPCHECK(c, hipEventElapsedTime(&ms, c->ev[1], c->ev[2]));
                    tune_ms += ms;
                    const double rate = cnt / std::max(ms, 1e-3f);
                    if (rate > best_rate) { best_rate = rate; sched = s; }
                    done += cnt;
                }
                dmi_host::schedule_store(o, sched);
                tune = false;
            }
        }
        if (!launched) HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
        launched = true;
        dmi_host::apply_schedule(p, sched);
        c->stats.lds_rows = std::min(p.rl, p.mw);
        c->stats.workgroups_per_cu = sched.per_cu;
        int rc = dmi_host::t2map_launch_raw(c->err, p, rb.vox + done, nvox - done, rb.counter, rb.counters, rb.status,
                                            rb.slots, rb.slot_doubles, c->num_cu, c->stream);
        if (rc) return rc;
    }
    unsigned long long cnt[4] = {0, 0, 0, 0};
    int status = 0;
    if (launched) {
        HIPCHECK(c, hipEventRecord(c->ev[2], c->stream));
        HIPCHECK(c, hipMemcpyAsync(cnt, rb.counters, sizeof cnt, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipMemcpyAsync(&status, rb.status, sizeof status, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHECK(c, hipEventRecord(c->ev[3], c->stream));
    HIPCHECK(c, hipEventSynchronize(c->ev[3]));
    float kms = 0, tms = 0;
    if (launched) HIPCHECK(c, hipEventElapsedTime(&kms, c->ev[1], c->ev[2]));
    HIPCHECK(c, hipEventElapsedTime(&tms, c->ev[0], c->ev[3]));
    c->stats.nnls_solves = (int64_t)cnt[1];
    c->stats.fp64_flops = (int64_t)cnt[0];
    c->stats.fp32_flops = (int64_t)cnt[2];
    c->stats.voxels_failed = (int64_t)cnt[3];
    c->stats.kernel_ms = kms + tune_ms;
    c->stats.total_ms = tms;
    (void)status;
    return DMI_OK;
}

// run_t2map_on recorded into the HIP graph being captured on ctx's stream: the same launches without a host read,
// wait or event. Each chunk's count of selected voxels stays on the device (counter[0], read by the voxel kernel,
// which takes its work from counter[1]); the schedule is the one measured for these sizes, or the default, and the
// persistent grid is sized for the whole chunk (workgroups beyond the selected voxels end at once). Statistics are
// not filled.
int record_t2map_on(dmi_ctx *c, const double *d_image, const dmi_t2map_opts *o, const double *d_alpha_in,
                    const dmi_t2map_out *d, const RunBufs &rb) {
    const int m = o->nTE, n = o->nT2;
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    if (!std::isnan(o->SetFlipAngle) && o->SaveNNLSBasis)
        HIPCHECK(c, hipMemcpy2DAsync(d->decaybasis, sizeof(double) * m, rb.g.AsT, sizeof(double) * dmi::basis_ld(m),
                                     sizeof(double) * m, n, hipMemcpyDeviceToDevice, c->stream));
    HIPCHECK(c, dmi::launch_fill(rb.counters, sizeof(unsigned long long) * 4, 0u, c->stream));
    HIPCHECK(c, dmi::launch_fill(rb.status, sizeof(int) * 4, 0u, c->stream));
    c->stats = dmi_stats{};
    dmi::T2mapParams p = dmi_host::t2map_params(o, rb.g, d_image, d_alpha_in, d);
    dmi_host::Schedule sched;
    int rc = dmi_host::t2map_schedule(c->err, o, sched);
    if (rc) return rc;
    dmi_host::apply_schedule(p, sched);
    unsigned *count = rb.counter, *work = rb.counter + 1;
    p.nvox_dev = count;
    for (int64_t v0 = 0; v0 < V; v0 += rb.vox_cap) {
        const int64_t v1 = std::min<int64_t>(V, v0 + rb.vox_cap);
        HIPCHECK(c, dmi::launch_fill(count, sizeof(unsigned), 0u, c->stream));
        HIPCHECK(c, dmi::launch_select(d_image, v0, v1, o->Threshold, rb.vox, count, c->stream));
        rc = dmi_host::t2map_launch_raw(c->err, p, rb.vox, v1 - v0, work, rb.counters, rb.status, rb.slots,
                                        rb.slot_doubles, c->num_cu, c->stream);
        if (rc) return rc;
    }
    return DMI_OK;
}

int run_t2map(dmi_ctx *c, const double *d_image, const dmi_t2map_opts *o, const double *d_alpha_in,
              const dmi_t2map_out *d) {
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    int rc = dmi_host::validate_outputs(c->err, o, d);
    if (rc) return rc;
    bool rec = false;
    if ((rc = capturing(c, rec))) return rc;
    if (rec)
        return dmi_host::errorf(c->err, DMI_EINVAL, "this entry point allocates and waits, which graph capture does "
                                "not allow: use dmi_t2map_f64_device_noalloc");
    HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    rc = c->grids.ensure(c->err, o, c->stream);
    if (rc) return rc;
    HIPCHECK(c, c->vox.ensure(sizeof(int64_t) * std::max<int64_t>(V, 1)));
    HIPCHECK(c, c->lb.counter.ensure(sizeof(unsigned) * 4));
    HIPCHECK(c, c->lb.counters.ensure(sizeof(unsigned long long) * 4));
    HIPCHECK(c, c->lb.status.ensure(sizeof(int) * 4));
    // slots for the largest persistent grid the voxel kernel can run with these options, under any schedule
    std::vector<dmi_host::Schedule> cands;
    rc = dmi_host::t2map_schedules(c->err, o, cands);
    if (rc) return rc;
    int64_t slot_doubles = 0;
    for (const auto &s : cands)
        slot_doubles = std::max<int64_t>(slot_doubles, dmi::t2map_slot_doubles(o->nTE, o->nT2, s.rl >= o->nTE + o->nT2 ? 0 : s.gld) *
                                                           std::max<int64_t>(1, std::min<int64_t>((int64_t)c->num_cu * std::max(s.per_cu, 1), V)));
    HIPCHECK(c, c->lb.scratch.ensure(sizeof(double) * (size_t)slot_doubles));
    RunBufs rb{c->grids.ptrs(), c->vox.as<int64_t>(), std::max<int64_t>(V, 1), c->lb.counter.as<unsigned>(),
               c->lb.counters.as<unsigned long long>(), c->lb.status.as<int>(), c->lb.scratch.as<double>(), slot_doubles};
    return run_t2map_on(c, d_image, o, d_alpha_in, d, rb);
}

// dmi_t2map_f64 on a volume of more than one chunk. The voxels are taken as consecutive ranges of kHostChunk linear
// indices, each handled as a volume of its own (MatrixSize (len, 1, 1): a voxel's result does not depend on the
// volume around it), in one of two device staging sets. The copy stream moves the next chunk's image rows and the
// outputs' current contents in, and the previous chunk's outputs out, while ctx's stream runs the voxel kernel on
// the current chunk; each chunk's count of selected voxels stays on the device, so no chunk waits for the host.
// hp[i] / rows[i]: the caller's per-voxel output arrays and their rows of V doubles (NULL: not requested).
constexpr int64_t kHostChunk = 524288;

struct PipeEvents {
    hipEvent_t in[2] = {nullptr, nullptr}, done[2] = {nullptr, nullptr};
    ~PipeEvents() {
        for (int b = 0; b < 2; ++b) {
            if (in[b]) (void)hipEventDestroy(in[b]);
            if (done[b]) (void)hipEventDestroy(done[b]);
        }
    }
};

int run_t2map_host_chunks(dmi_ctx *c, const double *image, const dmi_t2map_opts *o, const double *alpha_in,
                          double *const hp[12], const int64_t rows[12]) {
    const int m = o->nTE, n = o->nT2;
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    const int64_t C = kHostChunk, nch = (V + C - 1) / C;
    const bool fixed = !std::isnan(o->SetFlipAngle);
    HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    int rc = c->grids.ensure(c->err, o, c->stream);
    if (rc) return rc;
    if (!c->copy_stream) HIPCHECK(c, hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    hipStream_t cs = c->copy_stream;
    HIPCHECK(c, c->vox.ensure(sizeof(int64_t) * C));
    HIPCHECK(c, c->lb.counter.ensure(sizeof(unsigned) * 4));
    HIPCHECK(c, c->lb.counters.ensure(sizeof(unsigned long long) * 4));
    HIPCHECK(c, c->lb.status.ensure(sizeof(int) * 4));
    HIPCHECK(c, c->pcnt.ensure(sizeof(unsigned) * (size_t)nch));
    // each chunk as a volume of its own; the fixed basis of SetFlipAngle is copied once at the end
    dmi_t2map_opts oc = *o;
    oc.MatrixSize[0] = C;
    oc.MatrixSize[1] = 1;
    oc.MatrixSize[2] = 1;
    if (fixed) oc.SaveNNLSBasis = 0;
    std::vector<dmi_host::Schedule> cands;
    rc = dmi_host::t2map_schedules(c->err, &oc, cands);
    if (rc) return rc;
    int64_t slot_doubles = 0;
    for (const auto &s : cands)
        slot_doubles = std::max<int64_t>(slot_doubles, dmi::t2map_slot_doubles(m, n, s.rl >= m + n ? 0 : s.gld) *
                                                           std::max<int64_t>(1, std::min<int64_t>((int64_t)c->num_cu * std::max(s.per_cu, 1), C)));
    HIPCHECK(c, c->lb.scratch.ensure(sizeof(double) * (size_t)slot_doubles));
    for (int b = 0; b < 2; ++b) {
        HIPCHECK(c, c->pimg[b].ensure(sizeof(double) * (size_t)(C * m)));
        if (alpha_in) HIPCHECK(c, c->pain[b].ensure(sizeof(double) * (size_t)C));
        for (int i = 0; i < 12; ++i)
            if (hp[i]) HIPCHECK(c, c->pout[b][i].ensure(sizeof(double) * (size_t)(C * rows[i])));
    }
    RunBufs rb{c->grids.ptrs(), c->vox.as<int64_t>(), C, c->lb.counter.as<unsigned>(),
               c->lb.counters.as<unsigned long long>(), c->lb.status.as<int>(), c->lb.scratch.as<double>(), slot_doubles};
    unsigned *cnt = c->pcnt.as<unsigned>();
    auto span = [&](int64_t k, int64_t &v0, int64_t &len) { v0 = k * C; len = std::min<int64_t>(C, V - v0); };
    // chunk k's rows between the caller's arrays (row pitch V) and staging set b (row pitch len)
    auto stage = [&](int64_t k, int b, bool in) -> int {
        int64_t v0, len;
        span(k, v0, len);
        const size_t w = sizeof(double) * (size_t)len, hpitch = sizeof(double) * (size_t)V;
        if (in) {
            HIPCHECK(c, hipMemcpy2DAsync(c->pimg[b].p, w, image + v0, hpitch, w, m, hipMemcpyHostToDevice, cs));
            if (alpha_in) HIPCHECK(c, hipMemcpyAsync(c->pain[b].p, alpha_in + v0, w, hipMemcpyHostToDevice, cs));
        }
        for (int i = 0; i < 12; ++i) {
            if (!hp[i]) continue;
            if (in)
                HIPCHECK(c, hipMemcpy2DAsync(c->pout[b][i].p, w, hp[i] + v0, hpitch, w, (size_t)rows[i],
                                             hipMemcpyHostToDevice, cs));
            else
                HIPCHECK(c, hipMemcpy2DAsync(hp[i] + v0, hpitch, c->pout[b][i].p, w, w, (size_t)rows[i],
                                             hipMemcpyDeviceToHost, cs));
        }
        return DMI_OK;
    };
    auto chunk_opts = [&](int64_t k) {
        int64_t v0, len;
        span(k, v0, len);
        dmi_t2map_opts ok = oc;
        ok.MatrixSize[0] = len;
        return ok;
    };
    auto chunk_out = [&](int b) {
        double *d[12];
        for (int i = 0; i < 12; ++i) d[i] = hp[i] ? c->pout[b][i].as<double>() : nullptr;
        return dmi_t2map_out{d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], d[11]};
    };
    c->stats = dmi_stats{};
    dmi_host::Schedule sched;
    int64_t k0 = 0, voxels = 0;
    double kms0 = 0.0;
    const bool tune = oc.schedule == DMI_SCHED_AUTO && cands.size() > 1 && !dmi_host::schedule_lookup(&oc, sched);
    if (tune) {
        // the first chunk on its own, with the autotuner, which waits for its trial launches
        if ((rc = stage(0, 0, true))) return rc;
        HIPCHECK(c, hipStreamSynchronize(cs));
        const dmi_t2map_opts o0 = chunk_opts(0);
        const dmi_t2map_out d0 = chunk_out(0);
        rc = run_t2map_on(c, c->pimg[0].as<double>(), &o0, alpha_in ? c->pain[0].as<double>() : nullptr, &d0, rb);
        if (rc) return rc;
        if ((rc = stage(0, 0, false))) return rc;
        HIPCHECK(c, hipStreamSynchronize(cs));
        voxels = c->stats.voxels;
        kms0 = c->stats.kernel_ms;
        k0 = 1;
    } else {
        HIPCHECK(c, dmi::launch_fill(rb.counters, sizeof(unsigned long long) * 4, 0u, c->stream));
        HIPCHECK(c, dmi::launch_fill(rb.status, sizeof(int) * 4, 0u, c->stream));
    }
    rc = dmi_host::t2map_schedule(c->err, &oc, sched);
    if (rc) return rc;
    PipeEvents ev;
    for (int b = 0; b < 2; ++b) {
        HIPCHECK(c, hipEventCreateWithFlags(&ev.in[b], hipEventDisableTiming));
        HIPCHECK(c, hipEventCreateWithFlags(&ev.done[b], hipEventDisableTiming));
    }
    auto pipeline = [&]() -> int {
        if (k0 >= nch) return DMI_OK;
        HIPCHECK(c, dmi::launch_fill(cnt + k0, sizeof(unsigned) * (size_t)(nch - k0), 0u, c->stream));
        // the copy stream starts after the counts are cleared and the grids are in place
        HIPCHECK(c, hipEventRecord(ev.done[1], c->stream));
        HIPCHECK(c, hipStreamWaitEvent(cs, ev.done[1], 0));
        int r = stage(k0, (int)(k0 % 2), true);
        if (r) return r;
        HIPCHECK(c, hipEventRecord(ev.in[k0 % 2], cs));
        HIPCHECK(c, hipEventRecord(c->ev[1], c->stream));
        for (int64_t k = k0; k < nch; ++k) {
            const int b = (int)(k % 2), b1 = 1 - b;
            HIPCHECK(c, hipStreamWaitEvent(c->stream, ev.in[b], 0));
            const dmi_t2map_opts ok = chunk_opts(k);
            const dmi_t2map_out dk = chunk_out(b);
            const int64_t len = ok.MatrixSize[0];
            dmi::T2mapParams p = dmi_host::t2map_params(&ok, rb.g, c->pimg[b].as<double>(),
                                                        alpha_in ? c->pain[b].as<double>() : nullptr, &dk);
            dmi_host::apply_schedule(p, sched);
            p.nvox_dev = cnt + k;
            HIPCHECK(c, dmi::launch_select(c->pimg[b].as<double>(), 0, len, o->Threshold, rb.vox, cnt + k, c->stream));
            r = dmi_host::t2map_launch_raw(c->err, p, rb.vox, len, rb.counter + 1, rb.counters, rb.status, rb.slots,
                                           rb.slot_doubles, c->num_cu, c->stream);
            if (r) return r;
            HIPCHECK(c, hipEventRecord(ev.done[b], c->stream));
            if (k + 1 < nch) {
                // staging set b1 held chunk k - 1: its outputs go home before chunk k + 1 moves in
                if (k - 1 >= k0) {
                    HIPCHECK(c, hipStreamWaitEvent(cs, ev.done[b1], 0));
                    if ((r = stage(k - 1, b1, false))) return r;
                }
                if ((r = stage(k + 1, b1, true))) return r;
                HIPCHECK(c, hipEventRecord(ev.in[b1], cs));
            }
        }
        HIPCHECK(c, hipEventRecord(c->ev[2], c->stream));
        for (int64_t k = std::max(k0, nch - 2); k < nch; ++k) {
            HIPCHECK(c, hipStreamWaitEvent(cs, ev.done[k % 2], 0));
            if ((r = stage(k, (int)(k % 2), false))) return r;
        }
        return DMI_OK;
    };
    // on a failure both streams drain first, so no queued copy or kernel outlives the call
    if ((rc = pipeline())) {
        (void)hipStreamSynchronize(cs);
        (void)hipStreamSynchronize(c->stream);
        return rc;
    }
    HIPCHECK(c, hipStreamSynchronize(cs));
    HIPCHECK(c, hipEventRecord(c->ev[3], c->stream));
    HIPCHECK(c, hipEventSynchronize(c->ev[3]));
    std::vector<unsigned> counts((size_t)nch, 0u);
    unsigned long long tot[4] = {0, 0, 0, 0};
    HIPCHECK(c, hipMemcpy(counts.data(), cnt, sizeof(unsigned) * (size_t)nch, hipMemcpyDeviceToHost));
    HIPCHECK(c, hipMemcpy(tot, rb.counters, sizeof tot, hipMemcpyDeviceToHost));
    for (int64_t k = k0; k < nch; ++k) voxels += std::min<int64_t>(counts[(size_t)k], std::min<int64_t>(C, V - k * C));
    float kms = 0, tms = 0;
    if (k0 < nch) HIPCHECK(c, hipEventElapsedTime(&kms, c->ev[1], c->ev[2]));
    HIPCHECK(c, hipEventElapsedTime(&tms, c->ev[0], c->ev[3]));
    c->stats.voxels = voxels;
    c->stats.fp64_flops = (int64_t)tot[0];
    c->stats.nnls_solves = (int64_t)tot[1];
    c->stats.fp32_flops = (int64_t)tot[2];
    c->stats.voxels_failed = (int64_t)tot[3];
    c->stats.kernel_ms = kms0 + kms;
    c->stats.total_ms = tms;
    c->stats.lds_rows = std::min(sched.rl, m + n);
    c->stats.workgroups_per_cu = sched.per_cu;
    return DMI_OK;
}

// The allocation-free form: every buffer lives in the device's static scratch. The grids in it are keyed per
// device (they belong to whichever call last built them there).
std::mutex g_noalloc_mu;
dmi_host::GridKey g_noalloc_key[dmi::kNoAllocMaxDevices];

int run_t2map_noalloc(dmi_ctx *c, const double *d_image, const dmi_t2map_opts *o, const double *d_alpha_in,
                      const dmi_t2map_out *d) {
    int rc = dmi_host::validate_outputs(c->err, o, d);
    if (rc) return rc;
    if (!c->noalloc) return fail(c, DMI_EHIP, "static scratch address unavailable on this device");
    const int m = o->nTE, n = o->nT2, K = std::isnan(o->SetFlipAngle) ? o->nRefAngles : 1;
    if (m > dmi::kMaxEcho || n > dmi::kMaxT2 || K > dmi::kMaxAngles ||
        (int64_t)dmi::basis_ld(m) * n * K > dmi::kNoAllocBasis)
        return dmi_host::errorf(c->err, DMI_ECAPACITY,
                                "nTE = %d, nT2 = %d, nRefAngles = %d exceed the static scratch (at most %d x %d x %d)", m,
                                n, K, dmi::kMaxEcho, dmi::kMaxT2, dmi::kMaxAngles);
    double *S = c->noalloc;
    RunBufs rb{};
    rb.g.E2 = S + dmi::kNoAllocOffGrid;
    rb.g.logT2 = rb.g.E2 + 64;
    rb.g.alphas = rb.g.E2 + 128;
    rb.g.As = S + dmi::kNoAllocOffBasis;
    rb.g.dAs = rb.g.As + dmi::kNoAllocBasis;
    rb.g.AsT = rb.g.As + 2 * dmi::kNoAllocBasis;
    rb.g.dAsT = rb.g.As + 3 * dmi::kNoAllocBasis;
    rb.counters = reinterpret_cast<unsigned long long *>(S);       // S[0, 4)
    rb.status = reinterpret_cast<int *>(S + 4);                     // S[4, 6)
    rb.counter = reinterpret_cast<unsigned *>(S + 6);               // S[6, 8)
    rb.vox = reinterpret_cast<int64_t *>(S + dmi::kNoAllocOffVox);
    rb.vox_cap = dmi::kNoAllocVox;
    rb.slots = S + dmi::kNoAllocOffSlots;
    rb.slot_doubles = dmi::kNoAllocSlots;
    bool rec = false;
    if ((rc = capturing(c, rec))) return rc;
    if (!rec) HIPCHECK(c, hipEventRecord(c->ev[0], c->stream));
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu);
        const dmi_host::GridKey k = dmi_host::GridKey::of(o);
        dmi_host::GridKey &have = g_noalloc_key[c->device];
        if (!(have == k)) {
            // building the grids copies host tables, which a captured graph would read again at every replay
            if (rec)
                return dmi_host::errorf(c->err, DMI_EINVAL, "under graph capture the grids of these options must "
                                        "already be in the static scratch: make one call with them outside capture first");
            have = dmi_host::GridKey{};
            rc = dmi_host::build_grids(c->err, o, rb.g, c->stream);
            if (rc) return rc;
            have = k;
        }
    }
    return rec ? record_t2map_on(c, d_image, o, d_alpha_in, d, rb) : run_t2map_on(c, d_image, o, d_alpha_in, d, rb);
}

int run_t2part(dmi_ctx *c, const double *d_dist, const dmi_t2part_opts *o, double *sfr, double *sgm, double *mfr,
               double *mgm) {
    std::vector<double> lg, wts;
    dmi::T2partParams p{};
    int rc = dmi_host::t2part_prepare(c->err, o, lg, wts, p);
    if (rc) return rc;
    bool rec = false;
    if ((rc = capturing(c, rec))) return rc;
    if (rec) {
        // a captured graph reads the tables uploaded by an earlier call with the same options
        if (lg != c->tlog_host || wts != c->weights_host)
            return dmi_host::errorf(c->err, DMI_EINVAL, "under graph capture the T2 tables of these options must "
                                    "already be on the device: make one call with them outside capture first");
    } else {
        c->tlog_host.clear();
        c->weights_host.clear();
        HIPCHECK(c, c->tlog.ensure(sizeof(double) * p.n));
        HIPCHECK(c, hipMemcpyAsync(c->tlog.p, lg.data(), sizeof(double) * p.n, hipMemcpyHostToDevice, c->stream));
        if (!wts.empty()) {
            HIPCHECK(c, c->weights.ensure(sizeof(double) * p.n));
            HIPCHECK(c, hipMemcpyAsync(c->weights.p, wts.data(), sizeof(double) * p.n, hipMemcpyHostToDevice,
                                       c->stream));
        }
    }
    p.logT2 = c->tlog.as<double>();
    p.weights = wts.empty() ? nullptr : c->weights.as<double>();
    p.dist = d_dist; p.sfr = sfr; p.sgm = sgm; p.mfr = mfr; p.mgm = mgm;
    HIPCHECK(c, dmi::launch_t2part(p, c->stream));
    if (!rec) {
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        c->tlog_host = lg;
        c->weights_host = wts;
    }
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
    if (e != hipSuccess || ndev <= 0)
        return dmi_host::errorf(g_create_error, DMI_ENODEV, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev)
        return dmi_host::errorf(g_create_error, DMI_ENODEV, "device %d out of range (%d devices)", device, ndev);
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return dmi_host::errorf(g_create_error, DMI_EHIP, "hipGetDeviceProperties: %s", hipGetErrorString(e));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return dmi_host::errorf(g_create_error, DMI_ENODEV, "device %d is %s; this library is built for gfx950 only", device,
                                prop.gcnArchName);
    auto *c = new dmi_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    dmi_host::DeviceGuard guard(device);
    if ((e = guard.status) != hipSuccess || (e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete c;
        return dmi_host::errorf(g_create_error, DMI_EHIP, "stream creation failed: %s", hipGetErrorString(e));
    }
    c->own_stream = c->stream;
    for (auto &ev : c->ev) (void)hipEventCreate(&ev);
    if ((e = dmi::noalloc_scratch(device, &c->noalloc)) != hipSuccess) {
        for (auto &ev : c->ev) (void)hipEventDestroy(ev);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return dmi_host::errorf(g_create_error, DMI_EHIP, "static scratch lookup failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return DMI_OK;
}

int dmi_destroy(dmi_ctx *c) {
    if (!c) return DMI_OK;
    dmi_host::DeviceGuard guard(c->device);
    if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
    c->grids.release();
    c->lb.release();
    DevBuf *bufs[] = {&c->vox, &c->img, &c->ain, &c->weights, &c->tlog, &c->f32};
    for (auto *b : bufs) b->release();
    for (auto &b : c->out) b.release();
    for (auto &b : c->t2p) b.release();
    for (int k = 0; k < 2; ++k) {
        c->pimg[k].release();
        c->pain[k].release();
        for (auto &b : c->pout[k]) b.release();
    }
    c->pcnt.release();
    if (c->copy_stream) {
        (void)hipStreamSynchronize(c->copy_stream);
        (void)hipStreamDestroy(c->copy_stream);
    }
    for (auto &ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return DMI_OK;
}

int dmi_set_stream(dmi_ctx *c, void *stream) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    if (!stream) {
        c->stream = c->own_stream;
        return DMI_OK;
    }
    hipDevice_t dev = -1;
    HIPCHECK(c, hipStreamGetDevice(static_cast<hipStream_t>(stream), &dev));
    if (dev != c->device)
        return dmi_host::errorf(c->err, DMI_EINVAL, "the stream is on device %d, the context on device %d", (int)dev,
                                c->device);
    c->stream = static_cast<hipStream_t>(stream);
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
    DMI_ON_DEVICE(c->err, c->device);
    int rc = dmi_host::validate_t2map(c->err, o);
    if (rc) return rc;
    if (!d_image) return fail(c, DMI_EINVAL, "image is NULL");
    return run_t2map(c, d_image, o, d_alpha_in, d_out);
}

int dmi_t2map_f64_device_noalloc(dmi_ctx *c, const double *d_image, const dmi_t2map_opts *o,
                                 const double *d_alpha_in, const dmi_t2map_out *d_out) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    DMI_ON_DEVICE(c->err, c->device);
    int rc = dmi_host::validate_t2map(c->err, o);
    if (rc) return rc;
    if (!d_image) return fail(c, DMI_EINVAL, "image is NULL");
    return run_t2map_noalloc(c, d_image, o, d_alpha_in, d_out);
}

int dmi_t2map_f64(dmi_ctx *c, const double *image, const dmi_t2map_opts *o, const double *alpha_in,
                  const dmi_t2map_out *out) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    DMI_ON_DEVICE(c->err, c->device);
    if (int rc0 = no_capture(c, "dmi_t2map_f64")) return rc0;
    int rc = dmi_host::validate_t2map(c->err, o);
    if (rc) return rc;
    if (!image || !out) return fail(c, DMI_EINVAL, "image and out are required");
    const int m = o->nTE, n = o->nT2;
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    const bool fixed = !std::isnan(o->SetFlipAngle);
    if (V > kHostChunk && !(o->SaveNNLSBasis && !fixed)) {
        // copies overlapped with the voxel kernel, chunk by chunk; the per-voxel bases of SaveNNLSBasis take the
        // one-piece path below
        double *hp[12] = {out->dist, out->gdn, out->ggm, out->gva, out->fnr, out->snr, out->alpha,
                          o->SaveResidualNorm ? out->resnorm : nullptr, o->SaveDecayCurve ? out->decaycurve : nullptr,
                          o->SaveRegParam ? out->mu : nullptr, o->SaveRegParam ? out->chi2factor : nullptr, nullptr};
        const int64_t rows[12] = {n, 1, 1, 1, 1, 1, 1, 1, m, 1, 1, 0};
        rc = run_t2map_host_chunks(c, image, o, alpha_in, hp, rows);
        if (rc) return rc;
        if (fixed && o->SaveNNLSBasis)  // the shared nTE x nT2 basis, from its column-major copy in the grids
            HIPCHECK(c, hipMemcpy2D(out->decaybasis, sizeof(double) * m, c->grids.ptrs().AsT,
                                    sizeof(double) * dmi::basis_ld(m), sizeof(double) * m, n, hipMemcpyDeviceToHost));
        return DMI_OK;
    }
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

// Host Float32 array -> device FP64 (through the FP32 transfer buffer)
static int widen_in(dmi_ctx *c, const float *h, DevBuf &d, size_t n) {
    HIPCHECK(c, d.ensure(sizeof(double) * std::max<size_t>(n, 1)));
    HIPCHECK(c, c->f32.ensure(sizeof(float) * std::max<size_t>(n, 1)));
    HIPCHECK(c, hipMemcpyAsync(c->f32.p, h, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, dmi::launch_widen(c->f32.as<float>(), d.as<double>(), (int64_t)n, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));  // the transfer buffer is reused by the next array
    return DMI_OK;
}

// Device FP64 -> host Float32
static int narrow_out(dmi_ctx *c, const DevBuf &d, float *h, size_t n) {
    HIPCHECK(c, c->f32.ensure(sizeof(float) * std::max<size_t>(n, 1)));
    HIPCHECK(c, dmi::launch_narrow(d.as<double>(), c->f32.as<float>(), (int64_t)n, c->stream));
    HIPCHECK(c, hipMemcpyAsync(h, c->f32.p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return DMI_OK;
}

int dmi_t2map_f32(dmi_ctx *c, const float *image, const dmi_t2map_opts *o, const float *alpha_in,
                  const dmi_t2map_out_f32 *out) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    DMI_ON_DEVICE(c->err, c->device);
    if (int rc0 = no_capture(c, "dmi_t2map_f32")) return rc0;
    int rc = dmi_host::validate_t2map(c->err, o);
    if (rc) return rc;
    if (!image || !out) return fail(c, DMI_EINVAL, "image and out are required");
    const int m = o->nTE, n = o->nT2;
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    const bool fixed = !std::isnan(o->SetFlipAngle);
    if ((rc = widen_in(c, image, c->img, (size_t)V * m))) return rc;
    if (alpha_in && (rc = widen_in(c, alpha_in, c->ain, (size_t)V))) return rc;
    float *hp[12] = {out->dist, out->gdn, out->ggm, out->gva, out->fnr, out->snr, out->alpha,
                     o->SaveResidualNorm ? out->resnorm : nullptr, o->SaveDecayCurve ? out->decaycurve : nullptr,
                     o->SaveRegParam ? out->mu : nullptr, o->SaveRegParam ? out->chi2factor : nullptr,
                     o->SaveNNLSBasis ? out->decaybasis : nullptr};
    const size_t cnt[12] = {(size_t)V * n, (size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)V,
                            (size_t)V * m, (size_t)V, (size_t)V, fixed ? (size_t)m * n : (size_t)V * m * n};
    double *dp[12] = {nullptr};
    for (int i = 0; i < 12; ++i) {
        if (!hp[i]) continue;
        if ((rc = widen_in(c, hp[i], c->out[i], cnt[i]))) return rc;  // keeps the caller's values below threshold
        dp[i] = c->out[i].as<double>();
    }
    dmi_t2map_out d{dp[0], dp[1], dp[2], dp[3], dp[4], dp[5], dp[6], dp[7], dp[8], dp[9], dp[10], dp[11]};
    rc = run_t2map(c, c->img.as<double>(), o, alpha_in ? c->ain.as<double>() : nullptr, &d);
    if (rc && rc != DMI_ECAPACITY) return rc;
    for (int i = 0; i < 12; ++i) {
        if (!hp[i]) continue;
        int r2 = narrow_out(c, c->out[i], hp[i], cnt[i]);
        if (r2) return r2;
    }
    return rc;
}

int dmi_t2part_f32(dmi_ctx *c, const float *dist, const dmi_t2part_opts *o, float *sfr, float *sgm, float *mfr,
                   float *mgm) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    DMI_ON_DEVICE(c->err, c->device);
    if (int rc0 = no_capture(c, "dmi_t2part_f32")) return rc0;
    if (!o) return fail(c, DMI_EINVAL, "opts is NULL");
    if (!dist || !sfr || !sgm || !mfr || !mgm) return fail(c, DMI_EINVAL, "dist and all four outputs are required");
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    int rc = widen_in(c, dist, c->t2p[0], (size_t)V * std::max(o->nT2, 1));
    if (rc) return rc;
    float *hp[4] = {sfr, sgm, mfr, mgm};
    for (int i = 0; i < 4; ++i)
        if ((rc = widen_in(c, hp[i], c->t2p[1 + i], (size_t)V))) return rc;
    rc = run_t2part(c, c->t2p[0].as<double>(), o, c->t2p[1].as<double>(), c->t2p[2].as<double>(),
                    c->t2p[3].as<double>(), c->t2p[4].as<double>());
    if (rc) return rc;
    for (int i = 0; i < 4; ++i)
        if ((rc = narrow_out(c, c->t2p[1 + i], hp[i], (size_t)V))) return rc;
    return DMI_OK;
}

int dmi_epg_basis_set(dmi_ctx *c, const dmi_t2map_opts *o, double *out) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    DMI_ON_DEVICE(c->err, c->device);
    if (int rc0 = no_capture(c, "dmi_epg_basis_set")) return rc0;
    int rc = dmi_host::validate_t2map(c->err, o);
    if (rc) return rc;
    if (!out) return fail(c, DMI_EINVAL, "out is NULL");
    const bool search = std::isnan(o->SetFlipAngle);
    rc = c->grids.ensure(c->err, o, c->stream);
    if (rc) return rc;
    const int m = o->nTE, n = o->nT2, K = search ? o->nRefAngles : 1;
    std::vector<double> rm((size_t)m * n * K);
    HIPCHECK(c, hipMemcpy(rm.data(), c->grids.ptrs().As, sizeof(double) * rm.size(), hipMemcpyDeviceToHost));
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
    DMI_ON_DEVICE(c->err, c->device);
    if (!d_dist || !sfr || !sgm || !mfr || !mgm) return fail(c, DMI_EINVAL, "dist and all four outputs are required");
    return run_t2part(c, d_dist, o, sfr, sgm, mfr, mgm);
}

int dmi_t2part_f64(dmi_ctx *c, const double *dist, const dmi_t2part_opts *o, double *sfr, double *sgm, double *mfr,
                   double *mgm) {
    if (!c) return fail(nullptr, DMI_EINVAL, "ctx is NULL");
    std::lock_guard<std::mutex> lk(c->mu);
    DMI_ON_DEVICE(c->err, c->device);
    if (int rc0 = no_capture(c, "dmi_t2part_f64")) return rc0;
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
