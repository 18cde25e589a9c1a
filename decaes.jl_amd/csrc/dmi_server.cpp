// Request loop: continuous batching of T2mapSEcorr (+ T2partSEcorr) requests on one GPU.
//
// Two host threads drive the device. The scheduler admits queued requests in arrival order while their device
// footprint fits the HBM budget, binds each to a stream whose cached grid bases match its options (or rebinds an
// idle stream), selects its voxels, and keeps up to two voxel chunks in flight per stream, each from the stream's
// request with the fewest voxels left to launch (or from its requests in turn). The copy thread moves host arrays in and out (pageable copies block, so they stay
// off the scheduler) and frees a request's device memory once its outputs are home. Request memory is
// stream-ordered (hipMallocAsync / hipFreeAsync on the copy stream), so admission never waits on running kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "decaes_amd.h"
#include "dmi_host.h"
#include "dmi_internal.h"

namespace {

using Clock = std::chrono::steady_clock;
double ms_between(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); }

enum class St { Queued, Uploading, Uploaded, Selecting, Ready, Finishing, Downloading, Done };

constexpr int kOutputs = 12;

struct Request {
    int64_t id = 0;
    dmi_t2map_opts o{};
    bool has_part = false;
    dmi_t2part_opts po{};
    const double *h_image = nullptr, *h_alpha = nullptr;
    double *h_out[kOutputs] = {nullptr};
    double *h_part[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t cnt[kOutputs] = {0};  // element counts of the outputs (0 = not requested)
    int64_t V = 0, bytes = 0;
    // device arrays (stream-ordered allocations on the copy stream)
    double *d_image = nullptr, *d_alpha = nullptr, *d_out[kOutputs] = {nullptr}, *d_part[4] = {nullptr};
    double *d_tlog = nullptr, *d_wts = nullptr;
    int64_t *d_vox = nullptr;
    dmi::T2mapParams params{};
    dmi::T2partParams pp{};
    std::vector<double> tlog, wts;
    int stream = -1;
    St st = St::Queued;
    int64_t nvox = -1, launched = 0, finished = 0;
    bool tune = false;           // measure the launch schedules on this request's first voxels (none measured yet)
    int per_cu = 0;              // workgroups per CU of the request's chunks
    int64_t trial = 0;           // voxels run as schedule trials
    unsigned *h_nvox = nullptr;  // pinned
    hipEvent_t ev = nullptr;     // voxel selection done, later: request finished on its stream
    int status = DMI_OK;
    std::string err;
    Clock::time_point t_submit, t_admit, t_done;
};

struct Chunk {
    Request *r;
    int64_t count;
    hipEvent_t ev;
    int slot;
};

struct StreamState {
    hipStream_t s = nullptr;
    dmi_host::Grids grids;
    dmi_host::LaunchBufs lb;
    dmi_host::GridKey key;
    bool bound = false;
    std::vector<Request *> members;  // requests bound here and not yet finished on the device
    size_t rr = 0;
    std::deque<Chunk> inflight;
    int *h_status = nullptr;         // pinned, one slot per chunk in flight
    bool slot_busy[2] = {false, false};
};

constexpr int kMaxInflight = 2;
constexpr size_t kLatencyWindow = 65536;  // latencies of the most recent completed requests kept for p50/p99
constexpr size_t kErrorWindow = 256;      // error texts of waited-for failed requests kept for request_error

}  // namespace

struct dmi_server {
    int device = 0, num_cu = 0;
    dmi_server_config cfg{};
    hipStream_t copy = nullptr;
    std::vector<StreamState> streams;
    std::mutex mu;
    std::condition_variable cv_sched, cv_copy, cv_done;
    std::map<int64_t, std::shared_ptr<Request>> reqs;  // submitted and not yet returned by dmi_server_wait
    std::vector<Request *> active;    // admitted and not done (what the scheduler scans)
    std::deque<std::pair<int64_t, std::string>> waited_errors;  // recent failures already returned by wait
    std::deque<Request *> queued;     // waiting for admission
    std::deque<Request *> copy_jobs;  // uploads and downloads
    int64_t next_id = 1, used_bytes = 0, live = 0;
    int64_t submitted = 0, completed = 0, failed = 0, voxels = 0;
    std::vector<double> latencies;    // ring of the last kLatencyWindow latencies
    size_t lat_next = 0;
    double busy_ms = 0;
    Clock::time_point busy_since;
    bool stop = false;
    std::string err;
    std::thread sched, copier;
};

namespace {

// Caller holds S->mu.
void mark_done(dmi_server *S, Request *r) {
    r->st = St::Done;
    r->t_done = Clock::now();
    S->used_bytes -= r->bytes;
    if (--S->live == 0) S->busy_ms += ms_between(S->busy_since, r->t_done);
    if (r->status == DMI_OK) {
        ++S->completed;
        S->voxels += std::max<int64_t>(r->nvox, 0);
        const double lat = ms_between(r->t_submit, r->t_done);
        if (S->latencies.size() < kLatencyWindow) S->latencies.push_back(lat);
        else S->latencies[S->lat_next] = lat;
        S->lat_next = (S->lat_next + 1) % kLatencyWindow;
    } else {
        ++S->failed;
    }
    S->active.erase(std::remove(S->active.begin(), S->active.end(), r), S->active.end());
    S->cv_done.notify_all();
    S->cv_sched.notify_all();
    S->cv_copy.notify_all();
}

void fail_request(Request *r, int code, const std::string &msg) {
    if (r->status == DMI_OK) {
        r->status = code;
        r->err = msg;
    }
}

// Copy-thread side of a request's device memory.
void free_device(dmi_server *S, Request *r) {
    auto fr = [&](void *p) { if (p) (void)hipFreeAsync(p, S->copy); };
    fr(r->d_image); fr(r->d_alpha); fr(r->d_vox); fr(r->d_tlog); fr(r->d_wts);
    for (auto *p : r->d_out) fr(p);
    for (auto *p : r->d_part) fr(p);
    r->d_image = r->d_alpha = r->d_tlog = r->d_wts = nullptr;
    r->d_vox = nullptr;
    for (auto &p : r->d_out) p = nullptr;
    for (auto &p : r->d_part) p = nullptr;
    (void)hipStreamSynchronize(S->copy);
    if (r->h_nvox) { (void)hipHostFree(r->h_nvox); r->h_nvox = nullptr; }
    if (r->ev) { (void)hipEventDestroy(r->ev); r->ev = nullptr; }
}

int upload(dmi_server *S, Request *r) {
    std::string &e = r->err;
    hipStream_t s = S->copy;
    const size_t V = (size_t)r->V, m = (size_t)r->o.nTE;
    auto alloc = [&](double **p, size_t n) { return hipMallocAsync((void **)p, sizeof(double) * std::max<size_t>(n, 1), s); };
    DMI_HIP(e, alloc(&r->d_image, V * m));
    DMI_HIP(e, hipMemcpyAsync(r->d_image, r->h_image, sizeof(double) * V * m, hipMemcpyHostToDevice, s));
    if (r->h_alpha) {
        DMI_HIP(e, alloc(&r->d_alpha, V));
        DMI_HIP(e, hipMemcpyAsync(r->d_alpha, r->h_alpha, sizeof(double) * V, hipMemcpyHostToDevice, s));
    }
    // every output starts as the caller's array, so voxels at or below the threshold keep its contents
    for (int i = 0; i < kOutputs; ++i) {
        if (!r->cnt[i]) continue;
        DMI_HIP(e, alloc(&r->d_out[i], r->cnt[i]));
        DMI_HIP(e, hipMemcpyAsync(r->d_out[i], r->h_out[i], sizeof(double) * r->cnt[i], hipMemcpyHostToDevice, s));
    }
    if (r->has_part) {
        for (int i = 0; i < 4; ++i) {
            DMI_HIP(e, alloc(&r->d_part[i], V));
            DMI_HIP(e, hipMemcpyAsync(r->d_part[i], r->h_part[i], sizeof(double) * V, hipMemcpyHostToDevice, s));
        }
        DMI_HIP(e, alloc(&r->d_tlog, r->tlog.size()));
        DMI_HIP(e, hipMemcpyAsync(r->d_tlog, r->tlog.data(), sizeof(double) * r->tlog.size(), hipMemcpyHostToDevice, s));
        if (!r->wts.empty()) {
            DMI_HIP(e, alloc(&r->d_wts, r->wts.size()));
            DMI_HIP(e, hipMemcpyAsync(r->d_wts, r->wts.data(), sizeof(double) * r->wts.size(), hipMemcpyHostToDevice, s));
        }
    }
    DMI_HIP(e, hipMallocAsync((void **)&r->d_vox, sizeof(int64_t) * std::max<size_t>(V, 1), s));
    DMI_HIP(e, hipHostMalloc((void **)&r->h_nvox, sizeof(unsigned), hipHostMallocDefault));
    DMI_HIP(e, hipEventCreateWithFlags(&r->ev, hipEventDisableTiming));
    DMI_HIP(e, hipStreamSynchronize(s));
    return DMI_OK;
}

int download(dmi_server *S, Request *r) {
    std::string &e = r->err;
    hipStream_t s = S->copy;
    for (int i = 0; i < kOutputs; ++i)
        if (r->cnt[i]) DMI_HIP(e, hipMemcpyAsync(r->h_out[i], r->d_out[i], sizeof(double) * r->cnt[i], hipMemcpyDeviceToHost, s));
    if (r->has_part)
        for (int i = 0; i < 4; ++i)
            DMI_HIP(e, hipMemcpyAsync(r->h_part[i], r->d_part[i], sizeof(double) * (size_t)r->V, hipMemcpyDeviceToHost, s));
    DMI_HIP(e, hipStreamSynchronize(s));
    return DMI_OK;
}

void copy_loop(dmi_server *S) {
    (void)hipSetDevice(S->device);
    for (;;) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_copy.wait(lk, [&] { return !S->copy_jobs.empty() || (S->stop && S->live == 0 && S->queued.empty()); });
            if (S->copy_jobs.empty()) return;
            r = S->copy_jobs.front();
            S->copy_jobs.pop_front();
        }
        if (r->st == St::Uploading) {
            int rc = upload(S, r);
            if (rc) free_device(S, r);
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc) {
                fail_request(r, rc, r->err);
                mark_done(S, r);
            } else {
                r->st = St::Uploaded;
                S->cv_sched.notify_all();
            }
        } else if (r->st == St::Downloading) {
            int rc = r->status == DMI_OK || r->status == DMI_ECAPACITY ? download(S, r) : DMI_OK;
            free_device(S, r);
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc) fail_request(r, rc, r->err);
            mark_done(S, r);
        }
    }
}

// Scheduler-side helpers (called without S->mu; they touch only scheduler-owned state of r and st).
int bind_and_select(dmi_server *S, StreamState &st, Request *r) {
    std::string &e = r->err;
    int rc = st.grids.ensure(e, &r->o, st.s);
    if (rc) return rc;
    const bool fixed = !std::isnan(r->o.SetFlipAngle);
    if (fixed && r->o.SaveNNLSBasis)
        DMI_HIP(e, hipMemcpy2DAsync(r->d_out[11], sizeof(double) * r->o.nTE, st.grids.ptrs().AsT,
                                    sizeof(double) * dmi::basis_ld(r->o.nTE), sizeof(double) * r->o.nTE, r->o.nT2,
                                    hipMemcpyDeviceToDevice, st.s));
    // scratch for the largest grid any regularization method can launch with these sizes, so it never grows
    // while a chunk runs
    const int m = r->o.nTE, n = r->o.nT2;
    const int lda = dmi::t2map_lda(m, n);
    const size_t lds = dmi::t2map_lds_bytes(m, n, lda, fixed ? 1 : r->o.nRefAngles, r->o.legacy);  // as t2map_params sets K
    int grid = 0;
    for (int reg = DMI_REG_NONE; reg <= DMI_REG_MDP; ++reg) {
        int per_cu = 0;
        DMI_HIP(e, dmi::t2map_occupancy(reg, r->o.legacy, &per_cu, lds));
        grid = std::max(grid, per_cu * S->num_cu);
    }
    // the request's launch schedule (its explicit choice, or the one measured for these sizes, or all-LDS until the
    // first request of these sizes has measured the candidates, see tune_schedule)
    dmi_host::Schedule sched;
    rc = dmi_host::t2map_schedule(e, &r->o, sched);
    if (rc) return rc;
    std::vector<dmi_host::Schedule> cands;
    rc = dmi_host::t2map_schedules(e, &r->o, cands);
    if (rc) return rc;
    dmi_host::Schedule measured;
    r->tune = r->o.schedule == DMI_SCHED_AUTO && cands.size() > 1 && !dmi_host::schedule_lookup(&r->o, measured);
    size_t slots = (size_t)dmi::t2map_slot_doubles(m, n, 0) * std::max(grid, 1);
    for (const auto &c : cands) {  // room for every candidate, so a trial launch never grows the scratch either
        const int gld = c.rl >= m + n ? 0 : c.gld;
        slots = std::max(slots, (size_t)dmi::t2map_slot_doubles(m, n, gld) * std::max(c.per_cu * S->num_cu, 1));
    }
    DMI_HIP(e, st.lb.scratch.ensure(sizeof(double) * slots));
    DMI_HIP(e, st.lb.counter.ensure(sizeof(unsigned) * 4));
    DMI_HIP(e, st.lb.counters.ensure(sizeof(unsigned long long) * 4));
    DMI_HIP(e, st.lb.status.ensure(sizeof(int) * 4));
    DMI_HIP(e, dmi::launch_fill(st.lb.counter.p, sizeof(unsigned), 0u, st.s));
    DMI_HIP(e, dmi::launch_select(r->d_image, 0, r->V, r->o.Threshold, r->d_vox, st.lb.counter.as<unsigned>(), st.s));
    DMI_HIP(e, hipMemcpyAsync(r->h_nvox, st.lb.counter.p, sizeof(unsigned), hipMemcpyDeviceToHost, st.s));
    DMI_HIP(e, hipEventRecord(r->ev, st.s));
    dmi_t2map_out d{r->d_out[0], r->d_out[1], r->d_out[2], r->d_out[3], r->d_out[4], r->d_out[5],
                    r->d_out[6], r->d_out[7], r->d_out[8], r->d_out[9], r->d_out[10], r->d_out[11]};
    r->params = dmi_host::t2map_params(&r->o, st.grids.ptrs(), r->d_image, r->d_alpha, &d);
    dmi_host::apply_schedule(r->params, sched);
    r->per_cu = sched.per_cu;
    return DMI_OK;
}

// The autotuner of the one-call path (dmi_api.cpp), for the request loop: the first request of sizes with no
// measured schedule runs consecutive slices of its voxel list with each candidate schedule (8 voxels per resident
// workgroup each), timed with events on its stream, and the fastest is kept for the process; every schedule gives
// the same bits, so the trial voxels are simply done. The scheduler thread waits for the trials (once per set of
// sizes); chunks already queued on other streams keep running meanwhile.
int tune_schedule(dmi_server *S, StreamState &st, Request *r) {
    std::string &e = r->err;
    std::vector<dmi_host::Schedule> cands;
    int rc = dmi_host::t2map_schedules(e, &r->o, cands);
    if (rc) return rc;
    int64_t need = 0;
    for (const auto &c : cands) need += 8ll * S->num_cu * c.per_cu;
    if (r->nvox < 2 * need) return DMI_OK;  // too small to measure: all-LDS for this request
    hipEvent_t e0 = nullptr, e1 = nullptr;
    DMI_HIP(e, hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return dmi_host::errorf(e, DMI_EHIP, "hipEventCreate failed");
    }
    dmi_host::Schedule best = cands.front();
    double best_rate = -1.0;
    int64_t done = 0;
    for (const auto &c : cands) {
        const int64_t cnt = 8ll * S->num_cu * c.per_cu;
        dmi::T2mapParams ps = r->params;
        dmi_host::apply_schedule(ps, c);
        float ms = 0.0f;
        if (hipEventRecord(e0, st.s) != hipSuccess ||
            (rc = dmi_host::t2map_launch(e, ps, r->d_vox + done, cnt, st.lb, S->num_cu, st.s)) != DMI_OK ||
            hipEventRecord(e1, st.s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc ? rc : dmi_host::errorf(e, DMI_EHIP, "schedule trial failed");
        }
        const double rate = cnt / std::max(ms, 1e-3f);
        if (rate > best_rate) { best_rate = rate; best = c; }
        done += cnt;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    dmi_host::schedule_store(&r->o, best);
    dmi_host::apply_schedule(r->params, best);
    r->per_cu = best.per_cu;
    r->trial = done;
    r->launched += done;
    r->finished += done;
    return DMI_OK;
}

int launch_chunk(dmi_server *S, StreamState &st, Request *r, int64_t count, int slot) {
    std::string &e = r->err;
    DMI_HIP(e, dmi::launch_fill(st.lb.status.p, sizeof(int), 0u, st.s));
    int rc = dmi_host::t2map_launch(e, r->params, r->d_vox + r->launched, count, st.lb, S->num_cu, st.s);
    if (rc) return rc;
    DMI_HIP(e, hipMemcpyAsync(&st.h_status[slot], st.lb.status.p, sizeof(int), hipMemcpyDeviceToHost, st.s));
    return DMI_OK;
}

int launch_finish(StreamState &st, Request *r) {
    std::string &e = r->err;
    if (r->has_part) {
        dmi::T2partParams p = r->pp;
        p.dist = r->d_out[0];
        p.logT2 = r->d_tlog;
        p.weights = r->d_wts;
        p.sfr = r->d_part[0]; p.sgm = r->d_part[1]; p.mfr = r->d_part[2]; p.mgm = r->d_part[3];
        DMI_HIP(e, dmi::launch_t2part(p, st.s));
    }
    DMI_HIP(e, hipEventRecord(r->ev, st.s));
    return DMI_OK;
}

void sched_loop(dmi_server *S) {
    (void)hipSetDevice(S->device);
    const int64_t budget = S->cfg.hbm_budget_bytes;
    for (;;) {
        bool progress = false;
        std::vector<Request *> ready_to_bind, selecting, finishing;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->stop && S->live == 0 && S->queued.empty()) return;
            // admission in arrival order
            while (!S->queued.empty()) {
                Request *r = S->queued.front();
                if (r->bytes > budget) {
                    S->queued.pop_front();
                    S->active.push_back(r);
                    fail_request(r, DMI_ECAPACITY, "request needs " + std::to_string(r->bytes) +
                                                       " bytes of device memory, more than the budget of " + std::to_string(budget));
                    ++S->live;
                    if (S->live == 1) S->busy_since = Clock::now();
                    S->used_bytes += r->bytes;
                    mark_done(S, r);
                    continue;
                }
                if (S->used_bytes + r->bytes > budget) break;
                S->queued.pop_front();
                S->active.push_back(r);
                S->used_bytes += r->bytes;
                r->t_admit = Clock::now();
                if (S->live++ == 0) S->busy_since = r->t_admit;
                r->st = St::Uploading;
                S->copy_jobs.push_back(r);
                S->cv_copy.notify_all();
                progress = true;
            }
            for (Request *r : S->active) {
                if (r->st == St::Uploaded) ready_to_bind.push_back(r);
                else if (r->st == St::Selecting) selecting.push_back(r);
                else if (r->st == St::Finishing) finishing.push_back(r);
            }
        }
        // bind uploaded requests to streams (matching grids first, else an idle stream)
        for (Request *r : ready_to_bind) {
            const dmi_host::GridKey k = dmi_host::GridKey::of(&r->o);
            int pick = -1;
            for (size_t i = 0; i < S->streams.size() && pick < 0; ++i)
                if (S->streams[i].bound && S->streams[i].key == k) pick = (int)i;
            for (size_t i = 0; i < S->streams.size() && pick < 0; ++i)
                if (S->streams[i].members.empty() && S->streams[i].inflight.empty()) pick = (int)i;
            if (pick < 0) continue;
            StreamState &st = S->streams[pick];
            st.key = k;
            st.bound = true;
            int rc = bind_and_select(S, st, r);
            std::lock_guard<std::mutex> lk(S->mu);
            if (rc) {
                fail_request(r, rc, r->err);
                r->st = St::Downloading;  // nothing to copy back; frees the device memory
                S->copy_jobs.push_back(r);
                S->cv_copy.notify_all();
            } else {
                r->stream = pick;
                r->st = St::Selecting;
                st.members.push_back(r);
            }
            progress = true;
        }
        for (Request *r : selecting) {
            if (hipEventQuery(r->ev) != hipSuccess) continue;
            r->nvox = *r->h_nvox;
            progress = true;
            if (r->nvox == 0) {
                StreamState &st = S->streams[r->stream];
                int rc = launch_finish(st, r);
                std::lock_guard<std::mutex> lk(S->mu);
                if (rc) fail_request(r, rc, r->err);
                r->st = St::Finishing;
            } else {
                std::lock_guard<std::mutex> lk(S->mu);
                r->st = St::Ready;
            }
        }
        for (StreamState &st : S->streams) {
            // retire finished chunks in launch order
            while (!st.inflight.empty() && hipEventQuery(st.inflight.front().ev) == hipSuccess) {
                Chunk c = st.inflight.front();
                st.inflight.pop_front();
                (void)hipEventDestroy(c.ev);
                st.slot_busy[c.slot] = false;
                Request *r = c.r;
                r->finished += c.count;
                if (r->finished == r->nvox) {
                    int rc = launch_finish(st, r);
                    std::lock_guard<std::mutex> lk(S->mu);
                    if (rc) fail_request(r, rc, r->err);
                    r->st = St::Finishing;
                }
                progress = true;
            }
            // keep up to kMaxInflight chunks queued: the stream's request with the fewest voxels left to launch, or
            // its requests in turn (dmi_server_config.order)
            while ((int)st.inflight.size() < kMaxInflight && !st.members.empty()) {
                Request *r = nullptr;
                if (S->cfg.order == DMI_ORDER_TURNS) {
                    for (size_t t = 0; t < st.members.size(); ++t) {
                        Request *c = st.members[(st.rr + t) % st.members.size()];
                        if (c->st == St::Ready && c->launched < c->nvox) {
                            r = c;
                            st.rr = (st.rr + t + 1) % st.members.size();
                            break;
                        }
                    }
                } else {
                    for (Request *c : st.members)
                        if (c->st == St::Ready && c->launched < c->nvox &&
                            (!r || c->nvox - c->launched < r->nvox - r->launched))
                            r = c;
                }
                if (!r) break;
                if (r->tune) {
                    r->tune = false;
                    int rc = tune_schedule(S, st, r);
                    if (rc) {
                        std::lock_guard<std::mutex> lk(S->mu);
                        fail_request(r, rc, r->err);
                    }
                    progress = true;
                    if (rc || r->launched >= r->nvox) {
                        // a failed trial fails the request; either way nothing more is launched for it here
                        r->finished += r->nvox - r->launched;
                        r->launched = r->nvox;
                        if (r->finished == r->nvox) {
                            (void)launch_finish(st, r);
                            std::lock_guard<std::mutex> lk(S->mu);
                            r->st = St::Finishing;
                        }
                        continue;
                    }
                }
                const int slot = st.slot_busy[0] ? 1 : 0;
                const int64_t count = std::min<int64_t>(S->cfg.chunk_voxels, r->nvox - r->launched);
                hipEvent_t ev = nullptr;
                int rc = launch_chunk(S, st, r, count, slot);
                if (!rc && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess &&
                    hipEventRecord(ev, st.s) == hipSuccess) {
                    st.slot_busy[slot] = true;
                    st.inflight.push_back(Chunk{r, count, ev, slot});
                    r->launched += count;
                } else {
                    // a launch that cannot be queued fails the request; its chunks already queued still drain
                    if (ev) (void)hipEventDestroy(ev);
                    {
                        std::lock_guard<std::mutex> lk(S->mu);
                        fail_request(r, rc ? rc : DMI_EHIP, rc ? r->err : std::string("event record failed"));
                    }
                    r->finished += r->nvox - r->launched;
                    r->launched = r->nvox;
                    if (r->finished == r->nvox) {
                        (void)launch_finish(st, r);
                        std::lock_guard<std::mutex> lk(S->mu);
                        r->st = St::Finishing;
                    }
                }
                progress = true;
            }
        }
        for (Request *r : finishing) {
            if (hipEventQuery(r->ev) != hipSuccess) continue;
            StreamState &st = S->streams[r->stream];
            st.members.erase(std::remove(st.members.begin(), st.members.end(), r), st.members.end());
            std::lock_guard<std::mutex> lk(S->mu);
            r->st = St::Downloading;
            S->copy_jobs.push_back(r);
            S->cv_copy.notify_all();
            progress = true;
        }
        if (!progress) {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_sched.wait_for(lk, std::chrono::microseconds(100));
        }
    }
}

int serr(dmi_server *S, int code, const char *msg) {
    if (S) S->err = msg;
    return code;
}

}  // namespace

extern "C" {

int dmi_server_create(dmi_server **out, int device, const dmi_server_config *cfg) {
    if (!out) return DMI_EINVAL;
    *out = nullptr;
    if (cfg && cfg->order != DMI_ORDER_SHORTEST && cfg->order != DMI_ORDER_TURNS) return DMI_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) return DMI_ENODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return DMI_EHIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return DMI_ENODEV;
    dmi_host::DeviceGuard guard(device);
    if (guard.status != hipSuccess) return DMI_EHIP;
    auto S = std::make_unique<dmi_server>();
    S->device = device;
    S->num_cu = prop.multiProcessorCount;
    if (cfg) S->cfg = *cfg;
    if (S->cfg.max_streams <= 0) S->cfg.max_streams = 4;
    if (S->cfg.chunk_voxels <= 0) S->cfg.chunk_voxels = 65536;
    if (S->cfg.hbm_budget_bytes <= 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return DMI_EHIP;
        S->cfg.hbm_budget_bytes = (int64_t)(fr / 2);
    }
    // on a failure, the streams and pinned slots created so far are released before returning
    auto undo = [&S]() {
        for (auto &st : S->streams) {
            if (st.h_status) (void)hipHostFree(st.h_status);
            if (st.s) (void)hipStreamDestroy(st.s);
        }
        if (S->copy) (void)hipStreamDestroy(S->copy);
        return DMI_EHIP;
    };
    if (hipStreamCreateWithFlags(&S->copy, hipStreamNonBlocking) != hipSuccess) return undo();
    S->streams.resize(S->cfg.max_streams);
    for (auto &st : S->streams) {
        if (hipStreamCreateWithFlags(&st.s, hipStreamNonBlocking) != hipSuccess) return undo();
        if (hipHostMalloc((void **)&st.h_status, sizeof(int) * kMaxInflight, hipHostMallocDefault) != hipSuccess)
            return undo();
    }
    dmi_server *raw = S.release();
    raw->sched = std::thread(sched_loop, raw);
    raw->copier = std::thread(copy_loop, raw);
    *out = raw;
    return DMI_OK;
}

int dmi_server_destroy(dmi_server *S) {
    if (!S) return DMI_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_sched.notify_all();
        S->cv_copy.notify_all();
    }
    if (S->sched.joinable()) S->sched.join();
    if (S->copier.joinable()) S->copier.join();
    dmi_host::DeviceGuard guard(S->device);
    for (auto &st : S->streams) {
        if (st.s) (void)hipStreamSynchronize(st.s);
        for (auto &c : st.inflight) (void)hipEventDestroy(c.ev);
        st.grids.release();
        st.lb.release();
        if (st.h_status) (void)hipHostFree(st.h_status);
        if (st.s) (void)hipStreamDestroy(st.s);
    }
    if (S->copy) (void)hipStreamDestroy(S->copy);
    delete S;
    return DMI_OK;
}

int dmi_server_submit(dmi_server *S, const double *image, const dmi_t2map_opts *o, const double *alpha_in,
                      const dmi_t2map_out *out, const dmi_t2part_opts *popts, double *sfr, double *sgm, double *mfr,
                      double *mgm, int64_t *id) {
    if (!S) return DMI_EINVAL;
    if (!id) return serr(S, DMI_EINVAL, "id is NULL");
    std::string e;
    int rc = dmi_host::validate_t2map(e, o);
    if (!rc) rc = dmi_host::validate_outputs(e, o, out);
    if (!rc && !image) rc = dmi_host::errorf(e, DMI_EINVAL, "image is NULL");
    auto r = std::make_shared<Request>();
    if (!rc && popts) {
        if (!sfr || !sgm || !mfr || !mgm) rc = dmi_host::errorf(e, DMI_EINVAL, "T2part outputs sfr, sgm, mfr and mgm are required");
        else if (popts->nT2 != o->nT2 || popts->MatrixSize[0] != o->MatrixSize[0] || popts->MatrixSize[1] != o->MatrixSize[1] ||
                 popts->MatrixSize[2] != o->MatrixSize[2])
            rc = dmi_host::errorf(e, DMI_EINVAL, "T2partOptions MatrixSize and nT2 must match the T2mapOptions");
        else rc = dmi_host::t2part_prepare(e, popts, r->tlog, r->wts, r->pp);
    }
    if (rc) {
        std::lock_guard<std::mutex> lk(S->mu);
        S->err = e;
        return rc;
    }
    r->o = *o;
    r->has_part = popts != nullptr;
    if (popts) r->po = *popts;
    r->h_image = image;
    r->h_alpha = alpha_in;
    const int64_t V = o->MatrixSize[0] * o->MatrixSize[1] * o->MatrixSize[2];
    const size_t m = o->nTE, n = o->nT2;
    const bool fixed = !std::isnan(o->SetFlipAngle);
    r->V = V;
    double *hp[kOutputs] = {out->dist, out->gdn, out->ggm, out->gva, out->fnr, out->snr, out->alpha,
                            o->SaveResidualNorm ? out->resnorm : nullptr, o->SaveDecayCurve ? out->decaycurve : nullptr,
                            o->SaveRegParam ? out->mu : nullptr, o->SaveRegParam ? out->chi2factor : nullptr,
                            o->SaveNNLSBasis ? out->decaybasis : nullptr};
    const size_t cnt[kOutputs] = {V * n, (size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)V, (size_t)V,
                                  V * m, (size_t)V, (size_t)V, fixed ? m * n : V * m * n};
    for (int i = 0; i < kOutputs; ++i) {
        r->h_out[i] = hp[i];
        r->cnt[i] = hp[i] ? cnt[i] : 0;
    }
    if (popts) { r->h_part[0] = sfr; r->h_part[1] = sgm; r->h_part[2] = mfr; r->h_part[3] = mgm; }
    r->bytes = dmi_host::t2map_request_bytes(o, alpha_in != nullptr, popts != nullptr);
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    r->id = S->next_id++;
    *id = r->id;
    Request *raw = r.get();
    S->reqs[raw->id] = std::move(r);
    S->queued.push_back(raw);
    ++S->submitted;
    S->cv_sched.notify_all();
    return DMI_OK;
}

int dmi_server_wait(dmi_server *S, int64_t id, dmi_request_info *info) {
    if (!S) return DMI_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) { S->err = "unknown request id"; return DMI_EINVAL; }
    std::shared_ptr<Request> r = it->second;  // another thread's wait may forget the entry first
    S->cv_done.wait(lk, [&] { return r->st == St::Done; });
    if (info) {
        info->status = r->status;
        info->voxels = std::max<int64_t>(r->nvox, 0);
        info->queue_ms = r->t_admit.time_since_epoch().count() ? ms_between(r->t_submit, r->t_admit) : 0.0;
        info->latency_ms = ms_between(r->t_submit, r->t_done);
        info->workgroups_per_cu = r->per_cu;
        info->trial_voxels = r->trial;
    }
    // the request is returned: forget it (its error text stays available for a while)
    const int status = r->status;
    if (status != DMI_OK) {
        S->waited_errors.emplace_back(id, r->err);
        if (S->waited_errors.size() > kErrorWindow) S->waited_errors.pop_front();
    }
    S->reqs.erase(id);
    return status;
}

int dmi_server_poll(dmi_server *S, int64_t id, int32_t *done) {
    if (!S || !done) return DMI_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) {
        if (id > 0 && id < S->next_id) { *done = 1; return DMI_OK; }  // already returned by dmi_server_wait
        S->err = "unknown request id";
        return DMI_EINVAL;
    }
    *done = it->second->st == St::Done;
    return DMI_OK;
}

int dmi_server_get_stats(dmi_server *S, dmi_server_stats *out) {
    if (!S || !out) return DMI_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    *out = dmi_server_stats{};
    out->submitted = S->submitted;
    out->completed = S->completed;
    out->failed = S->failed;
    out->voxels = S->voxels;
    out->busy_ms = S->busy_ms + (S->live > 0 ? ms_between(S->busy_since, Clock::now()) : 0.0);
    std::vector<double> l = S->latencies;
    if (!l.empty()) {
        std::sort(l.begin(), l.end());
        auto q = [&](double f) {
            double pos = f * (l.size() - 1);
            size_t i = (size_t)pos;
            double t = pos - i;
            return i + 1 < l.size() ? l[i] * (1 - t) + l[i + 1] * t : l[i];
        };
        out->p50_ms = q(0.50);
        out->p99_ms = q(0.99);
    }
    return DMI_OK;
}

const char *dmi_server_last_error(const dmi_server *S) { return S ? S->err.c_str() : "server is NULL"; }

const char *dmi_server_request_error(dmi_server *S, int64_t id) {
    if (!S) return "server is NULL";
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it != S->reqs.end()) return it->second->err.c_str();
    for (auto w = S->waited_errors.rbegin(); w != S->waited_errors.rend(); ++w)
        if (w->first == id) return w->second.c_str();
    return "unknown request id";
}

}  // extern "C"
