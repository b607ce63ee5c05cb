// SGM serving loop behind include/vxh/sgm_server.h.
//
// Three parts, so that no host copy ever sits between a completion and the
// next admission:
//   submit (client threads)  copies the pair into a pinned staging buffer
//                            from a pool, takes a pinned result buffer from
//                            the same pool and queues the request;
//   scheduler thread         owns the device side: it polls in-flight
//                            requests for completion, then admits queued
//                            requests in arrival order while the sum of
//                            admitted cost-volume bytes (each request's
//                            pipeline workspace: the per-direction volumes)
//                            stays within the HBM budget, fewer than
//                            max_inflight run and (with a work budget) the
//                            admitted pixel-disparities stay within it, so
//                            small requests overlap while a large one runs
//                            nearly alone.  An admitted request takes a
//                            device-resident slot (inputs, output, census +
//                            volume scratch) from a pool of slots reused by
//                            capacity, as the reference's TexturePool reuses
//                            textures (vx-vision/src/pool.rs:61-97), and runs
//                            H2D copy -> the per-pair SGM pipeline -> D2H
//                            copies of the error word and the map on one of
//                            several HIP streams, so requests that arrive
//                            while others are on the GPU overlap with them;
//                            a completed
//                            request's slot goes back to the pool at once;
//   finisher threads         copy each completed map to the caller's
//                            buffer, check the request's error word, and
//                            return the staging buffers.
// Work never changes with the mix: each request runs the same kernels on its
// own buffers, so its result is the one it would get alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vxh/sgm.h"
#include "vxh/sgm_server.h"

namespace {

using Clock = std::chrono::steady_clock;
double now_s() {
    return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

// Device-resident buffers of one request at a time, reusable by any request
// of at most npx_cap pixels and ws_cap workspace bytes.
struct Slot {
    size_t npx_cap = 0;
    size_t ws_cap = 0;
    uint8_t *d_in = nullptr;    // left then right
    float *d_out = nullptr;
    void *d_ws = nullptr;
    size_t dev_bytes = 0;
    double last_used = 0;
};

// Pinned host buffer: a request's left || right pixels, or its map followed
// (at byte out_err_offset) by its error word.
struct Staging {
    uint8_t *ptr = nullptr;
    size_t cap = 0;
};

struct Request {
    uint64_t id = 0;
    uint32_t w = 0, h = 0;
    vxh_sgm_config cfg{};
    Staging in, res;
    float *out = nullptr;
    uint64_t vol_bytes = 0;
    size_t ws_bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    Slot *slot = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool done = false;
    vxh_status status = VXH_OK;
};

thread_local std::string g_srv_error;
vxh_status sfail(vxh_status st, const std::string &m) {
    g_srv_error = m;
    return st;
}

size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Result staging of an npx-pixel map: the map, then the error word.
size_t out_err_offset(size_t npx) { return round_up(npx * 4, 64); }

}  // namespace

// The work a request adds to the GPU: its pixel-disparities.
inline uint64_t request_pxd(const Request *r) {
    return (uint64_t)r->w * r->h * r->cfg.num_disparities;
}

struct vxh_server {
    int device = 0;
    vxh_server_config cfg{};
    vxh_context *ctx = nullptr;
    vxh_sgm *sgm = nullptr;
    std::vector<hipStream_t> streams;
    size_t rr = 0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;                  // guarded by mu
    std::unordered_map<uint64_t, Request *> by_id;  // guarded by mu
    uint64_t next_id = 1;
    bool stop = false;

    // Scheduler-thread state.
    std::vector<Request *> inflight;
    uint64_t admitted_bytes = 0;
    uint64_t inflight_pxd = 0;   // W*H*D of the admitted requests
    std::vector<hipEvent_t> events;

    // Completed requests waiting for the finisher (guarded by mu).
    std::deque<Request *> finished;
    std::condition_variable cv_finish;
    bool stop_finisher = false;

    // Slot pool (guarded by pool_mu): idle slots and the device bytes of all.
    std::mutex pool_mu;
    std::vector<Slot *> idle;
    uint64_t pool_bytes = 0;

    // Pinned staging pool (guarded by pin_mu).
    std::mutex pin_mu;
    std::vector<Staging> pin_idle;

    std::thread th;
    std::vector<std::thread> fins;   // finisher threads

    // Statistics, guarded by mu.
    vxh_server_stats st{};
    std::vector<double> lat_ms, queue_ms;
};

namespace {

void free_slot(Slot *s) {
    if (s->d_in) (void)hipFree(s->d_in);
    if (s->d_out) (void)hipFree(s->d_out);
    if (s->d_ws) (void)hipFree(s->d_ws);
    delete s;
}

// A new slot for npx pixels and ws workspace bytes (nullptr when out of
// memory).
Slot *new_slot(size_t npx, size_t ws) {
    Slot *s = new Slot();
    s->npx_cap = npx;
    s->ws_cap = ws;
    s->dev_bytes = 2 * npx + npx * 4 + ws;
    bool ok = hipMalloc(&s->d_in, 2 * npx) == hipSuccess &&
              hipMalloc(&s->d_out, npx * 4) == hipSuccess &&
              hipMalloc(&s->d_ws, ws) == hipSuccess;
    if (!ok) {
        free_slot(s);
        return nullptr;
    }
    return s;
}

// Takes the smallest idle slot that fits the request, or allocates one,
// first releasing least recently used idle slots when the pool would exceed
// its cap (1.5 x the budget).
Slot *acquire_slot(vxh_server *S, Request *r, bool *hit) {
    const size_t npx = (size_t)r->w * r->h;
    {
        std::lock_guard<std::mutex> g(S->pool_mu);
        size_t best = S->idle.size();
        for (size_t i = 0; i < S->idle.size(); ++i) {
            const Slot *s = S->idle[i];
            if (s->npx_cap >= npx && s->ws_cap >= r->ws_bytes &&
                (best == S->idle.size() || s->dev_bytes < S->idle[best]->dev_bytes))
                best = i;
        }
        if (best < S->idle.size()) {
            Slot *s = S->idle[best];
            S->idle.erase(S->idle.begin() + (long)best);
            *hit = true;
            return s;
        }
        *hit = false;
        const size_t need = 2 * npx + npx * 4 + r->ws_bytes;
        const uint64_t cap = S->cfg.hbm_budget_bytes + S->cfg.hbm_budget_bytes / 2;
        std::sort(S->idle.begin(), S->idle.end(),
                  [](Slot *a, Slot *b) { return a->last_used < b->last_used; });
        while (!S->idle.empty() && S->pool_bytes + need > cap) {
            Slot *v = S->idle.front();
            S->idle.erase(S->idle.begin());
            S->pool_bytes -= v->dev_bytes;
            free_slot(v);
        }
    }
    Slot *s = new_slot(npx, r->ws_bytes);
    if (s) {
        std::lock_guard<std::mutex> g(S->pool_mu);
        S->pool_bytes += s->dev_bytes;
    }
    return s;
}

void release_slot(vxh_server *S, Slot *s, double t) {
    std::lock_guard<std::mutex> g(S->pool_mu);
    s->last_used = t;
    S->idle.push_back(s);
}

bool take_staging(vxh_server *S, size_t bytes, Staging *out) {
    {
        std::lock_guard<std::mutex> g(S->pin_mu);
        size_t best = S->pin_idle.size();
        for (size_t i = 0; i < S->pin_idle.size(); ++i)
            if (S->pin_idle[i].cap >= bytes &&
                (best == S->pin_idle.size() || S->pin_idle[i].cap < S->pin_idle[best].cap))
                best = i;
        if (best < S->pin_idle.size()) {
            *out = S->pin_idle[best];
            S->pin_idle.erase(S->pin_idle.begin() + (long)best);
            return true;
        }
    }
    Staging s;
    s.cap = round_up(bytes, (size_t)1 << 20);
    if (hipHostMalloc((void **)&s.ptr, s.cap, hipHostMallocDefault) != hipSuccess) return false;
    *out = s;
    return true;
}

void give_staging(vxh_server *S, Staging s) {
    if (!s.ptr) return;
    std::lock_guard<std::mutex> g(S->pin_mu);
    S->pin_idle.push_back(s);
}

hipEvent_t take_event(vxh_server *S) {
    if (!S->events.empty()) {
        hipEvent_t e = S->events.back();
        S->events.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Enqueues one admitted request on a stream: inputs in, the pipeline, the
// error word and the map out, then the completion event.  Returns false when
// the device work could not be enqueued; r->stream then names the stream that
// may hold part of it.
bool launch(vxh_server *S, Request *r, bool busy) {
    Slot *s = r->slot;
    const size_t npx = (size_t)r->w * r->h;
    // Without a backlog the requests queue on the first stream: each one's
    // copies and kernels are enqueued behind the running one (no idle gap
    // between them) but never share the GPU with it, which keeps a large
    // request's latency at its own.  With a backlog they spread over every
    // stream and overlap.
    hipStream_t st = busy || S->cfg.backlog_requests == 0
                         ? S->streams[S->rr++ % S->streams.size()]
                         : S->streams[0];
    r->ev = take_event(S);
    if (!r->ev) return false;
    r->stream = st;
    if (hipMemcpyAsync(s->d_in, r->in.ptr, 2 * npx, hipMemcpyHostToDevice, st) != hipSuccess)
        return false;
    if (vxh_sgm_enqueue(S->sgm, s->d_in, s->d_in + npx, r->w, r->h, &r->cfg, s->d_out, s->d_ws,
                        s->ws_cap, st) != VXH_OK)
        return false;
    if (hipMemcpyAsync(r->res.ptr + out_err_offset(npx),
                       (const char *)s->d_ws + vxh_sgm_workspace_error_offset(), 4,
                       hipMemcpyDeviceToHost, st) != hipSuccess)
        return false;
    if (hipMemcpyAsync(r->res.ptr, s->d_out, npx * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        return false;
    return hipEventRecord(r->ev, st) == hipSuccess;
}

// Hands a request whose device work has ended (or never started) to the
// finisher; its slot goes back to the pool now.
void complete(vxh_server *S, Request *r, vxh_status status) {
    r->status = status;
    if (r->ev) S->events.push_back(r->ev);
    r->ev = nullptr;
    S->admitted_bytes -= r->vol_bytes;
    S->inflight_pxd -= request_pxd(r);
    if (r->slot) release_slot(S, r->slot, now_s());
    r->slot = nullptr;
    std::lock_guard<std::mutex> g(S->mu);
    S->finished.push_back(r);
    S->cv_finish.notify_one();
}

// Copies results out, checks error words, recycles buffers, records stats.
// Finisher threads: a map's copy into the caller's buffer (pageable memory,
// up to 33 MB for 4K) takes milliseconds when its pages are new, so requests
// that complete together are copied side by side instead of one after the
// other.
constexpr int kFinishers = 4;

void finisher(vxh_server *S) {
    (void)hipSetDevice(S->device);
    while (true) {
        Request *r = nullptr;
        {
            std::unique_lock<std::mutex> g(S->mu);
            S->cv_finish.wait(g, [&] { return S->stop_finisher || !S->finished.empty(); });
            if (S->finished.empty()) break;
            r = S->finished.front();
            S->finished.pop_front();
        }
        const size_t npx = (size_t)r->w * r->h;
        vxh_status status = r->status;
        if (status == VXH_OK) {
            if (*(const uint32_t *)(r->res.ptr + out_err_offset(npx)) != 0) status = VXH_ERR_GPU;
            else std::memcpy(r->out, r->res.ptr, npx * 4);
        }
        r->t_done = now_s();
        give_staging(S, r->in);
        give_staging(S, r->res);
        r->in = r->res = Staging{};
        std::lock_guard<std::mutex> g(S->mu);
        r->done = true;
        r->status = status;
        S->st.completed++;
        const double lat = (r->t_done - r->t_submit) * 1e3;
        S->lat_ms.push_back(lat);
        S->queue_ms.push_back((r->t_admit - r->t_submit) * 1e3);
        S->st.pxd_total += (double)r->w * r->h * r->cfg.num_disparities;
        S->st.last_complete_s = std::max(S->st.last_complete_s, r->t_done);
        S->cv_done.notify_all();
    }
}

void scheduler(vxh_server *S) {
    (void)hipSetDevice(S->device);
    while (true) {
        bool progressed = false;
        // Completions.
        for (size_t i = 0; i < S->inflight.size();) {
            Request *r = S->inflight[i];
            hipError_t q = hipEventQuery(r->ev);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            S->inflight.erase(S->inflight.begin() + (long)i);
            complete(S, r, q == hipSuccess ? VXH_OK : VXH_ERR_GPU);
            progressed = true;
        }
        // Admissions, in arrival order.
        while (true) {
            Request *r = nullptr;
            bool busy = false;
            {
                std::unique_lock<std::mutex> g(S->mu);
                if (S->pending.empty()) break;
                if (S->inflight.size() >= S->cfg.max_inflight) break;
                r = S->pending.front();
                if (S->admitted_bytes + r->vol_bytes > S->cfg.hbm_budget_bytes) break;
                // The work cap holds while anything runs (a request larger
                // than the cap runs alone); with a backlog of at least
                // backlog_requests waiting, the wider busy cap.
                busy = S->cfg.backlog_requests != 0 &&
                       S->pending.size() >= S->cfg.backlog_requests;
                if (!S->inflight.empty()) {
                    const uint64_t cap =
                        busy && S->cfg.busy_work_budget_pxd ? S->cfg.busy_work_budget_pxd
                                                            : S->cfg.work_budget_pxd;
                    if (cap != 0 && S->inflight_pxd + request_pxd(r) > cap) break;
                }
                S->pending.pop_front();
            }
            bool hit = false;
            r->slot = acquire_slot(S, r, &hit);
            if (!r->slot) {
                if (S->inflight.empty()) {
                    // Nothing will free memory: fail this request.
                    r->t_admit = now_s();
                    S->admitted_bytes += r->vol_bytes;
                    S->inflight_pxd += request_pxd(r);
                    complete(S, r, VXH_ERR_BUFFER_ALLOC);
                    progressed = true;
                    continue;
                }
                // Out of device memory: put it back and retry after completions.
                std::lock_guard<std::mutex> g(S->mu);
                S->pending.push_front(r);
                break;
            }
            r->t_admit = now_s();
            S->admitted_bytes += r->vol_bytes;
            S->inflight_pxd += request_pxd(r);
            {
                std::lock_guard<std::mutex> g(S->mu);
                if (hit) S->st.slot_hits++;
                else S->st.slot_misses++;
                S->st.peak_admitted_bytes = std::max(S->st.peak_admitted_bytes, S->admitted_bytes);
            }
            if (!launch(S, r, busy)) {
                // Part of the request may be on its stream already: the slot
                // goes back only after that work has drained.
                if (r->stream) (void)hipStreamSynchronize(r->stream);
                complete(S, r, VXH_ERR_GPU);
            } else {
                S->inflight.push_back(r);
                std::lock_guard<std::mutex> g(S->mu);
                S->st.peak_inflight = std::max<uint64_t>(S->st.peak_inflight, S->inflight.size());
                S->st.peak_inflight_pxd =
                    std::max(S->st.peak_inflight_pxd, (double)S->inflight_pxd);
            }
            progressed = true;
        }
        std::unique_lock<std::mutex> g(S->mu);
        if (S->stop && S->pending.empty() && S->inflight.empty()) break;
        if (progressed) continue;
        if (S->inflight.empty())
            S->cv_work.wait(g, [&] { return S->stop || !S->pending.empty(); });
        else
            S->cv_work.wait_for(g, std::chrono::microseconds(20));
    }
}

}  // namespace

extern "C" {

uint64_t vxh_server_request_bytes(uint32_t w, uint32_t h, const vxh_sgm_config *cfg) {
    if (!cfg) return 0;
    return (uint64_t)vxh_sgm_workspace_bytes(w, h, cfg);
}

vxh_status vxh_server_create(int device, const vxh_server_config *cfg, vxh_server **out) {
    if (!cfg || !out) return sfail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    *out = nullptr;
    if (cfg->max_inflight == 0 || cfg->num_streams == 0 || cfg->hbm_budget_bytes == 0)
        return sfail(VXH_ERR_INVALID_CONFIG, "invalid config: server limits must be positive");
    vxh_server *S = new vxh_server();
    S->device = device;
    S->cfg = *cfg;
    vxh_status st = vxh_context_create(device, &S->ctx);
    if (st == VXH_OK) st = vxh_sgm_create(S->ctx, &S->sgm);
    if (st != VXH_OK) {
        g_srv_error = vxh_last_error();
        if (S->sgm) vxh_sgm_destroy(S->sgm);
        if (S->ctx) vxh_context_destroy(S->ctx);
        delete S;
        return st;
    }
    for (uint32_t i = 0; i < cfg->num_streams; ++i) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            vxh_server_destroy(S);
            return sfail(VXH_ERR_GPU, "GPU error: hipStreamCreate");
        }
        S->streams.push_back(s);
    }
    S->th = std::thread(scheduler, S);
    for (int i = 0; i < kFinishers; ++i) S->fins.emplace_back(finisher, S);
    *out = S;
    return VXH_OK;
}

vxh_status vxh_server_reserve(vxh_server *S, uint32_t w, uint32_t h, const vxh_sgm_config *cfg,
                              uint32_t count) {
    if (!S) return sfail(VXH_ERR_INVALID_CONFIG, "invalid config: null server");
    vxh_status v = vxh_sgm_config_validate(cfg, w, h);
    if (v != VXH_OK) return sfail(v, vxh_last_error());
    const size_t npx = (size_t)w * h;
    const size_t ws = vxh_sgm_workspace_bytes(w, h, cfg);
    if (hipSetDevice(S->device) != hipSuccess)
        return sfail(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    std::vector<Slot *> made;
    std::vector<Staging> pinned;
    for (uint32_t i = 0; i < count; ++i) {
        Slot *s = new_slot(npx, ws);
        Staging p, q;
        const bool got_p = s && take_staging(S, 2 * npx, &p);
        const bool got_q = got_p && take_staging(S, out_err_offset(npx) + 64, &q);
        if (got_p) pinned.push_back(p);
        if (got_q) pinned.push_back(q);
        if (!got_q) {
            if (s) free_slot(s);
            for (Slot *m : made) free_slot(m);
            for (Staging &x : pinned) give_staging(S, x);
            return sfail(VXH_ERR_BUFFER_ALLOC, "buffer allocation failed (server reserve)");
        }
        made.push_back(s);
    }
    for (Staging &q : pinned) give_staging(S, q);
    std::lock_guard<std::mutex> g(S->pool_mu);
    for (Slot *s : made) {
        S->pool_bytes += s->dev_bytes;
        S->idle.push_back(s);
    }
    return VXH_OK;
}

vxh_status vxh_server_submit(vxh_server *S, const uint8_t *left, const uint8_t *right,
                             uint32_t w, uint32_t h, const vxh_sgm_config *cfg, float *out,
                             uint64_t *request_id) {
    const double t = now_s();
    if (!S || !left || !right || !out)
        return sfail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    vxh_status v = vxh_sgm_config_validate(cfg, w, h);
    if (v != VXH_OK) return sfail(v, vxh_last_error());
    const uint64_t vol_bytes = vxh_server_request_bytes(w, h, cfg);
    if (vol_bytes > S->cfg.hbm_budget_bytes) {
        std::lock_guard<std::mutex> g(S->mu);
        S->st.submitted++;
        S->st.rejected++;
        return sfail(VXH_ERR_BUFFER_ALLOC, "buffer allocation failed (request volume exceeds "
                                           "the server's HBM budget)");
    }
    const size_t npx = (size_t)w * h;
    Request *r = new Request();
    if (!take_staging(S, 2 * npx, &r->in) || !take_staging(S, out_err_offset(npx) + 64, &r->res)) {
        give_staging(S, r->in);
        delete r;
        return sfail(VXH_ERR_BUFFER_ALLOC, "buffer allocation failed (pinned staging)");
    }
    r->w = w;
    r->h = h;
    r->cfg = *cfg;
    r->out = out;
    r->vol_bytes = vol_bytes;
    r->ws_bytes = vxh_sgm_workspace_bytes(w, h, cfg);
    r->t_submit = t;
    std::memcpy(r->in.ptr, left, npx);
    std::memcpy(r->in.ptr + npx, right, npx);
    {
        std::lock_guard<std::mutex> g(S->mu);
        r->id = S->next_id++;
        if (S->st.submitted == 0 || S->st.first_submit_s == 0) S->st.first_submit_s = t;
        S->st.submitted++;
        S->by_id[r->id] = r;
        S->pending.push_back(r);
    }
    S->cv_work.notify_one();
    if (request_id) *request_id = r->id;
    return VXH_OK;
}

vxh_status vxh_server_wait(vxh_server *S, uint64_t id, double *latency_ms) {
    if (!S) return sfail(VXH_ERR_INVALID_CONFIG, "invalid config: null server");
    std::unique_lock<std::mutex> g(S->mu);
    auto it = S->by_id.find(id);
    if (it == S->by_id.end()) return sfail(VXH_ERR_INVALID_CONFIG, "invalid config: unknown request id");
    Request *r = it->second;
    S->cv_done.wait(g, [&] { return r->done; });
    if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
    vxh_status st = r->status;
    S->by_id.erase(it);
    delete r;
    if (st == VXH_ERR_BUFFER_ALLOC) return sfail(st, "buffer allocation failed (request slot)");
    if (st != VXH_OK)
        return sfail(st, "GPU error: request failed on the device (launch failure or an "
                         "expired split-scanline wait)");
    return VXH_OK;
}

vxh_status vxh_server_drain(vxh_server *S) {
    if (!S) return sfail(VXH_ERR_INVALID_CONFIG, "invalid config: null server");
    std::unique_lock<std::mutex> g(S->mu);
    S->cv_done.wait(g, [&] { return S->st.completed + S->st.rejected >= S->st.submitted; });
    return VXH_OK;
}

static double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (double)(v.size() - 1);
    const size_t lo = (size_t)pos;
    const size_t hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

vxh_status vxh_server_get_stats(vxh_server *S, vxh_server_stats *out) {
    if (!S || !out) return sfail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    std::lock_guard<std::mutex> g(S->mu);
    *out = S->st;
    out->p50_ms = percentile(S->lat_ms, 0.50);
    out->p99_ms = percentile(S->lat_ms, 0.99);
    out->max_ms = S->lat_ms.empty() ? 0.0 : *std::max_element(S->lat_ms.begin(), S->lat_ms.end());
    double q = 0;
    for (double v : S->queue_ms) q += v;
    out->mean_queue_ms = S->queue_ms.empty() ? 0.0 : q / (double)S->queue_ms.size();
    return VXH_OK;
}

void vxh_server_reset_stats(vxh_server *S) {
    if (!S) return;
    std::lock_guard<std::mutex> g(S->mu);
    const uint64_t outstanding = S->st.submitted - S->st.completed - S->st.rejected;
    S->st = vxh_server_stats{};
    S->st.submitted = outstanding;
    S->lat_ms.clear();
    S->queue_ms.clear();
}

void vxh_server_destroy(vxh_server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    if (S->th.joinable()) S->th.join();
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop_finisher = true;
    }
    S->cv_finish.notify_all();
    for (std::thread &t : S->fins)
        if (t.joinable()) t.join();
    (void)hipSetDevice(S->device);
    for (hipStream_t s : S->streams) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (Slot *s : S->idle) free_slot(s);
    for (Staging &p : S->pin_idle) (void)hipHostFree(p.ptr);
    for (hipEvent_t e : S->events) (void)hipEventDestroy(e);
    for (auto &kv : S->by_id) {
        if (kv.second->in.ptr) (void)hipHostFree(kv.second->in.ptr);
        if (kv.second->res.ptr) (void)hipHostFree(kv.second->res.ptr);
        delete kv.second;
    }
    if (S->sgm) vxh_sgm_destroy(S->sgm);
    if (S->ctx) vxh_context_destroy(S->ctx);
    delete S;
}

const char *vxh_server_last_error(void) { return g_srv_error.c_str(); }

}  // extern "C"
