// SGM serving loop behind include/vxh/sgm_server.h.
//
// One scheduler thread owns the device side: it polls in-flight requests for
// completion, then admits queued requests in arrival order while the sum of
// admitted cost-volume bytes (each request's pipeline workspace: the
// per-direction volumes) stays within the HBM budget and fewer than
// max_inflight are running.  An admitted request takes a device-resident slot
// (inputs, output, census + volume scratch) from a pool keyed by its size,
// like the reference's TexturePool (vx-vision/src/pool.rs:61-97), and runs
// H2D copy -> the per-pair SGM pipeline -> D2H copy on one of several HIP
// streams, so requests that arrive while others are on the GPU overlap with
// them.  Work never changes with the mix: each request runs the same kernels
// on its own buffers, so its result is the one it would get alone.
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


struct Slot {
    uint32_t w = 0, h = 0;
    size_t ws_bytes = 0;
    uint8_t *d_in = nullptr;    // left then right
    float *d_out = nullptr;
    void *d_ws = nullptr;
    uint8_t *h_in = nullptr;    // pinned staging
    float *h_out = nullptr;
    size_t dev_bytes = 0;
    double last_used = 0;
};

struct Request {
    uint64_t id = 0;
    uint32_t w = 0, h = 0;
    vxh_sgm_config cfg{};
    const uint8_t *left = nullptr, *right = nullptr;
    std::vector<uint8_t> in_copy;   // left||right, copied at submit
    float *out = nullptr;
    uint64_t vol_bytes = 0;
    size_t ws_bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    Slot *slot = nullptr;
    hipEvent_t ev = nullptr;
    bool done = false;
    vxh_status status = VXH_OK;
};

thread_local std::string g_srv_error;
vxh_status sfail(vxh_status st, const std::string &m) {
    g_srv_error = m;
    return st;
}

}  // namespace

struct vxh_server {
    int device = 0;
    vxh_server_config cfg{};
    vxh_context *ctx = nullptr;
    vxh_sgm *sgm = nullptr;
    std::vector<hipStream_t> streams;
    size_t rr = 0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> pending;             // guarded by mu
    std::unordered_map<uint64_t, Request *> by_id;  // guarded by mu
    uint64_t next_id = 1;
    bool stop = false;

    // Scheduler-thread state.
    std::vector<Request *> inflight;
    std::vector<Slot *> idle;
    uint64_t pool_bytes = 0;
    uint64_t admitted_bytes = 0;
    std::vector<hipEvent_t> events;
    std::thread th;

    // Statistics, guarded by mu.
    vxh_server_stats st{};
    std::vector<double> lat_ms, queue_ms;
};

namespace {

void free_slot(Slot *s) {
    if (s->d_in) (void)hipFree(s->d_in);
    if (s->d_out) (void)hipFree(s->d_out);
    if (s->d_ws) (void)hipFree(s->d_ws);
    if (s->h_in) (void)hipHostFree(s->h_in);
    if (s->h_out) (void)hipHostFree(s->h_out);
    delete s;
}

// Takes an idle slot of the request's size or allocates one, first releasing
// idle slots of other sizes when the pool would exceed its cap.
Slot *acquire_slot(vxh_server *S, Request *r, bool *hit) {
    for (size_t i = 0; i < S->idle.size(); ++i) {
        Slot *s = S->idle[i];
        if (s->w == r->w && s->h == r->h && s->ws_bytes >= r->ws_bytes) {
            S->idle.erase(S->idle.begin() + (long)i);
            *hit = true;
            return s;
        }
    }
    *hit = false;
    const size_t npx = (size_t)r->w * r->h;
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
    Slot *s = new Slot();
    s->w = r->w;
    s->h = r->h;
    s->ws_bytes = r->ws_bytes;
    s->dev_bytes = need;
    bool ok = hipMalloc(&s->d_in, 2 * npx) == hipSuccess &&
              hipMalloc(&s->d_out, npx * 4) == hipSuccess &&
              hipMalloc(&s->d_ws, r->ws_bytes) == hipSuccess &&
              hipHostMalloc(&s->h_in, 2 * npx, hipHostMallocDefault) == hipSuccess &&
              hipHostMalloc(&s->h_out, npx * 4, hipHostMallocDefault) == hipSuccess;
    if (!ok) {
        free_slot(s);
        return nullptr;
    }
    S->pool_bytes += need;
    return s;
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

// Enqueues one admitted request on a stream.  Returns false when the device
// work could not be enqueued (the request then completes with an error).
bool launch(vxh_server *S, Request *r) {
    Slot *s = r->slot;
    const size_t npx = (size_t)r->w * r->h;
    std::memcpy(s->h_in, r->in_copy.data(), 2 * npx);
    hipStream_t st = S->streams[S->rr++ % S->streams.size()];
    r->ev = take_event(S);
    if (!r->ev) return false;
    if (hipMemcpyAsync(s->d_in, s->h_in, 2 * npx, hipMemcpyHostToDevice, st) != hipSuccess)
        return false;
    if (vxh_sgm_enqueue(S->sgm, s->d_in, s->d_in + npx, r->w, r->h, &r->cfg, s->d_out, s->d_ws,
                        s->ws_bytes, st) != VXH_OK)
        return false;
    if (hipMemcpyAsync(s->h_out, s->d_out, npx * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
        return false;
    return hipEventRecord(r->ev, st) == hipSuccess;
}

void finish(vxh_server *S, Request *r, vxh_status status) {
    const size_t npx = (size_t)r->w * r->h;
    if (status == VXH_OK && r->slot) std::memcpy(r->out, r->slot->h_out, npx * 4);
    r->t_done = now_s();
    if (r->ev) S->events.push_back(r->ev);
    r->ev = nullptr;
    if (r->slot) {
        r->slot->last_used = r->t_done;
        S->idle.push_back(r->slot);
        r->slot = nullptr;
    }
    S->admitted_bytes -= r->vol_bytes;
    std::vector<uint8_t>().swap(r->in_copy);
    std::lock_guard<std::mutex> g(S->mu);
    r->done = true;
    r->status = status;
    S->st.completed++;
    const double lat = (r->t_done - r->t_submit) * 1e3;
    S->lat_ms.push_back(lat);
    S->queue_ms.push_back((r->t_admit - r->t_submit) * 1e3);
    S->st.pxd_total += (double)r->w * r->h * r->cfg.num_disparities;
    S->st.last_complete_s = r->t_done;
    S->cv_done.notify_all();
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
            finish(S, r, q == hipSuccess ? VXH_OK : VXH_ERR_GPU);
            progressed = true;
        }
        // Admissions, in arrival order.
        while (true) {
            Request *r = nullptr;
            {
                std::unique_lock<std::mutex> g(S->mu);
                if (S->pending.empty()) break;
                if (S->inflight.size() >= S->cfg.max_inflight) break;
                r = S->pending.front();
                if (S->admitted_bytes + r->vol_bytes > S->cfg.hbm_budget_bytes) break;
                S->pending.pop_front();
            }
            bool hit = false;
            r->slot = acquire_slot(S, r, &hit);
            if (!r->slot) {
                if (S->inflight.empty()) {
                    // Nothing will free memory: fail this request.
                    r->t_admit = now_s();
                    S->admitted_bytes += r->vol_bytes;
                    finish(S, r, VXH_ERR_BUFFER_ALLOC);
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
            {
                std::lock_guard<std::mutex> g(S->mu);
                if (hit) S->st.slot_hits++;
                else S->st.slot_misses++;
                S->st.peak_admitted_bytes = std::max(S->st.peak_admitted_bytes, S->admitted_bytes);
            }
            if (!launch(S, r)) {
                finish(S, r, VXH_ERR_GPU);
            } else {
                S->inflight.push_back(r);
                std::lock_guard<std::mutex> g(S->mu);
                S->st.peak_inflight = std::max<uint64_t>(S->st.peak_inflight, S->inflight.size());
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
    *out = S;
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
    Request *r = new Request();
    r->w = w;
    r->h = h;
    r->cfg = *cfg;
    r->out = out;
    r->vol_bytes = vxh_server_request_bytes(w, h, cfg);
    r->ws_bytes = vxh_sgm_workspace_bytes(w, h, cfg);
    r->t_submit = t;
    if (r->vol_bytes > S->cfg.hbm_budget_bytes) {
        std::lock_guard<std::mutex> g(S->mu);
        S->st.submitted++;
        S->st.rejected++;
        delete r;
        return sfail(VXH_ERR_BUFFER_ALLOC, "buffer allocation failed (request volume exceeds "
                                           "the server's HBM budget)");
    }
    const size_t npx = (size_t)w * h;
    r->in_copy.resize(2 * npx);
    std::memcpy(r->in_copy.data(), left, npx);
    std::memcpy(r->in_copy.data() + npx, right, npx);
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
    if (st != VXH_OK) return sfail(st, "GPU error: request failed on the device");
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
    (void)hipSetDevice(S->device);
    for (hipStream_t s : S->streams) {
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
    }
    for (Slot *s : S->idle) free_slot(s);
    for (hipEvent_t e : S->events) (void)hipEventDestroy(e);
    for (auto &kv : S->by_id) delete kv.second;
    if (S->sgm) vxh_sgm_destroy(S->sgm);
    if (S->ctx) vxh_context_destroy(S->ctx);
    delete S;
}

const char *vxh_server_last_error(void) { return g_srv_error.c_str(); }

}  // extern "C"
