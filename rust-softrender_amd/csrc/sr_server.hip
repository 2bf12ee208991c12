// sr_server.hip — the request server (sr_server_* in include/softrender_hip.h):
// continuous batching of independent frame requests over the batched
// triangle pipeline.
//
// One dispatcher thread owns every device call of the server.  Each turn it
//   1. retires finished batches (event query on each slot): a batch whose
//      bins overflowed is queued again on the same slot with the grown
//      capacity; otherwise its requests complete (their pixels are already
//      in pinned host buffers), their framebuffers return to the pool and
//      their bytes leave the in-flight budget;
//   2. when a slot is free, admits pending requests in arrival order while
//      the batch limits and the HBM budget (framebuffer + workspace bytes
//      per request) allow, and queues them as one batch on that slot's
//      stream: table upload, the whole pipeline, then the copies of every
//      frame to its pinned host buffer and a completion event;
//   3. otherwise sleeps briefly (or until a submit when nothing runs).
// Requests that arrive while batches run therefore join the next batch as
// soon as a slot frees up, and up to max_inflight batches overlap on the
// device.  Every request gets its own framebuffer, so a result does not
// depend on what else shared its batch.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "sr_runtime.hpp"

using namespace sr;
using namespace sr_rt;

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct HostBuf {
    uint32_t* rgba = nullptr;
    float* depth = nullptr;
    size_t px = 0;
    bool has_depth = false;
};

struct Req {
    uint64_t ticket = 0;
    sr_request r{};
    double t_arrival = 0, t_admit = 0, t_done = 0;
    int32_t status = SR_OK;
    size_t cost = 0;  // bytes charged to the budget while in flight
    uint64_t tris = 0;
    sr_framebuffer fb = nullptr;
    HostBuf host;
    uint64_t batch_id = 0;
    uint32_t batch_frames = 0;
};

struct Slot {
    BatchWork W;
    BatchPlan plan;
    std::vector<Req*> reqs;
    hipEvent_t done = nullptr;
    bool busy = false;
    uint64_t id = 0;
    int reruns = 0;
};

}  // namespace

struct sr_server_t {
    sr_context user = nullptr;  // meshes and textures live here
    sr_context ictx = nullptr;  // the dispatcher's own context (error text, policy)
    sr_server_config cfg{};
    std::thread th;

    std::mutex mu;  // guards everything below up to `stop`
    std::condition_variable cv_submit, cv_done;
    std::deque<Req*> pending, done;
    std::unordered_map<uint64_t, Req*> held;  // completed, handed out, not yet released
    std::map<size_t, std::vector<HostBuf>> host_pool;
    std::string err;
    uint64_t next_ticket = 1;
    sr_server_stats st{};
    bool stop = false;

    // dispatcher-only state
    std::vector<Slot> slots;
    std::map<std::pair<uint32_t, uint32_t>, std::vector<sr_framebuffer>> fb_pool;
    std::vector<sr_framebuffer> all_fbs;
    uint64_t in_flight = 0, next_batch = 1;
};

namespace {

size_t fb_bytes(uint32_t w, uint32_t h) { return (size_t)w * h * 9; }

uint64_t mesh_tris(sr_mesh m) { return mesh_count(m) / 3; }

size_t request_cost(const sr_server_t* S, const sr_request& r) {
    const uint64_t t = mesh_tris(r.mesh);
    return fb_bytes(r.width, r.height) + batch_bytes_estimate(t, r.mesh->nv, S->cfg.pairs_per_tri);
}

void set_err(sr_server_t* S, const std::string& m) {
    std::lock_guard<std::mutex> g(S->mu);
    S->err = m;
}

// pinned result buffers, pooled by pixel count (caller holds S->mu)
bool host_get(sr_server_t* S, size_t px, bool depth, HostBuf& out) {
    auto& v = S->host_pool[px];
    for (size_t i = 0; i < v.size(); i++) {
        if (v[i].has_depth >= depth) {
            out = v[i];
            v.erase(v.begin() + (long)i);
            return true;
        }
    }
    HostBuf b;
    b.px = px;
    if (hipHostMalloc((void**)&b.rgba, px * 4, hipHostMallocDefault) != hipSuccess) return false;
    if (depth && hipHostMalloc((void**)&b.depth, px * 4, hipHostMallocDefault) != hipSuccess) {
        hipHostFree(b.rgba);
        return false;
    }
    b.has_depth = depth;
    out = b;
    return true;
}

void host_put(sr_server_t* S, const HostBuf& b) {
    if (b.rgba) S->host_pool[b.px].push_back(b);
}

int fb_get(sr_server_t* S, uint32_t w, uint32_t h, sr_framebuffer* out) {
    auto& v = S->fb_pool[{w, h}];
    if (!v.empty()) {
        *out = v.back();
        v.pop_back();
        return SR_OK;
    }
    int rc = sr_fb_create(S->ictx, w, h, out);
    if (rc == SR_OK) S->all_fbs.push_back(*out);
    return rc;
}

void fb_put(sr_server_t* S, sr_framebuffer fb) { S->fb_pool[{fb->w, fb->h}].push_back(fb); }

// Queues the slot's plan and the copies of every frame to the host.
int launch_slot(sr_server_t* S, Slot& sl) {
    int rc;
    sr_context ctx = S->ictx;
    // size the bins from the configured estimate so overflow reruns are rare
    const uint64_t want = (uint64_t)((double)sl.plan.NT * S->cfg.pairs_per_tri * 1.25) + 1024;
    if (want > sl.W.cap) sl.W.cap = (uint32_t)std::min<uint64_t>(want, 0x7FFFFFF0ull);
    if ((rc = batch_enqueue(ctx, sl.W, sl.plan, false))) return rc;
    for (Req* q : sl.reqs) {
        const size_t px = (size_t)q->r.width * q->r.height;
        HIP_OK(ctx, hipMemcpyAsync(q->host.rgba, q->fb->color, px * 4, hipMemcpyDeviceToHost, sl.W.stream));
        if (q->r.want_depth)
            HIP_OK(ctx, hipMemcpyAsync(q->host.depth, q->fb->depth, px * 4, hipMemcpyDeviceToHost, sl.W.stream));
    }
    HIP_OK(ctx, hipEventRecord(sl.done, sl.W.stream));
    sl.busy = true;
    return SR_OK;
}

void finish_reqs(sr_server_t* S, std::vector<Req*>& reqs, int status) {
    const double t = now_s();
    for (Req* q : reqs) {
        q->status = status;
        q->t_done = t;
        if (q->fb) fb_put(S, q->fb);
        q->fb = nullptr;
        S->in_flight -= q->cost;
    }
    std::lock_guard<std::mutex> g(S->mu);
    for (Req* q : reqs) {
        S->done.push_back(q);
        if (status == SR_OK) S->st.completed++;
        else S->st.failed++;
    }
    S->st.in_flight_bytes = S->in_flight;
    reqs.clear();
    S->cv_done.notify_all();
}

// Retires finished batches; returns true if any slot changed state.
bool retire(sr_server_t* S) {
    bool any = false;
    for (Slot& sl : S->slots) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        any = true;
        sl.busy = false;
        if (q != hipSuccess) {
            set_err(S, std::string("server: batch failed: ") + hipGetErrorString(q));
            finish_reqs(S, sl.reqs, SR_ERR_DEVICE);
            continue;
        }
        bool again = false;
        int rc = batch_result(S->ictx, sl.W, sl.plan, &again);
        if (rc == SR_OK && again && sl.reruns < 4) {
            sl.reruns++;
            {
                std::lock_guard<std::mutex> g(S->mu);
                S->st.overflow_reruns++;
            }
            if ((rc = launch_slot(S, sl)) == SR_OK) continue;
        } else if (rc == SR_OK && again) {
            rc = SR_ERR_DEVICE;
        }
        if (rc != SR_OK) set_err(S, S->ictx->err);
        finish_reqs(S, sl.reqs, rc);
    }
    return any;
}

// Admits pending requests into a free slot; returns true if a batch started.
bool admit(sr_server_t* S) {
    Slot* free_slot = nullptr;
    for (Slot& sl : S->slots)
        if (!sl.busy) {
            free_slot = &sl;
            break;
        }
    if (!free_slot) return false;
    std::vector<Req*> take;
    {
        std::lock_guard<std::mutex> g(S->mu);
        uint64_t tris = 0, bytes = S->in_flight;
        while (!S->pending.empty() && take.size() < S->cfg.max_batch_frames) {
            Req* q = S->pending.front();
            if (!take.empty() && tris + q->tris > S->cfg.max_batch_tris) break;
            if (bytes + q->cost > S->cfg.hbm_budget) break;  // arrival order: wait for room
            S->pending.pop_front();
            take.push_back(q);
            tris += q->tris;
            bytes += q->cost;
        }
        // pinned result buffers
        for (size_t i = 0; i < take.size(); i++) {
            Req* q = take[i];
            if (!host_get(S, (size_t)q->r.width * q->r.height, q->r.want_depth != 0, q->host)) {
                // put back what cannot be served now
                for (size_t j = take.size(); j-- > i;) S->pending.push_front(take[j]);
                take.resize(i);
                break;
            }
        }
    }
    if (take.empty()) return false;
    const double t = now_s();
    const uint64_t bid = S->next_batch++;
    Slot& sl = *free_slot;
    sl.reqs = take;
    sl.reruns = 0;
    sl.id = bid;
    std::vector<sr_batch_item> items;
    items.reserve(take.size());
    int rc = SR_OK;
    for (Req* q : take) {
        S->in_flight += q->cost;
        q->t_admit = t;
        q->batch_id = bid;
        q->batch_frames = (uint32_t)take.size();
        if (rc == SR_OK) rc = fb_get(S, q->r.width, q->r.height, &q->fb);
        sr_batch_item it;
        it.fb = q->fb;
        it.mesh = q->r.mesh;
        it.desc = &q->r.desc;
        it.tex = q->r.tex;
        it.clear = 1;
        it.clear_rgba = q->r.clear_rgba;
        items.push_back(it);
    }
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->st.batches++;
        S->st.in_flight_bytes = S->in_flight;
        S->st.peak_in_flight_bytes = std::max<uint64_t>(S->st.peak_in_flight_bytes, S->in_flight);
        S->st.max_batch_frames_seen = std::max<uint32_t>(S->st.max_batch_frames_seen, (uint32_t)take.size());
    }
    if (rc == SR_OK) rc = batch_plan(S->ictx, items.data(), (uint32_t)items.size(), sl.plan);
    if (rc == SR_OK) rc = launch_slot(S, sl);
    if (rc != SR_OK) {
        set_err(S, S->ictx->err);
        if (sl.W.stream) hipStreamSynchronize(sl.W.stream);
        sl.busy = false;
        finish_reqs(S, sl.reqs, rc);
    }
    return true;
}

void cancel_pending(sr_server_t* S) {
    std::lock_guard<std::mutex> g(S->mu);
    const double t = now_s();
    while (!S->pending.empty()) {
        Req* q = S->pending.front();
        S->pending.pop_front();
        q->status = SR_ERR_CANCELLED;
        q->t_done = t;
        S->done.push_back(q);
        S->st.failed++;
    }
    S->cv_done.notify_all();
}

// Copies the dispatcher-owned figures into the shared stats.
void publish(sr_server_t* S) {
    uint32_t busy = 0;
    size_t dev = 0;
    for (const Slot& sl : S->slots) {
        busy += sl.busy ? 1u : 0u;
        dev += sl.W.device_bytes();
    }
    for (sr_framebuffer fb : S->all_fbs) dev += fb_bytes(fb->w, fb->h);
    std::lock_guard<std::mutex> g(S->mu);
    S->st.inflight_batches = busy;
    S->st.device_bytes = dev;
    S->st.in_flight_bytes = S->in_flight;
}

void dispatcher(sr_server_t* S) {
    hipSetDevice(S->ictx->device);
    for (;;) {
        bool stopping;
        {
            std::lock_guard<std::mutex> g(S->mu);
            stopping = S->stop;
        }
        if (stopping) cancel_pending(S);
        bool progressed = retire(S);
        if (!stopping) progressed |= admit(S);
        bool busy = false;
        for (Slot& sl : S->slots) busy |= sl.busy;
        if (progressed) publish(S);
        if (stopping && !busy) break;
        if (progressed) continue;
        if (busy) {
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        } else {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cv_submit.wait_for(lk, std::chrono::milliseconds(2),
                                  [&] { return S->stop || !S->pending.empty(); });
        }
    }
}

void free_host(HostBuf& b) {
    if (b.rgba) hipHostFree(b.rgba);
    if (b.depth) hipHostFree(b.depth);
    b = HostBuf();
}

}  // namespace

extern "C" {

double sr_server_clock(void) { return now_s(); }

int sr_server_create(sr_context ctx, const sr_server_config* cfg, sr_server* out) {
    if (!ctx || !out) return SR_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    sr_server_t* S = new (std::nothrow) sr_server_t();
    if (!S) return SR_ERR_OUT_OF_MEMORY;
    if (cfg) S->cfg = *cfg;
    if (!S->cfg.hbm_budget) S->cfg.hbm_budget = 16ull << 30;
    // 4 slots of at most 32 frames: a batch's frames are copied back while the
    // next batches render (2 slots of 256 left the copies of a large batch
    // unoverlapped: 3.3k against 6.7k frames/s with 384 requests queued)
    if (!S->cfg.max_inflight) S->cfg.max_inflight = 4;
    if (!S->cfg.max_batch_frames) S->cfg.max_batch_frames = 32;
    if (!S->cfg.max_batch_tris) S->cfg.max_batch_tris = 32ull << 20;
    if (!(S->cfg.pairs_per_tri > 0.f)) S->cfg.pairs_per_tri = 4.f;
    if (S->cfg.max_inflight > 8) return delete S, fail(ctx, SR_ERR_INVALID_ARGUMENT, "server: max_inflight > 8");
    S->user = ctx;
    int rc = sr_ctx_create(ctx->device, &S->ictx);
    if (rc) {
        delete S;
        return rc;
    }
    S->ictx->tile_policy = ctx->tile_policy;
    S->ictx->batch_cap_hint = ctx->batch_cap_hint;
    S->slots.resize(S->cfg.max_inflight);
    for (Slot& sl : S->slots) {
        if ((rc = sl.W.init(S->ictx, true)) ||
            hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
            for (Slot& x : S->slots) {
                x.W.release();
                if (x.done) hipEventDestroy(x.done);
            }
            sr_ctx_destroy(S->ictx);
            delete S;
            return rc ? rc : SR_ERR_DEVICE;
        }
    }
    S->st.budget_bytes = S->cfg.hbm_budget;
    S->th = std::thread(dispatcher, S);
    *out = S;
    return SR_OK;
}

int sr_server_destroy(sr_server S) {
    if (!S) return SR_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv_submit.notify_all();
    if (S->th.joinable()) S->th.join();
    hipSetDevice(S->ictx->device);
    for (Slot& sl : S->slots) {
        sl.W.release();
        if (sl.done) hipEventDestroy(sl.done);
    }
    for (sr_framebuffer fb : S->all_fbs) sr_fb_free(fb);
    for (auto& kv : S->host_pool)
        for (HostBuf& b : kv.second) free_host(b);
    for (auto& kv : S->held) {
        free_host(kv.second->host);
        delete kv.second;
    }
    for (Req* q : S->done) {
        free_host(q->host);
        delete q;
    }
    sr_ctx_destroy(S->ictx);
    delete S;
    return SR_OK;
}

int sr_server_submit(sr_server S, const sr_request* r, uint64_t* ticket) {
    if (!S || !r) return SR_ERR_INVALID_ARGUMENT;
    const char* why = nullptr;
    if (!r->mesh || r->mesh->ctx != S->user) why = "server: mesh missing or from another context";
    else if (r->tex && r->tex->ctx != S->user) why = "server: texture from another context";
    else if (r->width < 1 || r->height < 1 || r->width > 16384 || r->height > 16384) why = "server: bad framebuffer size";
    else if (!batchable(r->mesh, &r->desc, r->tex))
        why = "server: requests are triangle draws without geometry program, stencil Always/Keep, RGBA8 texture";
    else if (r->desc.vertex_program != SR_VP_MVP_NORMAL_UV && r->desc.vertex_program != SR_VP_MVP_NORMAL_WORLD)
        why = "server: unknown vertex program";
    else if (r->desc.fragment_program < SR_FP_FLAT || r->desc.fragment_program > SR_FP_UV_COLOR)
        why = "server: unknown fragment program";
    else if (r->desc.fragment_program == SR_FP_TEX_LAMBERT && !r->tex)
        why = "server: fragment program samples a texture";
    else if (r->desc.fragment_program == SR_FP_BLINN_PHONG && r->desc.vertex_program != SR_VP_MVP_NORMAL_WORLD)
        why = "server: Blinn-Phong needs world-space varyings";
    else if ((r->desc.fragment_program == SR_FP_TEX_LAMBERT || r->desc.fragment_program == SR_FP_UV_COLOR) &&
             r->desc.vertex_program != SR_VP_MVP_NORMAL_UV)
        why = "server: program needs uv varyings";
    else if (r->desc.exec_tile != 0 && r->desc.exec_tile != 8 && r->desc.exec_tile != 16 &&
             r->desc.exec_tile != 32 && r->desc.exec_tile != 64)
        why = "server: exec_tile must be 0, 8, 16, 32 or 64";
    if (why) {
        set_err(S, why);
        return SR_ERR_INVALID_ARGUMENT;
    }
    Req* q = new (std::nothrow) Req();
    if (!q) return SR_ERR_OUT_OF_MEMORY;
    q->r = *r;
    q->r.desc.exec_flags &= ~SR_EXEC_TIMING;
    q->tris = mesh_tris(r->mesh);
    q->cost = request_cost(S, q->r);
    q->t_arrival = r->arrival > 0 ? r->arrival : now_s();
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (q->cost > S->cfg.hbm_budget) {
            S->err = "server: request exceeds the HBM budget on its own";
            delete q;
            return SR_ERR_BUDGET;
        }
        q->ticket = S->next_ticket++;
        S->pending.push_back(q);
        S->st.submitted++;
        if (ticket) *ticket = q->ticket;
    }
    S->cv_submit.notify_one();
    return SR_OK;
}

int sr_server_poll(sr_server S, sr_completion* out, uint32_t max, uint32_t* n, int32_t timeout_ms) {
    if (!S || !n || (max && !out)) return SR_ERR_INVALID_ARGUMENT;
    *n = 0;
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done.empty() && timeout_ms != 0) {
        auto ready = [&] { return !S->done.empty(); };
        if (timeout_ms < 0) S->cv_done.wait(lk, ready);
        else S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    }
    while (*n < max && !S->done.empty()) {
        Req* q = S->done.front();
        S->done.pop_front();
        S->held[q->ticket] = q;
        sr_completion& c = out[(*n)++];
        memset(&c, 0, sizeof(c));
        c.ticket = q->ticket;
        c.tag = q->r.tag;
        c.status = q->status;
        c.width = q->r.width;
        c.height = q->r.height;
        c.batch_frames = q->batch_frames;
        c.batch_id = q->batch_id;
        c.rgba = q->status == SR_OK ? q->host.rgba : nullptr;
        c.depth = (q->status == SR_OK && q->r.want_depth) ? q->host.depth : nullptr;
        c.t_arrival = q->t_arrival;
        c.t_admit = q->t_admit;
        c.t_done = q->t_done;
    }
    return SR_OK;
}

int sr_server_release(sr_server S, uint64_t ticket) {
    if (!S) return SR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->held.find(ticket);
    if (it == S->held.end()) {
        S->err = "server: ticket not held (never polled, or released already)";
        return SR_ERR_INVALID_ARGUMENT;
    }
    host_put(S, it->second->host);
    delete it->second;
    S->held.erase(it);
    return SR_OK;
}

int sr_server_get_stats(sr_server S, sr_server_stats* out) {
    if (!S || !out) return SR_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(S->mu);
    *out = S->st;
    out->pending = (uint32_t)S->pending.size();
    return SR_OK;
}

const char* sr_server_last_error(sr_server S) {
    if (!S) return "null server";
    std::lock_guard<std::mutex> g(S->mu);
    thread_local std::string copy;
    copy = S->err;
    return copy.c_str();
}

}  // extern "C"
