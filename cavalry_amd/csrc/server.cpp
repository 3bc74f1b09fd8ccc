// Render server: continuous batching of render requests over resident scenes
// (cav_serve_*). No counterpart in the reference, which renders one scene
// per process; this is the serving layer around its integrators.
//
// One worker thread owns a persistent ray pool per scene. Every step it
// admits camera rays of the scene's pending requests into the pool
// (round-robin, within the pool's free task slots and the device arena),
// then enqueues one bounce of every task in the pool. Tasks of different
// requests, integrators and depths share every launch. A batch of samples
// admitted at step s has finished after step s + max_depth - 1 (step s for
// directlighting), so completion is known on the host by counting steps; the
// request's film is then enqueued and copied to pinned staging, and an event
// says when it is ready. The host runs at most `max_steps_ahead` steps ahead
// of the device. Requests whose buffers do not fit the arena wait (admission
// by memory).
//
// path / directlighting requests are admitted in chunks of contiguous sample
// ids (at most chunk_samples per request per step). rlpath learns from every
// sample of a bounce of its own round (the reference's render loop,
// ReinforcementLearningPathTracing.cu:802-950), so an rlpath request is
// admitted one whole round at a time (round_spp samples of every pixel) once
// its previous round has finished, with its own Q table, light statistics
// and syncDimension words in its arena block: its film is the one cav_render
// makes with spp_per_round = round_spp, whatever else shares the pool.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/cavalry.h"
#include "capi_internal.h"
#include "render.h"
#include "serve.h"

namespace cav {
SceneDev scene_dev(const cav_scene* sc);
int halton_dev(cav_ctx* ctx, HaltonDev* out);
void scene_hold(cav_scene* sc, int delta);
void scene_bounds(const cav_scene* sc, float out[6]);
}  // namespace cav

using namespace cav;

namespace {

constexpr int kSlots = kPoolSlots;
constexpr int kQBins = 128;  // direction bins per Q entry (QEntry::NUM_XY)

double now_ms(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// "what: HIP error string", for failure messages
std::string hip_msg(const char* what, hipError_t e) { return std::string(what) + ": " + hipGetErrorString(e); }

// first-fit sub-allocator over one pre-allocated block (no allocation calls
// while serving)
class Arena {
public:
    void init(char* base, size_t size) {
        base_ = base;
        free_.clear();
        if (size) free_[0] = size;
    }
    bool alloc(size_t n, size_t& off) {
        n = (n + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = (n + 255) & ~(size_t)255;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* at(size_t off) const { return base_ + off; }

private:
    char* base_ = nullptr;
    std::map<size_t, size_t> free_;
};

struct Request {
    int64_t id = 0;
    cav_scene* scene = nullptr;
    cav_camera cam{};
    cav_render_params prm{};
    float* out_rgb = nullptr;
    float* out_w = nullptr;
    int64_t npix = 0, nsamp = 0;
    int32_t status = CAV_SERVE_QUEUED;
    int32_t error = CAV_OK;
    double t_submit = 0, t_start = -1, t_done = -1;
    bool active = false;
    int slot = -1;
    int64_t next_sample = 0;  // path / directlighting: next sample id to admit
    int spp_admitted = 0;     // rlpath: samples per pixel admitted so far
    int round_spp = 0;        // rlpath: samples per pixel per round
    int grid = 32;            // rlpath: Q entries per axis
    int64_t outstanding = 0;  // admitted batches not yet finished
    size_t dev_off = 0, dev_len = 0, host_off = 0, host_len = 0;
    float *res = nullptr, *sxy = nullptr, *rgb = nullptr, *w = nullptr;
    float4* partial = nullptr;
    float *h_rgb = nullptr, *h_w = nullptr;
    uint32_t* h_ovf = nullptr;  // pool overflow counters at admission [0..1] and at the end [2..3]
    hipEvent_t ready = nullptr;
    bool finishing = false;
    bool cancel = false;
    bool final = false;  // a final status was set (the scene's hold released)
    std::string msg;     // why it failed
    bool rl() const { return prm.integrator == CAV_INTEGRATOR_RLPATH; }
    bool det() const { return rl() && (prm.flags & CAV_RENDER_DETERMINISTIC); }
    bool all_admitted() const { return rl() ? spp_admitted == prm.spp : next_sample == nsamp; }
};

struct LiveChunk {
    Request* req;
    int64_t done_step;
};

struct Pool {
    cav_scene* scene = nullptr;
    PoolDev dev{};
    hipStream_t stream = nullptr;
    std::vector<void*> bufs;
    int qin = 0;
    int64_t step = 0;
    int64_t reserved = 0;
    std::vector<std::pair<LiveChunk, int64_t>> live;  // batch, samples
    std::deque<Request*> pending;
    Request* slots[kSlots] = {};
    ReqDev* d_reqs = nullptr;
    ReqDev* h_reqs = nullptr;        // pinned staging of request entries, [ring][kSlots]
    RlReqState* h_rl = nullptr;      // pinned staging of rlpath states, [ring][kSlots]
    Chunk* d_chunks = nullptr;       // [ring][kSlots]
    Chunk* h_chunks = nullptr;       // pinned
    std::vector<hipEvent_t> ring;
    int64_t done_upto = 0;           // steps known complete on the device
    // the bounce sequence per queue parity and bounce kind (kBounce*)
    hipGraphExec_t graph[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    bool graphs = false;
};

}  // namespace

struct cav_server {
    cav_ctx* ctx = nullptr;
    cav_serve_config cfg{};
    std::chrono::steady_clock::time_point t0;
    HaltonDev halton{};
    std::mutex mu;
    std::atomic<int> api_waiting{0};  // API calls waiting for mu
    std::condition_variable cv, done_cv;
    bool stop = false;
    int64_t next_id = 1;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request*> incoming;
    std::vector<std::unique_ptr<Pool>> pools;
    Arena dev_arena, host_arena;
    char* dev_base = nullptr;
    char* host_base = nullptr;
    std::thread worker;
    unsigned long long pool_rays[2] = {0, 0};
    std::string error;       // fatal: the failing call and its HIP error
    std::string graph_note;  // why a pool runs without graphs
};

namespace {
// The lock an API call takes. It announces itself first: the worker re-takes
// mu right after each step, and std::mutex is not fair, so without this a
// submitter could wait until the pools drain.
struct ApiLock {
    std::unique_lock<std::mutex> lk;
    explicit ApiLock(cav_server* sv) : lk((sv->api_waiting.fetch_add(1), sv->mu)) { sv->api_waiting.fetch_sub(1); }
};

// Sets a request's final status once, releasing its hold on the scene.
void finalize(cav_server* sv, Request* r, int32_t status, int32_t error = CAV_OK, const std::string& msg = "") {
    if (r->final) return;
    r->final = true;
    r->status = status;
    r->error = error;
    if (!msg.empty()) r->msg = msg;
    r->t_done = now_ms(sv->t0);
    scene_hold(r->scene, -1);
    sv->done_cv.notify_all();
}

// rlpath samples per pixel per round: the caller's, or as many as the
// admission quantum holds (at least one)
int rl_round_spp(const cav_serve_config& c, const cav_render_params& prm, int64_t npix) {
    if (prm.spp_per_round > 0) return std::min(prm.spp_per_round, prm.spp);
    const int64_t k = c.chunk_samples / npix;
    return (int)std::max<int64_t>(1, std::min<int64_t>(prm.spp, k));
}

int rl_grid(const cav_render_params& prm) { return prm.q_grid > 0 ? prm.q_grid : 32; }

// device bytes of a request's arena block
size_t request_bytes(const Request& r, int nlights) {
    const size_t S = (size_t)r.nsamp, P = (size_t)r.npix;
    size_t b = 12 * S + 8 * S + 144 * P + 12 * P + 4 * P + 5 * 256;
    if (r.rl()) {
        const size_t E = (size_t)r.grid * r.grid * r.grid;
        b += 4 * (4 * E * kQBins + 256) + 5 * (4 * E * nlights + 256) + 256;
    }
    return b;
}

int alloc_dev(std::vector<void*>& list, void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return CAV_ERR_OOM;
    list.push_back(*p);
    return CAV_OK;
}

int make_pool(cav_server* sv, cav_scene* sc, Pool*& out) {
    std::unique_ptr<Pool> p(new Pool());
    p->scene = sc;
    const size_t C = (size_t)sv->cfg.pool_tasks;
    PoolDev& d = p->dev;
    d.scene = scene_dev(sc);
    d.launch = sv->ctx->launch[CAV_TRACE_WAVEFRONT];
    d.halton = sv->halton;
    d.cap = (int64_t)C;
    auto& L = p->bufs;
    int rc = CAV_OK;
    auto A = [&](auto*& ptr, size_t bytes) {
        if (rc == CAV_OK) rc = alloc_dev(L, reinterpret_cast<void**>(&ptr), bytes);
    };
    for (int k = 0; k < 2; ++k) {
        A(d.q[k].ray, 24 * C); A(d.q[k].sid, 4 * C); A(d.q[k].mult, 12 * C); A(d.q[k].pdf, 4 * C);
        A(d.q[k].delta, C); A(d.rid[k], 4 * C); A(d.depth[k], C);
    }
    A(d.sh.ray, 24 * C); A(d.sh.limit, 4 * C); A(d.sh.src, 4 * C); A(d.sh.dst, 4 * C); A(d.sh.sid, 4 * C);
    A(d.sh.contrib, 12 * C); A(d.sh_rid, 4 * C); A(d.vis, C); A(d.hit_prim, 4 * C); A(d.hit_t, 4 * C);
    A(d.hit_u, 4 * C); A(d.hit_v, 4 * C);
    A(d.counts, 64); A(d.rays, 64); A(d.trace_counters, 128);
    unsigned char* coop = nullptr;
    const uint32_t cc = any_spill_capacity((int64_t)C);
    A(coop, trace_coop_bytes(cc, cc));
    if (rc == CAV_OK) d.coop = trace_coop_in(coop, cc, cc);
    // rlpath tasks: per-slot learning state, queue and bounce state, records
    PoolRl& R = d.rl;
    A(R.req, sizeof(RlReqState) * kSlots); A(R.live, kSlots);
    for (int k = 0; k < 2; ++k) { A(R.q[k].prev_entry, 4 * C); A(R.q[k].prev_cell, 4 * C); A(R.q[k].dim, 4 * C); }
    A(R.sl.entry, 4 * C); A(R.sl.light, 4 * C); A(R.sl.sel, 4 * C); A(R.sl.lpdf, 4 * C); A(R.sl.emis, 4 * C);
    A(R.sl.ldir, 12 * C); A(R.sl.inc, 12 * C); A(R.sl.sh_idx, 4 * C); A(R.sl.ndir, 12 * C); A(R.sl.npdf, 4 * C);
    A(R.sl.ncell, 4 * C);
    R.rec_cap = 3 * (int64_t)C;  // at most 3 updates per task
    const size_t RC = (size_t)R.rec_cap, RT = (size_t)sort_tiles(R.rec_cap);
    for (int k = 0; k < 2; ++k) { A(R.rec_keys[k], 8 * RC); A(R.rec_idx[k], 4 * RC); }
    A(R.rec_val, 4 * RC); A(R.rec_count, 256); A(R.rec_hist, 4 * 8 * 256); A(R.rec_counts, 4 * 256 * RT);
    A(R.rec_offsets, 4 * 256 * RT);
    A(p->d_reqs, sizeof(ReqDev) * kSlots);
    const int ring = sv->cfg.max_steps_ahead + 2;
    A(p->d_chunks, sizeof(Chunk) * kSlots * ring);
    if (rc != CAV_OK) {
        for (void* q : L) (void)hipFree(q);
        sv->error = "ray pool allocation (" + std::to_string(C) + " tasks)";
        return rc;
    }
    d.reqs = p->d_reqs;
    bool ok = hipHostMalloc(reinterpret_cast<void**>(&p->h_reqs), sizeof(ReqDev) * kSlots * ring, 0) == hipSuccess &&
              hipHostMalloc(reinterpret_cast<void**>(&p->h_rl), sizeof(RlReqState) * kSlots * ring, 0) == hipSuccess &&
              hipHostMalloc(reinterpret_cast<void**>(&p->h_chunks), sizeof(Chunk) * kSlots * ring, 0) == hipSuccess &&
              hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) == hipSuccess;
    p->ring.assign(ring, nullptr);
    for (auto& e : p->ring) ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipMemsetAsync(d.counts, 0, 64, p->stream) == hipSuccess &&
         hipMemsetAsync(d.rays, 0, 64, p->stream) == hipSuccess &&
         hipMemsetAsync(d.trace_counters, 0, 128, p->stream) == hipSuccess &&
         hipMemsetAsync(R.req, 0, sizeof(RlReqState) * kSlots, p->stream) == hipSuccess &&
         hipMemsetAsync(R.live, 0, kSlots, p->stream) == hipSuccess &&
         hipStreamSynchronize(p->stream) == hipSuccess;
    if (!ok) {
        sv->error = "creating the ray pool's stream, events or pinned staging failed";
        return CAV_ERR_HIP;
    }
    scene_hold(sc, 1);
    out = p.get();
    sv->pools.push_back(std::move(p));
    return CAV_OK;
}

void destroy_pool(Pool* p) {
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    for (auto& row : p->graph)
        for (auto& g : row)
            if (g) (void)hipGraphExecDestroy(g);
    for (auto e : p->ring)
        if (e) (void)hipEventDestroy(e);
    for (void* q : p->bufs) (void)hipFree(q);
    if (p->h_reqs) (void)hipHostFree(p->h_reqs);
    if (p->h_rl) (void)hipHostFree(p->h_rl);
    if (p->h_chunks) (void)hipHostFree(p->h_chunks);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    scene_hold(p->scene, -1);
}

Pool* pool_for(cav_server* sv, cav_scene* sc) {
    for (auto& p : sv->pools)
        if (p->scene == sc) return p.get();
    return nullptr;
}

// A request slot, buffers and (rlpath) fresh tables for a request's first
// admission; false if they do not fit now. Deterministic rlpath requests take
// slots below kPoolDetSlots (their index in the update records), the others
// prefer the slots above.
bool activate(cav_server* sv, Pool* p, Request* r) {
    int slot = -1;
    if (r->det()) {
        for (int k = 0; k < kPoolDetSlots && slot < 0; ++k)
            if (!p->slots[k]) slot = k;
    } else {
        for (int k = kPoolDetSlots; k < kSlots && slot < 0; ++k)
            if (!p->slots[k]) slot = k;
        for (int k = 0; k < kPoolDetSlots && slot < 0; ++k)
            if (!p->slots[k]) slot = k;
    }
    if (slot < 0) return false;
    const SceneDev& sd = p->dev.scene;
    const size_t S = (size_t)r->nsamp, P = (size_t)r->npix;
    const size_t dlen = request_bytes(*r, sd.nlights);
    const size_t hlen = 16 * P + 3 * 256;
    size_t doff = 0, hoff = 0;
    if (!sv->dev_arena.alloc(dlen, doff)) return false;
    if (!sv->host_arena.alloc(hlen, hoff)) {
        sv->dev_arena.release(doff, dlen);
        return false;
    }
    auto carve = [](char*& c, size_t bytes) {
        char* q = c;
        c += (bytes + 255) & ~(size_t)255;
        return q;
    };
    char* c = sv->dev_arena.at(doff);
    r->res = reinterpret_cast<float*>(carve(c, 12 * S));
    r->sxy = reinterpret_cast<float*>(carve(c, 8 * S));
    r->partial = reinterpret_cast<float4*>(carve(c, 144 * P));
    r->rgb = reinterpret_cast<float*>(carve(c, 12 * P));
    r->w = reinterpret_cast<float*>(carve(c, 4 * P));
    const int ring = (int)p->ring.size();
    const size_t stage = (size_t)(p->step % ring) * kSlots + slot;
    // staged per step: an entry copied at an earlier step may still be queued
    RlReqState& st = p->h_rl[stage];
    std::memset(&st, 0, sizeof st);
    if (r->rl()) {
        const size_t E = (size_t)r->grid * r->grid * r->grid, NL = (size_t)sd.nlights;
        RlTables& t = st.t;
        t.Q = reinterpret_cast<float*>(carve(c, 4 * E * kQBins));
        t.newQ = reinterpret_cast<float*>(carve(c, 4 * E * kQBins));
        t.cnt = reinterpret_cast<int32_t*>(carve(c, 4 * E * kQBins));
        t.total = reinterpret_cast<int32_t*>(carve(c, 4 * E * kQBins));
        t.ldist = reinterpret_cast<float*>(carve(c, 4 * E * NL));
        t.lres = reinterpret_cast<float*>(carve(c, 4 * E * NL));
        t.lnew = reinterpret_cast<float*>(carve(c, 4 * E * NL));
        t.lcount = reinterpret_cast<int32_t*>(carve(c, 4 * E * NL));
        t.lnewc = reinterpret_cast<int32_t*>(carve(c, 4 * E * NL));
        st.dims = reinterpret_cast<int32_t*>(carve(c, 64));
        float b6[6];
        scene_bounds(p->scene, b6);
        for (int k = 0; k < 3; ++k) {
            st.gr.lo[k] = b6[k];
            st.gr.size[k] = b6[3 + k] - b6[k];
        }
        st.gr.g = r->grid;
        st.res = r->res;
        st.alpha_scale = r->prm.alpha_scale > 0.f ? r->prm.alpha_scale : 1.f;
        st.rlx = (uint32_t)slot;
        st.max_depth = r->prm.max_depth;
        st.det = r->det() ? 1 : 0;
    }
    char* h = sv->host_arena.at(hoff);
    r->h_rgb = reinterpret_cast<float*>(carve(h, 12 * P));
    r->h_w = reinterpret_cast<float*>(carve(h, 4 * P));
    r->h_ovf = reinterpret_cast<uint32_t*>(carve(h, 16));
    r->dev_off = doff; r->dev_len = dlen; r->host_off = hoff; r->host_len = hlen;
    r->slot = slot;
    p->slots[slot] = r;
    ReqDev& e = p->h_reqs[stage];
    std::memcpy(e.cam.m, r->cam.cam_to_world, sizeof e.cam.m);
    e.cam.depth = r->cam.depth;
    e.cam.width = r->cam.width;
    e.cam.height = r->cam.height;
    e.spp = r->prm.spp;
    e.max_depth = r->prm.integrator == CAV_INTEGRATOR_DIRECT ? 1 : r->prm.max_depth;
    e.integrator = r->prm.integrator;
    e.res = r->res;
    e.sxy = r->sxy;
    // the pool's earlier steps no longer read this slot (its previous owner
    // finished), so the stream-ordered copies are safe
    hipError_t he = hipMemcpyAsync(p->d_reqs + slot, &e, sizeof(ReqDev), hipMemcpyHostToDevice, p->stream);
    if (he == hipSuccess)
        he = hipMemcpyAsync(p->dev.rl.req + slot, &st, sizeof(RlReqState), hipMemcpyHostToDevice, p->stream);
    if (he == hipSuccess && r->rl()) he = pool_rl_init(p->dev, st, sd.nlights, p->stream);
    // the pool's overflow counters before the request's first bounce
    if (he == hipSuccess)
        he = hipMemcpyAsync(r->h_ovf, p->dev.trace_counters, 8, hipMemcpyDeviceToHost, p->stream);
    if (he != hipSuccess) sv->error = hip_msg("request admission", he);
    r->active = true;
    return true;
}

void release_request_memory(cav_server* sv, Request* r) {
    if (r->dev_len) sv->dev_arena.release(r->dev_off, r->dev_len);
    if (r->host_len) sv->host_arena.release(r->host_off, r->host_len);
    r->dev_len = r->host_len = 0;
}

// Records the pool's bounce sequences (each queue parity and bounce kind)
// into graphs, node by node (no stream capture). On failure the pool runs its
// launches directly and the reason is kept in `why`.
bool build_graphs(Pool* p, std::string& why) {
    for (int k = 0; k < 2 && why.empty(); ++k)
        for (int kind = kBouncePlain; kind <= kBounceRlDet && why.empty(); ++kind) {
            hipGraph_t g = nullptr;
            hipError_t e = hipGraphCreate(&g, 0);
            if (e != hipSuccess) { why = hip_msg("hipGraphCreate", e); break; }
            LaunchSink ls(g);
            pool_bounce(p->dev, k, kind, ls);
            if (ls.err != hipSuccess) {
                why = hip_msg((std::string("hipGraphAddKernelNode ") + ls.failed).c_str(), ls.err);
            } else if ((e = hipGraphInstantiate(&p->graph[k][kind], g, nullptr, nullptr, 0)) != hipSuccess) {
                p->graph[k][kind] = nullptr;
                why = hip_msg("hipGraphInstantiate", e);
            }
            (void)hipGraphDestroy(g);
        }
    p->graphs = why.empty();
    if (p->graphs) return true;
    for (auto& row : p->graph)
        for (auto& x : row)
            if (x) { (void)hipGraphExecDestroy(x); x = nullptr; }
    return false;
}

// Admission of one step: round-robin over the pending requests. A path or
// directlighting request gets one chunk (at most the quantum) and goes to the
// back, so room freed at the next step goes to the requests after it. An
// rlpath request gets its next round once the previous one has finished; a
// round that does not fit yet holds the queue, so the room it needs builds up.
int admit(cav_server* sv, Pool* p, Chunk* hc, int& nch, int64_t& total) {
    const int64_t cap = p->dev.cap;
    const size_t turns = p->pending.size();
    for (size_t t = 0; t < turns && nch < kSlots; ++t) {
        Request* r = p->pending.front();
        p->pending.pop_front();
        if (r->cancel && !r->active) {
            finalize(sv, r, CAV_SERVE_CANCELLED);
            continue;
        }
        if (r->rl() && r->outstanding > 0) {  // its current round is still in the pool
            p->pending.push_back(r);
            continue;
        }
        const int64_t room = cap - p->reserved;
        if (room <= 0) {
            p->pending.push_front(r);
            break;
        }
        int spp_now = 0;
        int64_t n = 0;
        if (r->rl()) {
            spp_now = std::min(r->round_spp, r->prm.spp - r->spp_admitted);
            n = r->npix * spp_now;
            if (n > room) {
                p->pending.push_front(r);
                break;
            }
        }
        if (!r->active && !activate(sv, p, r)) {
            p->pending.push_back(r);
            continue;
        }
        if (!sv->error.empty()) return CAV_ERR_HIP;
        if (r->t_start < 0) {
            r->t_start = now_ms(sv->t0);
            r->status = CAV_SERVE_RUNNING;
        }
        if (r->rl()) {
            hc[nch] = Chunk{r->slot, spp_now, r->spp_admitted, n, total};
            r->spp_admitted += spp_now;
        } else {
            n = std::min({room, sv->cfg.chunk_samples, r->nsamp - r->next_sample});
            hc[nch] = Chunk{r->slot, 0, r->next_sample, n, total};
            r->next_sample += n;
        }
        ++nch;
        total += n;
        r->outstanding += 1;
        p->reserved += n;
        const int64_t last = r->prm.integrator == CAV_INTEGRATOR_DIRECT ? 0 : r->prm.max_depth - 1;
        p->live.push_back({LiveChunk{r, p->step + last}, n});
        if (!r->all_admitted()) p->pending.push_back(r);
    }
    return CAV_OK;
}

// one step of one pool; called with sv->mu held (it touches request state),
// enqueues work only
int pool_step(cav_server* sv, Pool* p) {
    const int ring = (int)p->ring.size();
    // wait until the chunk/event ring slot of step - ring + 1 is free
    while (p->step - p->done_upto >= sv->cfg.max_steps_ahead) {
        const int64_t s = p->done_upto;
        const hipError_t e = hipEventSynchronize(p->ring[s % ring]);
        if (e != hipSuccess) { sv->error = hip_msg("hipEventSynchronize (bounce step)", e); return CAV_ERR_HIP; }
        p->done_upto = s + 1;
    }
    Chunk* hc = p->h_chunks + (size_t)(p->step % ring) * kSlots;
    int nch = 0;
    int64_t total = 0;
    if (admit(sv, p, hc, nch, total) != CAV_OK) return CAV_ERR_HIP;
    if (nch) {
        Chunk* dc = p->d_chunks + (size_t)(p->step % ring) * kSlots;
        hipError_t e = hipMemcpyAsync(dc, hc, sizeof(Chunk) * nch, hipMemcpyHostToDevice, p->stream);
        if (e != hipSuccess) { sv->error = hip_msg("hipMemcpyAsync (chunk list)", e); return CAV_ERR_HIP; }
        if ((e = pool_raygen(p->dev, dc, nch, total, p->qin, p->stream)) != hipSuccess) {
            sv->error = hip_msg("k_pool_raygen launch", e);
            return CAV_ERR_HIP;
        }
    }
    // the bounce kind: rlpath kernels only while rlpath tasks are in the
    // pool, record sorting only for deterministic ones
    int kind = kBouncePlain;
    for (const auto& lc : p->live)
        if (lc.first.req->rl()) kind = std::max(kind, lc.first.req->det() ? (int)kBounceRlDet : (int)kBounceRl);
    if (p->graphs) {
        const hipError_t e = hipGraphLaunch(p->graph[p->qin][kind], p->stream);
        if (e != hipSuccess) { sv->error = hip_msg("hipGraphLaunch (bounce step)", e); return CAV_ERR_HIP; }
    } else {
        LaunchSink ls(p->stream);
        pool_bounce(p->dev, p->qin, kind, ls);
        if (ls.err != hipSuccess) {
            sv->error = hip_msg((std::string(ls.failed) + " launch").c_str(), ls.err);
            return CAV_ERR_HIP;
        }
    }
    // batches whose last bounce was this step
    for (size_t k = 0; k < p->live.size();) {
        if (p->live[k].first.done_step != p->step) { ++k; continue; }
        Request* r = p->live[k].first.req;
        p->reserved -= p->live[k].second;
        p->live[k] = p->live.back();
        p->live.pop_back();
        if (--r->outstanding == 0 && r->all_admitted()) {
            film_launch(r->cam.width, r->cam.height, r->prm.spp, r->res, r->sxy, r->partial, r->rgb, r->w, p->stream);
            hipError_t e = hipGetLastError();
            const char* what = "film kernels";
            if (e == hipSuccess) {
                what = "hipMemcpyAsync (film to staging)";
                e = hipMemcpyAsync(r->h_rgb, r->rgb, 12 * (size_t)r->npix, hipMemcpyDeviceToHost, p->stream);
            }
            if (e == hipSuccess) e = hipMemcpyAsync(r->h_w, r->w, 4 * (size_t)r->npix, hipMemcpyDeviceToHost, p->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(r->h_ovf + 2, p->dev.trace_counters, 8, hipMemcpyDeviceToHost, p->stream);
            if (e == hipSuccess) {
                what = "request event";
                e = hipEventCreateWithFlags(&r->ready, hipEventDisableTiming);
            }
            if (e == hipSuccess) e = hipEventRecord(r->ready, p->stream);
            if (e != hipSuccess) { sv->error = hip_msg(what, e); return CAV_ERR_HIP; }
            r->finishing = true;
            p->slots[r->slot] = nullptr;
        }
    }
    const hipError_t e = hipEventRecord(p->ring[p->step % ring], p->stream);
    if (e != hipSuccess) { sv->error = hip_msg("hipEventRecord (bounce step)", e); return CAV_ERR_HIP; }
    ++p->step;
    p->qin ^= 1;
    return CAV_OK;
}

void finish_ready(cav_server* sv) {
    for (auto& kv : sv->reqs) {
        Request* r = kv.second.get();
        if (!r->finishing) continue;
        const hipError_t q = hipEventQuery(r->ready);
        if (q == hipErrorNotReady) continue;
        r->finishing = false;
        (void)hipEventDestroy(r->ready);
        r->ready = nullptr;
        // any traversal of the pool cut short while this request was in it
        // (the counter is pool-wide, so a neighbour's overflow fails it too)
        const bool ovf = r->h_ovf[2] != r->h_ovf[0] || r->h_ovf[3] != r->h_ovf[1];
        int32_t st = CAV_SERVE_DONE, err = CAV_OK;
        std::string msg;
        if (q != hipSuccess) {
            st = CAV_SERVE_FAILED;
            err = CAV_ERR_HIP;
            msg = hip_msg("request completion event", q);
        } else if (r->cancel) {
            st = CAV_SERVE_CANCELLED;
        } else if (ovf) {
            st = CAV_SERVE_FAILED;
            err = CAV_ERR_OVERFLOW;
            msg = "traversal stack overflow in the request's ray pool";
        } else {
            std::memcpy(r->out_rgb, r->h_rgb, 12 * (size_t)r->npix);
            std::memcpy(r->out_w, r->h_w, 4 * (size_t)r->npix);
        }
        release_request_memory(sv, r);
        finalize(sv, r, st, err, msg);
    }
}

void worker_main(cav_server* sv) {
    (void)hipSetDevice(sv->ctx->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        while (!sv->incoming.empty()) {
            Request* r = sv->incoming.front();
            sv->incoming.pop_front();
            Pool* p = pool_for(sv, r->scene);
            int rc = CAV_OK;
            std::string why;
            if (!p) {
                rc = make_pool(sv, r->scene, p);
                // without graphs the pool launches its kernels directly; the reason is kept
                if (rc == CAV_OK && sv->cfg.use_graphs && !build_graphs(p, why)) sv->graph_note = why;
            }
            if (rc != CAV_OK) {
                finalize(sv, r, CAV_SERVE_FAILED, rc, sv->error.empty() ? "ray pool allocation failed" : sv->error);
                continue;
            }
            p->pending.push_back(r);
        }
        bool busy = false;
        for (auto& up : sv->pools) {
            Pool* p = up.get();
            if (p->live.empty() && p->pending.empty()) continue;
            busy = true;
            if (pool_step(sv, p) != CAV_OK) {
                if (sv->error.empty()) sv->error = "a HIP call failed while serving";
                for (auto& kv : sv->reqs)
                    if (!kv.second->final) finalize(sv, kv.second.get(), CAV_SERVE_FAILED, CAV_ERR_HIP, sv->error);
                return;
            }
        }
        finish_ready(sv);
        bool finishing = false;
        for (auto& kv : sv->reqs) finishing = finishing || kv.second->finishing;
        if (sv->stop && !busy && !finishing && sv->incoming.empty()) break;
        if (busy) {
            // let waiting submitters and pollers in between steps
            if (sv->api_waiting.load() > 0) {
                lk.unlock();
                while (sv->api_waiting.load() > 0) std::this_thread::yield();
                lk.lock();
            }
        } else if (finishing) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(100));
            lk.lock();
        } else {
            sv->cv.wait(lk, [&] { return sv->stop || !sv->incoming.empty(); });
        }
    }
    for (auto& up : sv->pools) {
        unsigned long long t[2] = {0, 0};
        (void)hipStreamSynchronize(up->stream);
        if (hipMemcpy(t, up->dev.rays, 16, hipMemcpyDeviceToHost) == hipSuccess) {
            sv->pool_rays[0] += t[0];
            sv->pool_rays[1] += t[1];
        }
    }
}

}  // namespace

extern "C" {

int cav_serve_open(cav_ctx* ctx, const cav_serve_config* cfg, cav_server** out) {
    if (!ctx || !out) return CAV_ERR_INVALID;
    *out = nullptr;
    std::unique_ptr<cav_server> sv(new cav_server());
    sv->ctx = ctx;
    if (cfg) sv->cfg = *cfg;
    cav_serve_config& c = sv->cfg;
    if (c.arena_bytes <= 0) c.arena_bytes = (int64_t)16 << 30;
    if (c.pool_tasks <= 0) c.pool_tasks = (int64_t)4 << 20;
    if (c.max_steps_ahead <= 0) c.max_steps_ahead = 4;
    if (c.chunk_samples <= 0) c.chunk_samples = c.pool_tasks / 4;
    if (c.host_arena_bytes <= 0) c.host_arena_bytes = (int64_t)1 << 30;
    if (c.pool_tasks >= ((int64_t)1 << 31) / 3) return fail(ctx, CAV_ERR_INVALID, "pool_tasks must be below 2^31 / 3");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, CAV_ERR_HIP, "hipSetDevice");
    int rc = halton_dev(ctx, &sv->halton);
    if (rc != CAV_OK) return rc;
    if (hipMalloc(reinterpret_cast<void**>(&sv->dev_base), (size_t)c.arena_bytes) != hipSuccess)
        return fail(ctx, CAV_ERR_OOM, "server device arena of " + std::to_string(c.arena_bytes) + " bytes");
    if (hipHostMalloc(reinterpret_cast<void**>(&sv->host_base), (size_t)c.host_arena_bytes, 0) != hipSuccess) {
        (void)hipFree(sv->dev_base);
        return fail(ctx, CAV_ERR_OOM, "server pinned host arena");
    }
    sv->dev_arena.init(sv->dev_base, (size_t)c.arena_bytes);
    sv->host_arena.init(sv->host_base, (size_t)c.host_arena_bytes);
    sv->t0 = std::chrono::steady_clock::now();
    cav_server* raw = sv.get();
    raw->worker = std::thread(worker_main, raw);
    *out = sv.release();
    return CAV_OK;
}

int cav_serve_submit(cav_server* sv, cav_scene* scene, const cav_camera* cam, const cav_render_params* prm,
                     float* film_rgb, float* film_weight, int64_t* id) {
    if (!sv || !scene || !cam || !prm || !film_rgb || !film_weight || !id) return CAV_ERR_INVALID;
    if (prm->integrator != CAV_INTEGRATOR_PATH && prm->integrator != CAV_INTEGRATOR_RLPATH &&
        prm->integrator != CAV_INTEGRATOR_DIRECT)
        return fail(sv->ctx, CAV_ERR_INVALID, "unknown integrator");
    if (prm->spp < 1 || cam->width < 1 || cam->height < 1 ||
        (prm->integrator != CAV_INTEGRATOR_DIRECT && prm->max_depth < 1))
        return fail(sv->ctx, CAV_ERR_INVALID, "spp, max_depth and resolution must be positive");
    const int64_t npix = (int64_t)cam->width * cam->height;
    if (npix * prm->spp >= ((int64_t)1 << 32)) return fail(sv->ctx, CAV_ERR_INVALID, "more than 2^32 samples");
    if (prm->integrator == CAV_INTEGRATOR_PATH &&
        2 + 10LL * prm->max_depth + (prm->max_depth > 4 ? prm->max_depth - 4 : 0) > 1024)
        return fail(sv->ctx, CAV_ERR_INVALID, "max_depth exceeds the 1024 Halton dimensions");
    if (prm->integrator == CAV_INTEGRATOR_RLPATH && 2 + 14LL * prm->max_depth > 1024)
        return fail(sv->ctx, CAV_ERR_INVALID, "max_depth exceeds the 1024 Halton dimensions");
    if (prm->integrator != CAV_INTEGRATOR_DIRECT && prm->max_depth > 255)
        return fail(sv->ctx, CAV_ERR_INVALID, "max_depth above 255");
    if (prm->spp_per_round < 0) return fail(sv->ctx, CAV_ERR_INVALID, "spp_per_round must be >= 0");
    std::unique_ptr<Request> r(new Request());
    r->scene = scene;
    r->cam = *cam;
    r->prm = *prm;
    r->out_rgb = film_rgb;
    r->out_w = film_weight;
    r->npix = npix;
    r->nsamp = npix * prm->spp;
    ApiLock lock(sv);
    if (sv->stop) return fail(sv->ctx, CAV_ERR_INVALID, "the server is closing");
    if (!sv->error.empty()) return fail(sv->ctx, CAV_ERR_HIP, "the server failed: " + sv->error);
    if (r->rl()) {
        const int g = rl_grid(*prm);
        if (g < 1 || g > 64) return fail(sv->ctx, CAV_ERR_INVALID, "q_grid must be in [1, 64]");
        if (!(prm->alpha_scale >= 0.f && prm->alpha_scale <= 1.f))
            return fail(sv->ctx, CAV_ERR_INVALID, "alpha_scale must be in [0, 1]");
        r->grid = g;
        r->round_spp = rl_round_spp(sv->cfg, *prm, npix);
        if (npix * r->round_spp > sv->cfg.pool_tasks)
            return fail(sv->ctx, CAV_ERR_INVALID, "one rlpath round (every pixel) exceeds the pool's task slots");
    }
    const size_t need = request_bytes(*r, scene_dev(scene).nlights);
    if (need > (size_t)sv->cfg.arena_bytes || 16 * (size_t)npix + 3 * 256 > (size_t)sv->cfg.host_arena_bytes)
        return fail(sv->ctx, CAV_ERR_OOM, "request larger than the server's arenas");
    r->id = sv->next_id++;
    r->t_submit = now_ms(sv->t0);
    *id = r->id;
    Request* raw = r.get();
    scene_hold(scene, 1);  // until the request's final status
    sv->reqs.emplace(raw->id, std::move(r));
    sv->incoming.push_back(raw);
    sv->cv.notify_all();
    return CAV_OK;
}

int cav_serve_poll(cav_server* sv, int64_t id, int32_t wait, cav_serve_result* out) {
    if (!sv || !out) return CAV_ERR_INVALID;
    ApiLock lock(sv);
    std::unique_lock<std::mutex>& lk = lock.lk;
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return CAV_ERR_NOT_FOUND;
    Request* r = it->second.get();
    auto final_state = [&] {
        return r->status == CAV_SERVE_DONE || r->status == CAV_SERVE_CANCELLED || r->status == CAV_SERVE_FAILED;
    };
    if (wait) sv->done_cv.wait(lk, final_state);
    out->status = r->status;
    out->error = r->error;
    out->submit_ms = r->t_submit;
    out->start_ms = r->t_start;
    out->done_ms = r->t_done;
    out->closest_rays = 0;
    out->shadow_rays = 0;
    if (r->status == CAV_SERVE_FAILED) fail(sv->ctx, r->error, "request " + std::to_string(id) + ": " + r->msg);
    if (final_state()) sv->reqs.erase(it);  // the id is released with its final status
    return CAV_OK;
}

int cav_serve_cancel(cav_server* sv, int64_t id) {
    if (!sv) return CAV_ERR_INVALID;
    ApiLock lock(sv);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return CAV_ERR_NOT_FOUND;
    it->second->cancel = true;
    sv->cv.notify_all();
    return CAV_OK;
}

int cav_serve_close(cav_server* sv, cav_serve_stats* stats) {
    if (!sv) return CAV_ERR_INVALID;
    {
        ApiLock lock(sv);
        sv->stop = true;
        sv->cv.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    if (stats) {
        stats->pool_closest_rays = (int64_t)sv->pool_rays[0];
        stats->pool_shadow_rays = (int64_t)sv->pool_rays[1];
        stats->rl_closest_rays = 0;
        stats->rl_shadow_rays = 0;
        int64_t steps = 0;
        for (auto& p : sv->pools) steps += p->step;
        stats->steps = steps;
        stats->pools = (int32_t)sv->pools.size();
        int32_t g = 0;
        for (auto& p : sv->pools) g += p->graphs ? 1 : 0;
        stats->graph_pools = g;
    }
    for (auto& p : sv->pools) destroy_pool(p.get());
    for (auto& kv : sv->reqs) {
        Request* r = kv.second.get();
        if (r->ready) (void)hipEventDestroy(r->ready);
        // requests a failed server never finished still hold their scenes
        if (!r->final) finalize(sv, r, CAV_SERVE_FAILED, CAV_ERR_HIP, sv->error);
    }
    if (sv->dev_base) (void)hipFree(sv->dev_base);
    if (sv->host_base) (void)hipHostFree(sv->host_base);
    delete sv;
    return CAV_OK;
}

}  // extern "C"
