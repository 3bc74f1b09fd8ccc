// C ABI for scenes and rendering (cav_scene_*, cav_render); see
// include/cavalry.h for the reference counterparts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cavalry.h"
#include "capi_internal.h"
#include "render.h"

using namespace cav;

#define RCHK(ctx, call)                                                \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return cav::hip_fail((ctx), e_, #call);  \
    } while (0)

struct cav_scene {
    cav_ctx* ctx = nullptr;
    cav_bvh* bvh = nullptr;
    int ntris = 0, nmeshes = 0, nlights = 0;
    float* soup = nullptr;
    int32_t* tri_mesh = nullptr;
    cav_mesh* meshes = nullptr;
    float* area = nullptr;
    int32_t* lights = nullptr;
    float* uv = nullptr;
    float* normals = nullptr;        // float[ntris][9] or null
    uint8_t* mesh_normals = nullptr; // per mesh
    uint32_t* texels = nullptr;      // every texture's RGBA8 texels
    TexDev* tex = nullptr;
    EnvDev* env = nullptr;           // environment-type lights
    int env_miss = -1;
    int light_dims = 3;
    std::vector<cav_mesh> host_meshes;
    std::vector<float> host_area;
    float bounds[6] = {0, 0, 0, 0, 0, 0};  // Scene::sceneBounds (Core/Scene.cu:131-136)
    std::atomic<int> users{0};  // server pools and unfinished server requests on this scene
};

namespace {

// device bytes per sample of a round (queues, shadow tasks, hits) and per
// sample of the whole render (radiance + film position)
constexpr double kQueueBytes = 2 * (24 + 4 + 12 + 4 + 1) + (24 + 4 + 4 + 4 + 4 + 12) + 1 + 16 + 4;
constexpr double kRlQueueBytes = 2 * 12 + (4 * 5 + 12 * 3 + 4 + 4 + 4) + 16 + 2 * 4 * 256.0 / 4096;
constexpr double kResultBytes = 12 + 8;

// Per-context render state: Halton tables and the sample/queue buffers,
// grown on demand and kept between renders.
struct RenderState {
    uint32_t* primes = nullptr;
    uint32_t* start = nullptr;
    uint16_t* perm = nullptr;
    int64_t cap = 0;      // samples per round the queue buffers hold
    int64_t tot_cap = 0;  // samples the per-sample result buffers hold
    int64_t pix_cap = 0;  // pixels the film buffers hold
    std::vector<void*> bufs;
    RenderArgs a{};
    hipEvent_t ev[2] = {nullptr, nullptr};
    // rlpath: per-slot state (sized like the sample buffers) and the tables
    std::vector<void*> rl_bufs, rl_tab, rl_rec;
    int64_t rl_cap = 0, rl_cells = 0, rl_lcells = 0, rl_rec_cap = 0;
    RlArgs rl{};
};

void free_list(std::vector<void*>& v) {
    for (void* p : v) (void)hipFree(p);
    v.clear();
}

void free_bufs(RenderState* rs) {
    free_list(rs->bufs);
    rs->cap = 0;
    rs->tot_cap = 0;
    rs->pix_cap = 0;
    free_list(rs->rl_bufs);
    free_list(rs->rl_tab);
    free_list(rs->rl_rec);
    rs->rl_cap = rs->rl_cells = rs->rl_lcells = rs->rl_rec_cap = 0;
}

// Halton digit permutations (Samplers/HaltonSampler.cu:7-57): the first 1024
// primes, each shuffled by Fisher-Yates with the Microsoft C runtime rand()
// stream (state * 214013 + 2531011, bits 16..30) seeded with 533799.
void halton_tables(std::vector<uint32_t>& primes, std::vector<uint32_t>& start,
                   std::vector<uint16_t>& perm) {
    primes.clear();
    for (uint32_t c = 2; primes.size() < 1024; ++c) {
        bool p = true;
        for (uint32_t q : primes) {
            if (q * q > c) break;
            if (c % q == 0) { p = false; break; }
        }
        if (p) primes.push_back(c);
    }
    start.resize(1024);
    uint32_t tot = 0;
    for (int k = 0; k < 1024; ++k) { start[k] = tot; tot += primes[k]; }
    perm.resize(tot);
    uint32_t state = 533799u;
    for (int k = 0; k < 1024; ++k) {
        const uint32_t p = primes[k];
        uint16_t* v = perm.data() + start[k];
        for (uint32_t j = 0; j < p; ++j) v[j] = (uint16_t)j;
        for (int i = (int)p - 1; i > 0; --i) {
            state = state * 214013u + 2531011u;
            const int r = (int)((state >> 16) & 0x7fffu);
            std::swap(v[i], v[r % (i + 1)]);
        }
    }
}

int ensure_state(cav_ctx* ctx, RenderState*& rs) {
    if (!ctx->render) {
        rs = new RenderState();
        ctx->render = rs;
        std::vector<uint32_t> primes, start;
        std::vector<uint16_t> perm;
        halton_tables(primes, start, perm);
        int rc = dev_alloc(ctx, (void**)&rs->primes, 4 * 1024, "halton primes");
        if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&rs->start, 4 * 1024, "halton starts");
        if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&rs->perm, 2 * perm.size(), "halton perms");
        if (rc != CAV_OK) return rc;
        RCHK(ctx, hipMemcpy(rs->primes, primes.data(), 4 * 1024, hipMemcpyHostToDevice));
        RCHK(ctx, hipMemcpy(rs->start, start.data(), 4 * 1024, hipMemcpyHostToDevice));
        RCHK(ctx, hipMemcpy(rs->perm, perm.data(), 2 * perm.size(), hipMemcpyHostToDevice));
        RCHK(ctx, hipEventCreate(&rs->ev[0]));
        RCHK(ctx, hipEventCreate(&rs->ev[1]));
    }
    rs = static_cast<RenderState*>(ctx->render);
    return CAV_OK;
}

template <typename T>
int grab_into(cav_ctx* ctx, std::vector<void*>& list, T*& p, size_t bytes, const char* what) {
    void* q = nullptr;
    const int rc = dev_alloc(ctx, &q, bytes, what);
    if (rc != CAV_OK) return rc;
    list.push_back(q);
    p = static_cast<T*>(q);
    return CAV_OK;
}
template <typename T>
int grab(cav_ctx* ctx, RenderState* rs, T*& p, size_t bytes, const char* what) {
    return grab_into(ctx, rs->bufs, p, bytes, what);
}

// Deterministic rlpath: records of up to `cap` table updates per bounce and
// the work space of their radix sort.
int ensure_rl_rec(cav_ctx* ctx, RenderState* rs, int64_t cap) {
    RlArgs& r = rs->rl;
    if (cap <= rs->rl_rec_cap) return CAV_OK;
    (void)hipStreamSynchronize(ctx->stream);
    free_list(rs->rl_rec);
    rs->rl_rec_cap = 0;
    auto& L = rs->rl_rec;
    const size_t C = (size_t)cap, T = (size_t)sort_tiles(cap);
    int rc = CAV_OK;
    for (int k = 0; k < 2 && rc == CAV_OK; ++k) {
        rc = grab_into(ctx, L, r.rec_keys[k], 8 * C, "rl update keys");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.rec_idx[k], 4 * C, "rl update order");
    }
    if (rc == CAV_OK) rc = grab_into(ctx, L, r.rec_val, 4 * C, "rl update values");
    if (rc == CAV_OK) rc = grab_into(ctx, L, r.rec_count, 256, "rl update count");
    if (rc == CAV_OK) rc = grab_into(ctx, L, r.rec_hist, 4 * 8 * 256, "rl update sort histogram");
    if (rc == CAV_OK) rc = grab_into(ctx, L, r.rec_counts, 4 * 256 * T, "rl update sort counts");
    if (rc == CAV_OK) rc = grab_into(ctx, L, r.rec_offsets, 4 * 256 * T, "rl update sort offsets");
    if (rc != CAV_OK) { free_list(L); return rc; }
    rs->rl_rec_cap = cap;
    r.rec_cap = cap;
    return CAV_OK;
}

// rlpath buffers: per-slot state for nsamp tasks, a Q table of `cells` bins
// and light statistics of `lcells` (entry, light) pairs.
int ensure_rl(cav_ctx* ctx, RenderState* rs, int64_t nsamp, int64_t cells, int64_t lcells) {
    RlArgs& r = rs->rl;
    if (nsamp > rs->rl_cap) {
        (void)hipStreamSynchronize(ctx->stream);
        free_list(rs->rl_bufs);
        rs->rl_cap = 0;
        const size_t S = (size_t)nsamp;
        auto& L = rs->rl_bufs;
        int rc = CAV_OK;
        for (int k = 0; k < 2 && rc == CAV_OK; ++k) {
            rc = grab_into(ctx, L, r.q[k].prev_entry, 4 * S, "rl previous entry");
            if (rc == CAV_OK) rc = grab_into(ctx, L, r.q[k].prev_cell, 4 * S, "rl previous bin");
            if (rc == CAV_OK) rc = grab_into(ctx, L, r.q[k].dim, 4 * S, "rl dimension");
        }
        RlSlots& sl = r.slots;
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.entry, 4 * S, "rl entry");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.light, 4 * S, "rl light");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.sel, 4 * S, "rl light selection pdf");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.lpdf, 4 * S, "rl light pdf");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.emis, 4 * S, "rl emission");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.ldir, 12 * S, "rl light direction");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.inc, 12 * S, "rl incident");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.sh_idx, 4 * S, "rl shadow slot");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.ndir, 12 * S, "rl next direction");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.npdf, 4 * S, "rl next pdf");
        if (rc == CAV_OK) rc = grab_into(ctx, L, sl.ncell, 4 * S, "rl next bin");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.dims, 64, "rl dimensions");
        const size_t T = (size_t)sort_tiles(nsamp);
        for (int k = 0; k < 2 && rc == CAV_OK; ++k) {
            rc = grab_into(ctx, L, r.skeys[k], 4 * S, "rl shading-order keys");
            if (rc == CAV_OK) rc = grab_into(ctx, L, r.svals[k], 4 * S, "rl shading order");
        }
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.shist, 4 * 8 * 256, "rl shading-order histogram");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.scounts, 4 * 256 * T, "rl shading-order counts");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.soffsets, 4 * 256 * T, "rl shading-order offsets");
        if (rc != CAV_OK) { free_list(L); return rc; }
        rs->rl_cap = nsamp;
    }
    if (cells > rs->rl_cells || lcells > rs->rl_lcells) {
        (void)hipStreamSynchronize(ctx->stream);
        free_list(rs->rl_tab);
        rs->rl_cells = rs->rl_lcells = 0;
        auto& L = rs->rl_tab;
        const size_t C = (size_t)cells, LC = (size_t)lcells;
        int rc = grab_into(ctx, L, r.Q, 4 * C, "Q table");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.newQ, 4 * C, "Q proposals");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.cnt, 4 * C, "Q proposal counts");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.total, 4 * C, "Q total counts");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.ldist, 4 * LC, "light CDFs");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.lres, 4 * LC, "light results");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.lnew, 4 * LC, "light new results");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.lcount, 4 * LC, "light counts");
        if (rc == CAV_OK) rc = grab_into(ctx, L, r.lnewc, 4 * LC, "light new counts");
        if (rc != CAV_OK) { free_list(L); return rc; }
        rs->rl_cells = cells;
        rs->rl_lcells = lcells;
    }
    return CAV_OK;
}

// Queue-sized buffers for nq samples per round; per-sample results for the
// whole render (ntot samples) and the film (npix pixels).
int ensure_buffers(cav_ctx* ctx, RenderState* rs, int64_t nq, int64_t ntot, int64_t npix) {
    if (nq <= rs->cap && ntot <= rs->tot_cap && npix <= rs->pix_cap) return CAV_OK;
    (void)hipStreamSynchronize(ctx->stream);
    free_bufs(rs);
    const size_t S = (size_t)nq, T = (size_t)ntot, P = (size_t)npix;
    RenderArgs& a = rs->a;
    int rc = CAV_OK;
    for (int k = 0; k < 2 && rc == CAV_OK; ++k) {
        rc = grab(ctx, rs, a.q[k].ray, 24 * S, "ray queue");
        if (rc == CAV_OK) rc = grab(ctx, rs, a.q[k].sid, 4 * S, "ray queue sid");
        if (rc == CAV_OK) rc = grab(ctx, rs, a.q[k].mult, 12 * S, "ray queue throughput");
        if (rc == CAV_OK) rc = grab(ctx, rs, a.q[k].pdf, 4 * S, "ray queue pdf");
        if (rc == CAV_OK) rc = grab(ctx, rs, a.q[k].delta, S, "ray queue delta");
    }
    if (rc == CAV_OK) rc = grab(ctx, rs, a.sh.ray, 24 * S, "shadow rays");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.sh.limit, 4 * S, "shadow limits");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.sh.src, 4 * S, "shadow src");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.sh.dst, 4 * S, "shadow dst");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.sh.sid, 4 * S, "shadow sid");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.sh.contrib, 12 * S, "shadow contributions");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.vis, S, "visibility");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.hit_prim, 4 * S, "hit prim");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.hit_t, 4 * S, "hit t");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.hit_u, 4 * S, "hit u");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.hit_v, 4 * S, "hit v");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.res, 12 * T, "sample radiance");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.sxy, 8 * T, "sample positions");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.counts, 64, "queue counts");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.rays, 64, "ray tallies");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.trace_counters, 128, "trace counters");
    unsigned char* coop = nullptr;
    const uint32_t cc = any_spill_capacity(nq);
    if (rc == CAV_OK) rc = grab(ctx, rs, coop, trace_coop_bytes(cc, cc), "long-ray pass buffers");
    if (rc == CAV_OK) a.coop = trace_coop_in(coop, cc, cc);
    if (rc == CAV_OK) rc = grab(ctx, rs, a.film_rgb, 12 * P, "film");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.film_w, 4 * P, "film weights");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.film_partial, 9 * 32 * P, "film partial sums");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.sort_perm, 4 * S, "material order");
    if (rc == CAV_OK) rc = grab(ctx, rs, a.sort_hist, 4 * (size_t)sort_hist_words(nq), "material histogram");
    if (rc != CAV_OK) { free_bufs(rs); return rc; }
    rs->cap = nq;
    rs->tot_cap = ntot;
    rs->pix_cap = npix;
    return CAV_OK;
}

}  // namespace

namespace cav {
void scene_hold(cav_scene* sc, int delta) { sc->users.fetch_add(delta); }
void scene_bounds(const cav_scene* sc, float out[6]) { std::memcpy(out, sc->bounds, sizeof sc->bounds); }

SceneDev scene_dev(const cav_scene* sc) {
    return SceneDev{sc->bvh->rec,  sc->bvh->root, sc->ntris,   sc->soup,
                    sc->tri_mesh,   sc->meshes,   sc->area,      sc->lights,  sc->nmeshes,
                    sc->nlights,    sc->normals,  sc->mesh_normals, sc->uv,   sc->texels,
                    sc->tex,        sc->env,      sc->env_miss,     sc->light_dims};
}

int halton_dev(cav_ctx* ctx, HaltonDev* out) {
    RenderState* rs = nullptr;
    const int rc = ensure_state(ctx, rs);
    if (rc == CAV_OK) *out = HaltonDev{rs->primes, rs->start, rs->perm};
    return rc;
}

void render_release(cav_ctx* ctx) {
    RenderState* rs = static_cast<RenderState*>(ctx->render);
    if (!rs) return;
    free_bufs(rs);
    for (void* p : {(void*)rs->primes, (void*)rs->start, (void*)rs->perm})
        if (p) (void)hipFree(p);
    for (auto& e : rs->ev)
        if (e) (void)hipEventDestroy(e);
    delete rs;
    ctx->render = nullptr;
}
}  // namespace cav

extern "C" {

int cav_scene_create(cav_ctx* ctx, const float* tris, int n, const cav_mesh* meshes, int nmeshes,
                     const float* tri_uv, int bvh_opt_rounds, int morton_bits, cav_scene** out) {
    cav_scene_desc d{};
    d.tris = tris;
    d.ntris = n;
    d.meshes = meshes;
    d.nmeshes = nmeshes;
    d.tri_uv = tri_uv;
    d.bvh_opt_rounds = bvh_opt_rounds;
    d.morton_bits = morton_bits;
    return cav_scene_create_desc(ctx, &d, out);
}

int cav_scene_create_desc(cav_ctx* ctx, const cav_scene_desc* desc, cav_scene** out) {
    if (!ctx || !out || !desc) return CAV_ERR_INVALID;
    *out = nullptr;
    const float* tris = desc->tris;
    const int n = desc->ntris;
    const cav_mesh* meshes = desc->meshes;
    const int nmeshes = desc->nmeshes;
    const float* tri_uv = desc->tri_uv;
    if (!tris || !meshes || n < 2 || nmeshes < 1) return fail(ctx, CAV_ERR_INVALID, "scene needs >= 2 triangles and >= 1 mesh");
    // meshes must tile the soup in order
    std::vector<int32_t> tri_mesh(n, -1);
    int next = 0;
    for (int m = 0; m < nmeshes; ++m) {
        const cav_mesh& M = meshes[m];
        if (M.first_tri != next || M.num_tris < 1 || M.first_tri + M.num_tris > n)
            return fail(ctx, CAV_ERR_INVALID, "meshes must be contiguous triangle ranges covering the soup in order");
        if (M.material != CAV_MAT_MATTE && M.material != CAV_MAT_MIRROR && M.material != CAV_MAT_GLASS &&
            M.material != CAV_MAT_PLASTIC && M.material != CAV_MAT_METAL && M.material != CAV_MAT_SUBSTRATE)
            return fail(ctx, CAV_ERR_INVALID, "unsupported material " + std::to_string(M.material));
        for (int t = M.first_tri; t < M.first_tri + M.num_tris; ++t) tri_mesh[t] = m;
        next += M.num_tris;
    }
    if (next != n) return fail(ctx, CAV_ERR_INVALID, "meshes do not cover every triangle");
    // the light list: area-light meshes and environment-type lights in the
    // scene's order (Scene::lightsHost)
    const int nenv = desc->nenv;
    if (nenv < 0 || (nenv > 0 && !desc->env_lights)) return fail(ctx, CAV_ERR_INVALID, "bad environment light list");
    std::vector<int32_t> lights;
    if (desc->light_order) {
        std::vector<uint8_t> seen_mesh((size_t)nmeshes, 0), seen_env((size_t)nenv, 0);
        for (int k = 0; k < desc->nlight_order; ++k) {
            const int32_t e = desc->light_order[k];
            const bool ok = e >= 0 ? (e < nmeshes && meshes[e].is_light && !seen_mesh[(size_t)e]++)
                                   : (-1 - e < nenv && !seen_env[(size_t)(-1 - e)]++);
            if (!ok) return fail(ctx, CAV_ERR_INVALID, "light_order entry " + std::to_string(k) + " is not a new light");
            lights.push_back(e);
        }
        for (int m = 0; m < nmeshes; ++m)
            if (meshes[m].is_light && !seen_mesh[(size_t)m])
                return fail(ctx, CAV_ERR_INVALID, "light_order misses area light mesh " + std::to_string(m));
        for (int k = 0; k < nenv; ++k)
            if (!seen_env[(size_t)k]) return fail(ctx, CAV_ERR_INVALID, "light_order misses environment light " + std::to_string(k));
    } else {
        for (int m = 0; m < nmeshes; ++m)
            if (meshes[m].is_light) lights.push_back(m);
        for (int k = 0; k < nenv; ++k) lights.push_back(-1 - k);
    }
    const int nlights = (int)lights.size();
    if (nlights == 0) return fail(ctx, CAV_ERR_INVALID, "scene has no light");
    const int background = nenv > 0 ? desc->env_background : -1;  // ignored without environment lights
    if (background < -1 || background >= nenv)
        return fail(ctx, CAV_ERR_INVALID, "env_background must name an environment light or be -1");
    // textures: packed into one texel buffer
    const int ntex = desc->ntextures;
    if (ntex < 0 || (ntex > 0 && !desc->textures)) return fail(ctx, CAV_ERR_INVALID, "bad texture list");
    std::vector<TexDev> tex_host((size_t)ntex);
    int64_t ntexels = 0;
    for (int k = 0; k < ntex; ++k) {
        const cav_texture& T = desc->textures[k];
        // sides up to 2^16 keep u W - 0.5 below W - 0.5 in float for every u < 1
        if (!T.rgba || T.width < 1 || T.height < 1 || T.width > 65536 || T.height > 65536)
            return fail(ctx, CAV_ERR_INVALID, "texture " + std::to_string(k) + " must be 1 to 65536 texels on each side");
        tex_host[(size_t)k] = TexDev{ntexels, T.width, T.height};
        ntexels += (int64_t)T.width * T.height;
    }
    bool any_tex = false;
    for (int m = 0; m < nmeshes; ++m) {
        const int32_t ti = meshes[m].texture;
        if (ti < -1 || ti >= ntex)
            return fail(ctx, CAV_ERR_INVALID, "mesh " + std::to_string(m) + " names texture " + std::to_string(ti) +
                                                  " of " + std::to_string(ntex));
        any_tex = any_tex || ti >= 0;
    }
    for (int k = 0; k < nenv; ++k) {
        const int32_t ti = desc->env_lights[k].texture;
        if (ti < -1 || ti >= ntex)
            return fail(ctx, CAV_ERR_INVALID, "environment light " + std::to_string(k) + " names texture " +
                                                  std::to_string(ti) + " of " + std::to_string(ntex));
        any_tex = any_tex || ti >= 0;
    }
    RCHK(ctx, hipSetDevice(ctx->device));
    cav_scene* sc = new cav_scene();
    // union of the triangle boxes, min/max as IEEE minimumNumber/maximumNumber
    for (int k = 0; k < 3; ++k) {
        sc->bounds[k] = tris[k];
        sc->bounds[3 + k] = tris[k];
    }
    for (int64_t t = 0; t < n; ++t)
        for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {
                const float v = tris[t * 9 + j * 3 + k];
                if (v < sc->bounds[k] || (v == sc->bounds[k] && std::signbit(v))) sc->bounds[k] = v;
                if (v > sc->bounds[3 + k] || (v == sc->bounds[3 + k] && !std::signbit(v))) sc->bounds[3 + k] = v;
            }
    sc->ctx = ctx;
    sc->ntris = n;
    sc->nmeshes = nmeshes;
    sc->nlights = nlights;
    sc->host_meshes.assign(meshes, meshes + nmeshes);
    // TriangleMesh::computeArea (Core/TriangleMesh.cu:189-199)
    sc->host_area.resize(nmeshes);
    for (int m = 0; m < nmeshes; ++m) {
        float a = 0;
        for (int t = meshes[m].first_tri; t < meshes[m].first_tri + meshes[m].num_tris; ++t) {
            const float* p = tris + (int64_t)t * 9;
            const float e1x = p[3] - p[0], e1y = p[4] - p[1], e1z = p[5] - p[2];
            const float e2x = p[6] - p[0], e2y = p[7] - p[1], e2z = p[8] - p[2];
            const float cx = e1y * e2z - e1z * e2y, cy = e1z * e2x - e1x * e2z, cz = e1x * e2y - e1y * e2x;
            a += std::sqrt(cx * cx + cy * cy + cz * cz) * 0.5f;
        }
        sc->host_area[m] = a;
    }
    sc->env_miss = background;
    // Halton dimensions of one light sample after randInt + rand4: the
    // reference's syncDimension takes each bounce to the queue's maximum, 3
    // (triangle, then a point in it) once any task samples a triangle of an
    // area light, else 2 (a direction: environment lights, a sphere light
    // seen from outside). Fixed per scene: 3 with any area light. A hit on a
    // sphere light's own (tessellated, so inside) surface samples its
    // triangles, so a bounce whose tasks all draw directions is the one case
    // where the reference uses 2 and this uses 3.
    sc->light_dims = 2;
    for (int32_t e : lights) sc->light_dims = e >= 0 ? 3 : sc->light_dims;
    int rc = dev_alloc(ctx, (void**)&sc->soup, 36 * (size_t)n, "scene soup");
    if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&sc->tri_mesh, 4 * (size_t)n, "triangle meshes");
    if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&sc->meshes, sizeof(cav_mesh) * nmeshes, "meshes");
    if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&sc->area, 4 * (size_t)nmeshes, "mesh areas");
    if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&sc->lights, 4 * (size_t)nlights, "lights");
    static_assert(sizeof(EnvDev) == sizeof(cav_env_light), "EnvDev mirrors cav_env_light");
    if (rc == CAV_OK && nenv > 0) rc = dev_alloc(ctx, (void**)&sc->env, sizeof(EnvDev) * (size_t)nenv, "environment lights");
    if (rc == CAV_OK && tri_uv) rc = dev_alloc(ctx, (void**)&sc->uv, 24 * (size_t)n, "uv");
    std::vector<uint8_t> mflags;
    if (desc->tri_normals) {
        mflags.assign(nmeshes, 1);
        if (desc->mesh_normals) mflags.assign(desc->mesh_normals, desc->mesh_normals + nmeshes);
        if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&sc->normals, 36 * (size_t)n, "vertex normals");
        if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&sc->mesh_normals, (size_t)nmeshes, "mesh normal flags");
    }
    if (any_tex) {
        if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&sc->tex, sizeof(TexDev) * (size_t)ntex, "textures");
        if (rc == CAV_OK) rc = dev_alloc(ctx, (void**)&sc->texels, 4 * (size_t)ntexels, "texels");
    }
    hipError_t e = hipSuccess;
    if (rc == CAV_OK && any_tex) {
        e = hipMemcpy(sc->tex, tex_host.data(), sizeof(TexDev) * (size_t)ntex, hipMemcpyHostToDevice);
        for (int k = 0; k < ntex && e == hipSuccess; ++k) {
            const cav_texture& T = desc->textures[k];
            e = hipMemcpy(sc->texels + tex_host[(size_t)k].off, T.rgba, 4 * (size_t)T.width * T.height,
                          hipMemcpyHostToDevice);
        }
        if (e != hipSuccess) rc = hip_fail(ctx, e, "upload textures");
    }
    if (rc == CAV_OK) {
        e = hipMemcpy(sc->soup, tris, 36 * (size_t)n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(sc->tri_mesh, tri_mesh.data(), 4 * (size_t)n, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(sc->meshes, meshes, sizeof(cav_mesh) * nmeshes, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(sc->area, sc->host_area.data(), 4 * (size_t)nmeshes, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(sc->lights, lights.data(), 4 * (size_t)nlights, hipMemcpyHostToDevice);
        if (e == hipSuccess && nenv > 0)
            e = hipMemcpy(sc->env, desc->env_lights, sizeof(EnvDev) * (size_t)nenv, hipMemcpyHostToDevice);
        if (e == hipSuccess && tri_uv) e = hipMemcpy(sc->uv, tri_uv, 24 * (size_t)n, hipMemcpyHostToDevice);
        if (e == hipSuccess && desc->tri_normals)
            e = hipMemcpy(sc->normals, desc->tri_normals, 36 * (size_t)n, hipMemcpyHostToDevice);
        if (e == hipSuccess && desc->tri_normals)
            e = hipMemcpy(sc->mesh_normals, mflags.data(), (size_t)nmeshes, hipMemcpyHostToDevice);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "upload scene");
    }
    if (rc == CAV_OK)
        rc = cav_bvh_build_device(ctx, sc->soup, n, nullptr, sc->tri_mesh, desc->morton_bits, desc->bvh_opt_rounds,
                                  &sc->bvh);
    if (rc != CAV_OK) { cav_scene_destroy(sc); return rc; }
    *out = sc;
    return CAV_OK;
}

int cav_scene_destroy(cav_scene* sc) {
    if (!sc) return CAV_ERR_INVALID;
    if (sc->users.load() > 0)
        return fail(sc->ctx, CAV_ERR_BUSY, "the scene is in use by a server (close the server first)");
    if (sc->ctx) (void)hipStreamSynchronize(sc->ctx->stream);
    if (sc->bvh) cav_bvh_destroy(sc->bvh);
    for (void* p : {(void*)sc->soup, (void*)sc->tri_mesh, (void*)sc->meshes, (void*)sc->area,
                    (void*)sc->lights, (void*)sc->uv, (void*)sc->normals, (void*)sc->mesh_normals,
                    (void*)sc->texels, (void*)sc->tex, (void*)sc->env})
        if (p) (void)hipFree(p);
    delete sc;
    return CAV_OK;
}

const cav_bvh* cav_scene_bvh(const cav_scene* sc) { return sc ? sc->bvh : nullptr; }

int cav_render(cav_ctx* ctx, cav_scene* sc, const cav_camera* cam, const cav_render_params* prm,
               float* film_rgb, float* film_weight, cav_render_stats* stats) {
    if (!ctx || !sc || !cam || !prm || !film_rgb || !film_weight) return CAV_ERR_INVALID;
    const bool rl = prm->integrator == CAV_INTEGRATOR_RLPATH;
    const bool direct = prm->integrator == CAV_INTEGRATOR_DIRECT;
    if (prm->integrator != CAV_INTEGRATOR_PATH && !rl && !direct)
        return fail(ctx, CAV_ERR_INVALID, "unknown integrator " + std::to_string(prm->integrator));
    const int grid = prm->q_grid == 0 ? 32 : prm->q_grid;
    const float alpha_scale = prm->alpha_scale == 0.f ? 1.f : prm->alpha_scale;
    if (rl && (grid < 1 || grid > 64))
        return fail(ctx, CAV_ERR_INVALID, "q_grid must be in [1, 64]");
    if (rl && !(alpha_scale > 0.f && alpha_scale <= 1.f))
        return fail(ctx, CAV_ERR_INVALID, "alpha_scale must be in (0, 1]");
    if (prm->spp < 1 || (prm->max_depth < 1 && !direct) || cam->width < 1 || cam->height < 1)
        return fail(ctx, CAV_ERR_INVALID, "spp, max_depth and resolution must be positive");
    const int64_t npix = (int64_t)cam->width * cam->height;
    const int64_t nsamp = npix * prm->spp;
    if (nsamp >= (int64_t)1 << 32) return fail(ctx, CAV_ERR_INVALID, "more than 2^32 samples in one render");
    // dimensions used: path 2 + 10 per bounce + 1 per bounce after the fourth;
    // rlpath at most 2 + 14 per bounce
    const int64_t dims = rl ? 2 + 14LL * prm->max_depth
                            : 2 + 10LL * prm->max_depth + (prm->max_depth > 4 ? prm->max_depth - 4 : 0);
    if (dims > 1024) return fail(ctx, CAV_ERR_INVALID, "max_depth exceeds the 1024 Halton dimensions");
    if (prm->spp_per_round < 0) return fail(ctx, CAV_ERR_INVALID, "spp_per_round must be >= 0");
    RCHK(ctx, hipSetDevice(ctx->device));
    RenderState* rs = nullptr;
    int rc = ensure_state(ctx, rs);
    if (rc != CAV_OK) return rc;
    // spp per round (Samplers/DecideSampleCount.cu:7-23, sppBySpp order): as
    // many as the free device memory holds with 10% slack, or the caller's
    // fixed count
    int spp_round = prm->spp;
    if (prm->spp_per_round > 0) {
        spp_round = std::min(prm->spp_per_round, prm->spp);
    } else {
        size_t fr = 0, tot = 0;
        RCHK(ctx, hipMemGetInfo(&fr, &tot));
        const double held = (double)rs->cap * kQueueBytes + (rl ? (double)rs->rl_cap * kRlQueueBytes : 0.0);
        const double per_spp = (double)npix * (kQueueBytes + (rl ? kRlQueueBytes : 0)) * 1.1;
        const double avail = (double)fr + held - (double)nsamp * kResultBytes - (512.0 * (1 << 20));
        const double fit = avail / per_spp;
        if (fit < 1.0) return fail(ctx, CAV_ERR_OOM, "one sample per pixel does not fit in device memory");
        spp_round = (int)std::min<double>(prm->spp, fit);
    }
    const int64_t nq = npix * spp_round;
    if ((rc = ensure_buffers(ctx, rs, nq, nsamp, npix)) != CAV_OK) return rc;
    const int64_t nent = (int64_t)grid * grid * grid;
    if (rl && (rc = ensure_rl(ctx, rs, nq, nent * 128, nent * sc->nlights)) != CAV_OK) return rc;
    const bool det = rl && (prm->flags & CAV_RENDER_DETERMINISTIC);
    if (det) {
        // update keys hold a Q bin or an (entry, light) index in 25 bits
        if (nent * 128 > ((int64_t)1 << 25) || nent * sc->nlights > ((int64_t)1 << 25))
            return fail(ctx, CAV_ERR_INVALID, "deterministic rlpath: more than 2^25 Q bins or light statistics");
        // at most 3 updates per task: light statistics, previous bin, direct proposal
        if ((rc = ensure_rl_rec(ctx, rs, 3 * nq)) != CAV_OK) return rc;
    }
    RenderArgs a = rs->a;
    a.scene = scene_dev(sc);
    a.launch = ctx->launch[CAV_TRACE_WAVEFRONT];
    a.halton = HaltonDev{rs->primes, rs->start, rs->perm};
    std::memcpy(a.cam.m, cam->cam_to_world, sizeof a.cam.m);
    a.cam.depth = cam->depth;
    a.cam.width = cam->width;
    a.cam.height = cam->height;
    a.width = cam->width;
    a.height = cam->height;
    a.spp = prm->spp;
    a.max_depth = prm->max_depth;
    a.nsamp = nsamp;
    if (!(prm->flags & CAV_RENDER_SORT_BY_MATERIAL)) a.sort_perm = nullptr;
    a.film_f64 = (prm->flags & CAV_RENDER_FILM_F64) != 0;
    hipStream_t s = ctx->stream;
    RCHK(ctx, hipMemsetAsync(a.rays, 0, 16, s));
    RCHK(ctx, hipMemsetAsync(a.trace_counters, 0, 32, s));
    RCHK(ctx, hipEventRecord(rs->ev[0], s));
    RlArgs r = rs->rl;
    if (!det) {
        r.rec_keys[0] = r.rec_keys[1] = nullptr;
        r.rec_cap = 0;
    } else {
        r.rec_cap = 3 * nq;
    }
    r.grid = grid;
    r.alpha_scale = alpha_scale;
    std::memcpy(r.bounds, sc->bounds, sizeof r.bounds);
    int rounds = 0;
    for (int done = 0; done < prm->spp; done += spp_round, ++rounds) {
        a.spp_done = done;
        a.spp_round = std::min(spp_round, prm->spp - done);
        a.nq = npix * a.spp_round;
        if (rl) RCHK(ctx, render_rlpath(a, r, done == 0, s));
        else if (direct) RCHK(ctx, render_direct(a, s));
        else RCHK(ctx, render_path(a, s));
    }
    render_film(a, s);
    RCHK(ctx, hipGetLastError());
    RCHK(ctx, hipEventRecord(rs->ev[1], s));
    unsigned long long tally[2] = {0, 0};
    uint32_t tc[2] = {0, 0};
    RCHK(ctx, hipMemcpyAsync(film_rgb, a.film_rgb, 12 * (size_t)npix, hipMemcpyDeviceToHost, s));
    RCHK(ctx, hipMemcpyAsync(film_weight, a.film_w, 4 * (size_t)npix, hipMemcpyDeviceToHost, s));
    RCHK(ctx, hipMemcpyAsync(tally, a.rays, 16, hipMemcpyDeviceToHost, s));
    RCHK(ctx, hipMemcpyAsync(tc, a.trace_counters, 8, hipMemcpyDeviceToHost, s));
    RCHK(ctx, hipStreamSynchronize(s));
    float ms = 0.f;
    RCHK(ctx, hipEventElapsedTime(&ms, rs->ev[0], rs->ev[1]));
    if (stats) {
        stats->ms_total = ms;
        stats->closest_rays = (int64_t)tally[0];
        stats->shadow_rays = (int64_t)tally[1];
        stats->mrays_per_s = ms > 0 ? (double)(tally[0] + tally[1]) / (ms * 1e3) : 0.0;
        stats->rounds = rounds;
    }
    if (tc[0] || tc[1]) return fail(ctx, CAV_ERR_OVERFLOW, "traversal stack overflow while rendering");
    return CAV_OK;
}

}  // extern "C"
