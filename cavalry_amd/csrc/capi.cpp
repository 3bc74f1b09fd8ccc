// C ABI implementation: contexts, BVH objects and traversal entry points.
// Declarations and their reference counterparts are in include/cavalry.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cavalry.h"
#include "capi_internal.h"
#include "lbvh.h"

using namespace cav;

const char* cav_version(void) { return "cavalry_amd 0.1 (gfx950)"; }

namespace cav {

int fail(cav_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(cav_ctx* ctx, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(ctx, CAV_ERR_HIP, m);
}

int dev_alloc(cav_ctx* ctx, void** p, size_t bytes, const char* what) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(ctx, CAV_ERR_OOM, std::string("hipMalloc ") + what + " (" + std::to_string(bytes) + " B): " + hipGetErrorString(e));
    }
    return CAV_OK;
}

}  // namespace cav

#define HIPCHK(ctx, call)                                              \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return cav::hip_fail((ctx), e_, #call);  \
    } while (0)

static void bvh_free(cav_bvh* h) {
    void* ptrs[] = {h->b.keys[0], h->b.keys[1], h->b.vals[0], h->b.vals[1], h->b.hist, h->b.counts,
                    h->b.scene_ord, h->b.nbox, h->b.topo, h->b.visited, h->b.ref,
                    h->b.partial, h->b.leaf_ids, h->b.counters, h->rec, h->root, h->mesh,
                    h->b.lvl_list, h->b.lvl_cnt, h->b.lvl_cursor, h->b.lvl_off,
                    h->b.bounds_part, h->b.gate64};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (h->spill) (void)hipFree(h->spill);
    if (h->b.h_lvl) (void)hipHostFree(h->b.h_lvl);
    delete[] h->b.h_off;
    if (h->events)
        for (auto& e : h->mk.ev) (void)hipEventDestroy(e);
}

static int bvh_alloc(cav_ctx* ctx, cav_bvh* h, int n) {
    const int64_t m = 2 * (int64_t)n - 1;
    const int64_t tiles = sort_tiles(n);
    struct { void** p; size_t bytes; const char* what; } req[] = {
        {&h->b.keys[0], sizeof(uint64_t) * n, "keys0"},
        {&h->b.keys[1], sizeof(uint64_t) * n, "keys1"},
        {(void**)&h->b.vals[0], sizeof(uint32_t) * n, "vals0"},
        {(void**)&h->b.vals[1], sizeof(uint32_t) * n, "vals1"},
        {(void**)&h->b.hist, sizeof(uint32_t) * (kSortHistWords + 16), "sort histograms"},
        {(void**)&h->b.counts, sizeof(uint32_t) * 8 * 256 * tiles, "sort look-back"},
        {(void**)&h->b.scene_ord, 64, "scene"},
        {(void**)&h->b.bounds_part, sizeof(uint32_t) * 6 * 1024, "bounds partials"},
        {(void**)&h->b.nbox, sizeof(NodeBox) * m, "nbox"},
        {(void**)&h->b.topo, sizeof(int4) * m, "topo"},
        {(void**)&h->b.visited, sizeof(uint32_t) * m, "visited"},
        {(void**)&h->b.gate64, sizeof(uint64_t) * bottom_up_gate_words(n), "split gates"},
        {(void**)&h->b.ref, sizeof(int32_t) * m, "ref"},
        {(void**)&h->b.partial, sizeof(uint32_t) * (lbvh_partials(m) + 1), "partial"},
        {(void**)&h->b.leaf_ids, sizeof(uint32_t) * n, "leaf_ids"},
        {(void**)&h->b.counters, 256, "counters"},
        {(void**)&h->rec, sizeof(TRec) * 2 * (size_t)n, "traversal records"},
        {(void**)&h->root, 64, "root"},
        {(void**)&h->b.lvl_list, sizeof(uint32_t) * m, "level list"},
        {(void**)&h->b.lvl_cnt, sizeof(uint32_t) * (kMaxLevels + 2), "level counts"},
        {(void**)&h->b.lvl_cursor, sizeof(uint32_t) * (kMaxLevels + 2), "level cursors"},
        {(void**)&h->b.lvl_off, sizeof(uint32_t) * (kMaxLevels + 2), "level offsets"},
    };
    for (auto& r : req) {
        int rc = dev_alloc(ctx, r.p, r.bytes, r.what);
        if (rc != CAV_OK) return rc;
    }
    // the bottom-up build leaves every split gate at zero for the next build;
    // the traversal counters only ever add to (or max into) their words
    HIPCHK(ctx, hipMemsetAsync(h->b.gate64, 0, sizeof(uint64_t) * bottom_up_gate_words(n), ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(h->b.counters, 0, 256, ctx->stream));
    HIPCHK(ctx, hipHostMalloc((void**)&h->b.h_lvl, sizeof(uint32_t) * (kMaxLevels + 2), hipHostMallocDefault));
    h->b.h_off = new uint32_t[kMaxLevels + 2]();
    for (auto& e : h->mk.ev) HIPCHK(ctx, hipEventCreate(&e));
    h->events = true;
    h->mk.on = true;
    return CAV_OK;
}

static void host_to_ord(const float* f6, uint32_t* o6) {
    for (int i = 0; i < 6; ++i) {
        uint32_t u;
        std::memcpy(&u, &f6[i], 4);
        o6[i] = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    }
}

static float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) { (void)hipGetLastError(); return -1.f; }
    return ms;
}

// Runs the whole build on the context stream and waits for it.
static int bvh_run(cav_ctx* ctx, cav_bvh* h, const float* tris_dev, const float* scene_bounds,
                   int morton_bits, int opt_rounds) {
    hipStream_t s = ctx->stream;
    const int n = h->n;
    Marker* mk = &h->mk;
    h->morton_bits = morton_bits;
    h->opt_rounds = opt_rounds;
    mk->start(s);
    if (scene_bounds) {
        uint32_t o6[6];
        host_to_ord(scene_bounds, o6);
        HIPCHK(ctx, hipMemcpyAsync(h->b.scene_ord, o6, sizeof o6, hipMemcpyHostToDevice, s));
        mark(mk, "bounds_upload");
    } else {
        HIPCHK(ctx, lbvh_scene_bounds(tris_dev, n, h->b.scene_ord, h->b.bounds_part, s, mk));
    }
    HIPCHK(ctx, lbvh_build(h->b, tris_dev, n, morton_bits, s, mk));
    float root_lbvh[8];
    uint32_t sort_err = 0;
    HIPCHK(ctx, hipMemcpyAsync(root_lbvh, h->b.nbox, sizeof(NodeBox), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipMemcpyAsync(&sort_err, h->b.hist + kSortHistWords + 8, 4, hipMemcpyDeviceToHost, s));
    mark(mk, "readback");
    for (int r = 0; r < opt_rounds; ++r) HIPCHK(ctx, treelet_round_impl(h->b, n, r == 0, s, mk));
    if (opt_rounds > 0) HIPCHK(ctx, lbvh_refs(h->b, n, s, mk, "flatten_refs"));
    HIPCHK(ctx, lbvh_flatten(h->b, tris_dev, h->mesh, n, opt_rounds == 0, h->rec, h->root, s, mk));
    float root_final[8];
    HIPCHK(ctx, hipMemcpyAsync(root_final, h->b.nbox, sizeof(NodeBox), hipMemcpyDeviceToHost, s));
    HIPCHK(ctx, hipStreamSynchronize(s));
    if (sort_err) return fail(ctx, CAV_ERR_HIP, "radix sort: a look-back wait exceeded its bound");
    cav_bvh_stats st{};
    st.n = n;
    st.morton_bits = morton_bits;
    st.opt_rounds = opt_rounds;
    h->nk = 0;
    for (int i = 1; i < mk->n; ++i) {
        const char* nm = mk->name[i];
        const float ms = elapsed(mk->ev[i - 1], mk->ev[i]);
        h->kname[h->nk] = nm;
        h->kms[h->nk++] = ms;
        const std::string k(nm);
        if (k.rfind("bounds", 0) == 0) st.ms_bounds += ms;
        else if (k == "morton") st.ms_morton += ms;
        else if (k.rfind("sort", 0) == 0) st.ms_sort += ms;
        else if (k == "karras") st.ms_karras += ms;
        else if (k.rfind("refit", 0) == 0 || k == "levels") st.ms_refit += ms;
        else if (k.rfind("treelet", 0) == 0) st.ms_treelet += ms;
        else if (k.rfind("flatten", 0) == 0) st.ms_flatten += ms;
    }
    st.ms_total = mk->n > 1 ? elapsed(mk->ev[0], mk->ev[mk->n - 1]) : 0.f;
    st.root_cost_lbvh = root_lbvh[7];
    st.root_cost_final = root_final[7];
    h->stats = st;
    return CAV_OK;
}

extern "C" {

int cav_ctx_create(int device, cav_ctx** out) {
    if (!out) return CAV_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return CAV_ERR_HIP;
    if (device < 0 || device >= count) return CAV_ERR_INVALID;
    e = hipSetDevice(device);
    if (e != hipSuccess) return CAV_ERR_HIP;
    // the no-alloc entry point's staging array on this device (its failure
    // only disables that entry point, which then reports it)
    (void)noalloc_init(device);
    cav_ctx* c = new cav_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return CAV_ERR_HIP; }
    *out = c;
    return CAV_OK;
}

int cav_ctx_destroy(cav_ctx* ctx) {
    if (!ctx) return CAV_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx_release(ctx);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CAV_OK;
}

const char* cav_last_error(const cav_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* cav_ctx_stream(cav_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int cav_ctx_synchronize(cav_ctx* ctx) {
    if (!ctx) return CAV_ERR_INVALID;
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    return CAV_OK;
}

int cav_bvh_build_device(cav_ctx* ctx, const float* tris_dev, int n, const float* scene_bounds,
                         const int32_t* tri_mesh_dev, int morton_bits, int opt_rounds,
                         cav_bvh** out) {
    if (!ctx || !out) return CAV_ERR_INVALID;
    *out = nullptr;
    if (!tris_dev) return fail(ctx, CAV_ERR_INVALID, "triangles is null");
    if (n < 2) return fail(ctx, CAV_ERR_INVALID, "BVH needs at least 2 triangles");
    if (n >= (1 << 30)) return fail(ctx, CAV_ERR_INVALID, "BVH takes fewer than 2^30 triangles");
    if (morton_bits != 30 && morton_bits != 63)
        return fail(ctx, CAV_ERR_INVALID, "morton_bits must be 30 or 63");
    if (opt_rounds < 0) return fail(ctx, CAV_ERR_INVALID, "opt_rounds must be >= 0");
    if ((reinterpret_cast<uintptr_t>(tris_dev) & 15u) != 0)
        return fail(ctx, CAV_ERR_INVALID, "device triangles must be 16-byte aligned");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    cav_bvh* h = new cav_bvh();
    h->ctx = ctx;
    h->n = n;
    int rc = bvh_alloc(ctx, h, n);
    if (rc == CAV_OK && tri_mesh_dev) {
        rc = dev_alloc(ctx, (void**)&h->mesh, sizeof(int32_t) * n, "mesh");
        if (rc == CAV_OK) {
            hipError_t e = hipMemcpyAsync(h->mesh, tri_mesh_dev, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, ctx->stream);
            if (e != hipSuccess) rc = hip_fail(ctx, e, "copy mesh ids");
        }
    }
    if (rc == CAV_OK) rc = bvh_run(ctx, h, tris_dev, scene_bounds, morton_bits, opt_rounds);
    if (rc != CAV_OK) { bvh_free(h); delete h; return rc; }
    *out = h;
    return CAV_OK;
}

int cav_bvh_build(cav_ctx* ctx, const float* tris, int n, const float* scene_bounds,
                  const int32_t* tri_mesh, int morton_bits, int opt_rounds, cav_bvh** out) {
    if (!ctx || !out) return CAV_ERR_INVALID;
    *out = nullptr;
    if (!tris) return fail(ctx, CAV_ERR_INVALID, "triangles is null");
    if (n < 2) return fail(ctx, CAV_ERR_INVALID, "BVH needs at least 2 triangles");
    if (n >= (1 << 30)) return fail(ctx, CAV_ERR_INVALID, "BVH takes fewer than 2^30 triangles");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    float* dt = nullptr;
    int32_t* dm = nullptr;
    int rc = dev_alloc(ctx, (void**)&dt, sizeof(float) * 9 * (size_t)n, "triangles");
    if (rc != CAV_OK) return rc;
    if (tri_mesh) rc = dev_alloc(ctx, (void**)&dm, sizeof(int32_t) * (size_t)n, "mesh");
    hipError_t e = hipSuccess;
    if (rc == CAV_OK) e = hipMemcpy(dt, tris, sizeof(float) * 9 * (size_t)n, hipMemcpyHostToDevice);
    if (rc == CAV_OK && e == hipSuccess && dm) e = hipMemcpy(dm, tri_mesh, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice);
    if (rc == CAV_OK && e != hipSuccess) rc = hip_fail(ctx, e, "upload triangles");
    if (rc == CAV_OK) rc = cav_bvh_build_device(ctx, dt, n, scene_bounds, dm, morton_bits, opt_rounds, out);
    (void)hipFree(dt);
    if (dm) (void)hipFree(dm);
    return rc;
}

int cav_bvh_rebuild_device(cav_ctx* ctx, cav_bvh* h, const float* tris_dev,
                           const float* scene_bounds, int morton_bits, int opt_rounds) {
    if (!ctx || !h || !tris_dev) return CAV_ERR_INVALID;
    if ((reinterpret_cast<uintptr_t>(tris_dev) & 15u) != 0)
        return fail(ctx, CAV_ERR_INVALID, "device triangles must be 16-byte aligned");
    if (morton_bits != 30 && morton_bits != 63)
        return fail(ctx, CAV_ERR_INVALID, "morton_bits must be 30 or 63");
    if (opt_rounds < 0) return fail(ctx, CAV_ERR_INVALID, "opt_rounds must be >= 0");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    return bvh_run(ctx, h, tris_dev, scene_bounds, morton_bits, opt_rounds);
}

int cav_bvh_destroy(cav_bvh* h) {
    if (!h) return CAV_ERR_INVALID;
    if (h->ctx) { (void)hipSetDevice(h->ctx->device); (void)hipStreamSynchronize(h->ctx->stream); }
    bvh_free(h);
    delete h;
    return CAV_OK;
}

int cav_bvh_stats_get(const cav_bvh* h, cav_bvh_stats* out) {
    if (!h || !out) return CAV_ERR_INVALID;
    *out = h->stats;
    return CAV_OK;
}

int cav_bvh_kernel_times(const cav_bvh* h, const char** names, float* ms, int max) {
    if (!h || max < 0) return -CAV_ERR_INVALID;
    const int k = h->nk < max ? h->nk : max;
    for (int i = 0; i < k; ++i) {
        if (names) names[i] = h->kname[i];
        if (ms) ms[i] = h->kms[i];
    }
    return h->nk;
}

int cav_bvh_export(cav_ctx* ctx, const cav_bvh* h, float* box6, float* sa, float* cost,
                   int32_t* left, int32_t* right, int32_t* parent, int32_t* prim,
                   uint64_t* sorted_codes, int32_t* order) {
    if (!ctx || !h) return CAV_ERR_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    const int n = h->n;
    const int64_t m = 2 * (int64_t)n - 1;
    std::vector<NodeBox> nb(m);
    std::vector<int4> tp(m);
    HIPCHK(ctx, hipMemcpy(nb.data(), h->b.nbox, sizeof(NodeBox) * m, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(tp.data(), h->b.topo, sizeof(int4) * m, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < m; ++i) {
        if (box6) {
            box6[i * 6 + 0] = nb[i].a.x; box6[i * 6 + 1] = nb[i].a.y; box6[i * 6 + 2] = nb[i].a.z;
            box6[i * 6 + 3] = nb[i].b.x; box6[i * 6 + 4] = nb[i].b.y; box6[i * 6 + 5] = nb[i].b.z;
        }
        if (sa) sa[i] = nb[i].a.w;
        if (cost) cost[i] = nb[i].b.w;
        if (left) left[i] = tp[i].x;
        if (right) right[i] = tp[i].y;
        if (parent) parent[i] = tp[i].z;
        if (prim) prim[i] = tp[i].w;
    }
    if (sorted_codes) {
        if (h->morton_bits == 63) {
            HIPCHK(ctx, hipMemcpy(sorted_codes, h->b.keys[h->b.sorted], sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
        } else {
            std::vector<uint32_t> k(n);
            HIPCHK(ctx, hipMemcpy(k.data(), h->b.keys[h->b.sorted], sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
            for (int i = 0; i < n; ++i) sorted_codes[i] = k[i];
        }
    }
    if (order) HIPCHK(ctx, hipMemcpy(order, h->b.order, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
    return CAV_OK;
}

int cav_trace_set_launch(cav_ctx* ctx, int32_t which, const cav_trace_launch* l) {
    if (!ctx || !l) return CAV_ERR_INVALID;
    if (which != CAV_TRACE_STANDALONE && which != CAV_TRACE_WAVEFRONT)
        return fail(ctx, CAV_ERR_INVALID, "which must be CAV_TRACE_STANDALONE or CAV_TRACE_WAVEFRONT");
    if (l->workgroups_per_cu < 0 || l->workgroups_per_cu > 64)
        return fail(ctx, CAV_ERR_INVALID, "workgroups_per_cu must be in 0..64");
    if (l->ray_chunk < 64 || l->ray_chunk > 65536)
        return fail(ctx, CAV_ERR_INVALID, "ray_chunk must be in 64..65536");
    if (l->lds_stack != 32 && l->lds_stack != 48) return fail(ctx, CAV_ERR_INVALID, "lds_stack must be 32 or 48");
    const auto size_ok = [](int32_t v) { return v == 0 || v == 64 || v == 128 || v == 256; };
    if (!size_ok(l->workgroup_size)) return fail(ctx, CAV_ERR_INVALID, "workgroup_size must be 0, 64, 128 or 256");
    if (!size_ok(l->shade_workgroup_size))
        return fail(ctx, CAV_ERR_INVALID, "shade_workgroup_size must be 0, 64, 128 or 256");
    if (l->handoff_steps < 0 || l->handoff_steps > (1 << 30))
        return fail(ctx, CAV_ERR_INVALID, "handoff_steps must be in 0..2^30");
    if (l->any_hit_step < 0 || l->any_hit_step > CAV_ANY_STEP_PAIR)
        return fail(ctx, CAV_ERR_INVALID, "any_hit_step must be 0, CAV_ANY_STEP_NODE or CAV_ANY_STEP_PAIR");
    if (l->any_hit_whole_line != 0 && l->any_hit_whole_line != 1)
        return fail(ctx, CAV_ERR_INVALID, "any_hit_whole_line must be 0 or 1");
    ctx->launch[which] = TraceLaunch{l->workgroups_per_cu, l->ray_chunk, l->lds_stack,
                                     l->workgroup_size ? l->workgroup_size : 256,
                                     l->shade_workgroup_size ? l->shade_workgroup_size : 256,
                                     l->handoff_steps ? l->handoff_steps : 1024, l->any_hit_step,
                                     l->any_hit_whole_line};
    return CAV_OK;
}

int cav_trace_get_launch(cav_ctx* ctx, int32_t which, cav_trace_launch* out) {
    if (!ctx || !out) return CAV_ERR_INVALID;
    if (which != CAV_TRACE_STANDALONE && which != CAV_TRACE_WAVEFRONT)
        return fail(ctx, CAV_ERR_INVALID, "which must be CAV_TRACE_STANDALONE or CAV_TRACE_WAVEFRONT");
    const TraceLaunch& l = ctx->launch[which];
    *out = cav_trace_launch{l.per_cu, l.chunk, l.lds_depth, l.block, l.shade_block, l.handoff, l.any_step,
                            l.any_full};
    return CAV_OK;
}

// Grows the BVH's long-ray pass buffers (any-hit spill records and
// closest-hit ray lists) to hold a batch of nrays (the stream is drained
// before old buffers are freed); *coop receives them.
static int ensure_spill(cav_ctx* ctx, const cav_bvh* h, int64_t nrays, TraceCoop* coop) {
    const uint32_t need = any_spill_capacity(nrays);
    if (h->spill_cap < need) {
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
        if (h->spill) (void)hipFree(h->spill);
        h->spill = nullptr;
        h->spill_cap = 0;
        void* p = nullptr;
        const int rc = dev_alloc(ctx, &p, trace_coop_bytes(need, need), "long-ray pass buffers");
        if (rc != CAV_OK) return rc;
        h->spill = p;
        h->spill_cap = need;
    }
    *coop = trace_coop_in(h->spill, h->spill_cap, h->spill_cap);
    return CAV_OK;
}

int cav_trace_closest_async(cav_ctx* ctx, const cav_bvh* h, const float* rays, int64_t nrays,
                            int32_t* prim, float* t) {
    if (!ctx || !h || (nrays > 0 && (!rays || !prim || !t))) return CAV_ERR_INVALID;
    TraceCoop coop;
    const int rc = ensure_spill(ctx, h, nrays, &coop);
    if (rc != CAV_OK) return rc;
    HIPCHK(ctx, trace_closest(h->rec, h->n, h->root, rays, nrays, prim, t, nullptr, nullptr,
                              h->b.counters, ctx->launch[CAV_TRACE_STANDALONE], ctx->stream, &coop));
    return CAV_OK;
}

int cav_trace_occluded_async(cav_ctx* ctx, const cav_bvh* h, const float* rays,
                             const float* limit, int64_t nrays, uint8_t* visible) {
    if (!ctx || !h || (nrays > 0 && (!rays || !limit || !visible))) return CAV_ERR_INVALID;
    TraceCoop coop;
    const int rc = ensure_spill(ctx, h, nrays, &coop);
    if (rc != CAV_OK) return rc;
    HIPCHK(ctx, trace_occluded(h->rec, h->n, rays, limit, nullptr, nullptr, nrays, visible,
                               h->b.counters, ctx->launch[CAV_TRACE_STANDALONE], ctx->stream, &coop));
    return CAV_OK;
}

int cav_trace_counters(cav_ctx* ctx, const cav_bvh* h, uint32_t* out4, int32_t reset) {
    if (!ctx || !h || !out4) return CAV_ERR_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    uint32_t w[8];
    HIPCHK(ctx, hipMemcpyAsync(w, h->b.counters, sizeof w, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    out4[0] = w[0];
    out4[1] = w[1];
    out4[2] = w[6];
    out4[3] = w[7];
    if (reset) {
        HIPCHK(ctx, hipMemsetAsync(h->b.counters, 0, 8, ctx->stream));
        HIPCHK(ctx, hipMemsetAsync(h->b.counters + 6, 0, 8, ctx->stream));
        HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return CAV_OK;
}

int cav_trace_fetches(cav_ctx* ctx, const cav_bvh* h, uint64_t* out6, int32_t reset) {
    if (!ctx || !h || !out6) return CAV_ERR_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    uint32_t w[11] = {};  // words 20..30: the long-ray passes of the last launches
    HIPCHK(ctx, hipMemcpyAsync(out6, h->b.counters + 16, 16, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(w, h->b.counters + 20, sizeof w, hipMemcpyDeviceToHost, ctx->stream));
    if (reset) HIPCHK(ctx, hipMemsetAsync(h->b.counters + 16, 0, 16, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    out6[2] = w[0] < h->spill_cap ? w[0] : h->spill_cap;  // any-hit rays spilled
    out6[3] = w[2];                                         // of those, traced again on one lane
    out6[4] = w[6] < h->spill_cap ? w[6] : h->spill_cap;  // closest-hit rays handed over
    out6[5] = w[10];                                        // of those, traced again on one lane
    return CAV_OK;
}

int cav_bvh_max_depth(cav_ctx* ctx, const cav_bvh* h, int32_t* depth) {
    if (!ctx || !h || !depth) return CAV_ERR_INVALID;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    // counters[8] is free scratch (the traversal uses words 0..7 and 16..30)
    uint32_t* word = h->b.counters + 8;
    HIPCHK(ctx, levels_max_depth(h->b, h->n, word, ctx->stream));
    uint32_t d = 0;
    HIPCHK(ctx, hipMemcpyAsync(&d, word, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    *depth = (int32_t)d;
    return CAV_OK;
}

int cav_trace_closest_device(cav_ctx* ctx, const cav_bvh* h, const float* rays, int64_t nrays,
                             int32_t* prim, float* t, float* u, float* v) {
    if (!ctx || !h) return CAV_ERR_INVALID;
    if (nrays < 0 || (nrays > 0 && (!rays || !prim || !t)))
        return fail(ctx, CAV_ERR_INVALID, "null ray or output buffer");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemsetAsync(h->b.counters, 0, 4, ctx->stream));
    TraceCoop coop;
    const int rc = ensure_spill(ctx, h, nrays, &coop);
    if (rc != CAV_OK) return rc;
    HIPCHK(ctx, trace_closest(h->rec, h->n, h->root, rays, nrays, prim, t, u, v, h->b.counters,
                              ctx->launch[CAV_TRACE_STANDALONE], ctx->stream, &coop));
    uint32_t ovf = 0;
    HIPCHK(ctx, hipMemcpyAsync(&ovf, h->b.counters, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ovf) return fail(ctx, CAV_ERR_OVERFLOW, "traversal stack overflow in " + std::to_string(ovf) + " pushes");
    return CAV_OK;
}

int cav_trace_occluded_device(cav_ctx* ctx, const cav_bvh* h, const float* rays,
                              const float* limit, const int32_t* src, const int32_t* dst,
                              int64_t nrays, uint8_t* visible) {
    if (!ctx || !h) return CAV_ERR_INVALID;
    if (nrays < 0 || (nrays > 0 && (!rays || !limit || !visible)))
        return fail(ctx, CAV_ERR_INVALID, "null ray, limit or output buffer");
    HIPCHK(ctx, hipSetDevice(ctx->device));
    HIPCHK(ctx, hipMemsetAsync(h->b.counters + 1, 0, 4, ctx->stream));
    TraceCoop coop;
    const int rc = ensure_spill(ctx, h, nrays, &coop);
    if (rc != CAV_OK) return rc;
    HIPCHK(ctx, trace_occluded(h->rec, h->n, rays, limit, src, dst, nrays, visible, h->b.counters,
                               ctx->launch[CAV_TRACE_STANDALONE], ctx->stream, &coop));
    uint32_t ovf = 0;
    HIPCHK(ctx, hipMemcpyAsync(&ovf, h->b.counters + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    if (ovf) return fail(ctx, CAV_ERR_OVERFLOW, "traversal stack overflow in " + std::to_string(ovf) + " pushes");
    return CAV_OK;
}

int cav_trace_closest(cav_ctx* ctx, const cav_bvh* h, const float* rays, int64_t nrays,
                      int32_t* prim, float* t, float* u, float* v) {
    if (!ctx || !h) return CAV_ERR_INVALID;
    if (nrays < 0 || (nrays > 0 && (!rays || !prim || !t)))
        return fail(ctx, CAV_ERR_INVALID, "null ray or output buffer");
    if (nrays == 0) return CAV_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t r = (size_t)nrays;
    void *dr = nullptr, *dp = nullptr, *dt = nullptr, *du = nullptr, *dv = nullptr;
    int rc = dev_alloc(ctx, &dr, 24 * r, "rays");
    if (rc == CAV_OK) rc = dev_alloc(ctx, &dp, 4 * r, "hit prim");
    if (rc == CAV_OK) rc = dev_alloc(ctx, &dt, 4 * r, "hit t");
    if (rc == CAV_OK && u) rc = dev_alloc(ctx, &du, 4 * r, "hit u");
    if (rc == CAV_OK && v) rc = dev_alloc(ctx, &dv, 4 * r, "hit v");
    hipError_t e = hipSuccess;
    if (rc == CAV_OK) e = hipMemcpy(dr, rays, 24 * r, hipMemcpyHostToDevice);
    if (rc == CAV_OK && e != hipSuccess) rc = hip_fail(ctx, e, "upload rays");
    int trc = CAV_OK;
    if (rc == CAV_OK) {
        trc = cav_trace_closest_device(ctx, h, (const float*)dr, nrays, (int32_t*)dp, (float*)dt, (float*)du, (float*)dv);
        if (trc != CAV_OK && trc != CAV_ERR_OVERFLOW) rc = trc;
    }
    if (rc == CAV_OK) {
        e = hipMemcpy(prim, dp, 4 * r, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipMemcpy(t, dt, 4 * r, hipMemcpyDeviceToHost);
        if (e == hipSuccess && u) e = hipMemcpy(u, du, 4 * r, hipMemcpyDeviceToHost);
        if (e == hipSuccess && v) e = hipMemcpy(v, dv, 4 * r, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "download hits");
    }
    for (void* p : {dr, dp, dt, du, dv}) if (p) (void)hipFree(p);
    return rc != CAV_OK ? rc : trc;
}

int cav_trace_occluded(cav_ctx* ctx, const cav_bvh* h, const float* rays, const float* limit,
                       const int32_t* src, const int32_t* dst, int64_t nrays, uint8_t* visible) {
    if (!ctx || !h) return CAV_ERR_INVALID;
    if (nrays < 0 || (nrays > 0 && (!rays || !limit || !visible)))
        return fail(ctx, CAV_ERR_INVALID, "null ray, limit or output buffer");
    if (nrays == 0) return CAV_OK;
    HIPCHK(ctx, hipSetDevice(ctx->device));
    const size_t r = (size_t)nrays;
    void *dr = nullptr, *dl = nullptr, *ds = nullptr, *dd = nullptr, *dvis = nullptr;
    int rc = dev_alloc(ctx, &dr, 24 * r, "rays");
    if (rc == CAV_OK) rc = dev_alloc(ctx, &dl, 4 * r, "limits");
    if (rc == CAV_OK && src) rc = dev_alloc(ctx, &ds, 4 * r, "src mesh");
    if (rc == CAV_OK && dst) rc = dev_alloc(ctx, &dd, 4 * r, "dst mesh");
    if (rc == CAV_OK) rc = dev_alloc(ctx, &dvis, r, "visible");
    hipError_t e = hipSuccess;
    if (rc == CAV_OK) e = hipMemcpy(dr, rays, 24 * r, hipMemcpyHostToDevice);
    if (rc == CAV_OK && e == hipSuccess) e = hipMemcpy(dl, limit, 4 * r, hipMemcpyHostToDevice);
    if (rc == CAV_OK && e == hipSuccess && src) e = hipMemcpy(ds, src, 4 * r, hipMemcpyHostToDevice);
    if (rc == CAV_OK && e == hipSuccess && dst) e = hipMemcpy(dd, dst, 4 * r, hipMemcpyHostToDevice);
    if (rc == CAV_OK && e != hipSuccess) rc = hip_fail(ctx, e, "upload rays");
    int trc = CAV_OK;
    if (rc == CAV_OK) {
        trc = cav_trace_occluded_device(ctx, h, (const float*)dr, (const float*)dl, (const int32_t*)ds,
                                        (const int32_t*)dd, nrays, (uint8_t*)dvis);
        if (trc != CAV_OK && trc != CAV_ERR_OVERFLOW) rc = trc;
    }
    if (rc == CAV_OK) {
        e = hipMemcpy(visible, dvis, r, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(ctx, e, "download visibility");
    }
    for (void* p : {dr, dl, ds, dd, dvis}) if (p) (void)hipFree(p);
    return rc != CAV_OK ? rc : trc;
}

}  // extern "C"
