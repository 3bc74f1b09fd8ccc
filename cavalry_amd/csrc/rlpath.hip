// Wavefront `rlpath` integrator: path tracing guided by a Q-table of
// incident radiance (expected-SARSA, Dahm & Keller) and a learnt per-cell
// light-selection distribution.
//
// Reference: Integrators/ReinforcementLearningPathTracing.cu (QEntry :10-148,
// findQEntry :155-165, intersectScene :276-302, genNextRay :305-375,
// computeLighting :388-460, finishLightingTraining :463-514,
// computeQDistributions :517-606, accumulateLightingContribs :609-647,
// syncDimension :678-712, Q-table kernels :715-767, render :802-950) and
// Distribution1D (Utils/RandomUtils.h:120-183).
//
// Learning couples every sample of a bounce: the Q proposals and light
// statistics of all tasks are folded into the tables before any task picks
// its next direction. One bounce is
//   closest hits -> k_dim_max (the reference's first syncDimension) ->
//   k_lighting (emission with MIS, learnt light pick, shadow task) ->
//   any-hit traversal -> k_qdist (NEE contribution, 128-bin Q distribution,
//   Q and light proposals, next-bin sample, MIS with the Q pdf) ->
//   k_finish_q (commit + average per entry) -> k_finish_light (light CDFs) ->
//   k_next (second syncDimension, Russian roulette, next ray).
// Halton dimensions follow the reference, including the batch-wide maximum
// taken by syncDimension, so the device keeps the maximum in a word instead
// of the host reading it back; the light sample's share is the scene's
// light_dims (see SceneDev). Tasks keep their queue slot through a
// bounce, so no lighting queue is built: per-slot arrays carry the results
// from one kernel to the next.
//
// Proposals are float atomics, as in the reference; the order in which they
// land varies between runs, so films agree with the CPU restatement exactly
// only at max depth 1 and statistically beyond.
#include <cfloat>

#include "../../include/cavalry.h"
#include "dev_common.h"
#include "radix_sort.h"
#include "render.h"
#include "serve.h"
#include "shade.h"

namespace cav {
namespace {

using namespace shade;

constexpr int kNX = 16, kNY = 8, kXY = kNX * kNY;  // QEntry::NUM_X / NUM_Y / NUM_XY
constexpr int kEntryWaves = 4;                      // entries per k_finish_q block

// float -> int as the reference's conversion behaves: NaN gives 0, out of
// range saturates.
__device__ inline int f2i(float v) {
    if (isnan(v)) return 0;
    if (v >= 2147483648.f) return 2147483647;
    if (v <= -2147483648.f) return (-2147483647 - 1);
    return (int)v;
}
__device__ inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// Correctly rounded single-precision acos / cos / sin via binary64.
__device__ inline float acos_f(float x) { return (float)acos((double)x); }
__device__ inline float cos_f(float x) { return (float)cos((double)x); }
__device__ inline float sin_f(float x) { return (float)sin((double)x); }
// both at once (one argument reduction); each rounded once to float
__device__ inline void sincos_f(float x, float& s, float& c) {
    double sd, cd;
    sincos((double)x, &sd, &cd);
    s = (float)sd;
    c = (float)cd;
}

__device__ inline float luminance(v3 c) {
    return (c.x * 0.212671f + c.y * 0.715160f) + c.z * 0.072169f;
}

// QEntry::dirToCellIndex (:113-128)
__device__ inline int dir_to_cell(v3 d) {
    const float u = d.z;
    const float y = (u + 1.f) / 2.f;
    const int ti = clampi(f2i(y * kNY), 0, kNY - 1);
    d.x = d.x / sqrtf(1.f - u * u);
    d.y = d.y / sqrtf(1.f - u * u);
    float v = acos_f(d.x);
    if (d.y < 0) v = (float)((2.0 * kPi) - (double)v);
    v = (float)((double)v / (2.0 * kPi));
    const int pi = clampi(f2i(v * kNX), 0, kNX - 1);
    return ti * kNX + pi;
}

// findQEntry (:155-165)
__device__ inline int q_entry(const Grid& gr, v3 p) {
    const float gf = (float)gr.g;
    const float ex = gr.size[0] / gf, ey = gr.size[1] / gf, ez = gr.size[2] / gf;
    const int x = clampi(f2i((p.x - gr.lo[0]) / ex), 0, gr.g - 1);
    const int y = clampi(f2i((p.y - gr.lo[1]) / ey), 0, gr.g - 1);
    const int z = clampi(f2i((p.z - gr.lo[2]) / ez), 0, gr.g - 1);
    return (x * gr.g + y) * gr.g + z;
}

// Distribution1D::getCorrespondingIndex (Utils/RandomUtils.h:130-150)
__device__ inline int dist_index(const float* cdf, int n, float f) {
    int l = 0, r = n - 1;
    while (l < r) {
        const int m = (l + r) / 2;
        if (cdf[m] >= f) {
            if (m == 0) return 0;
            if (cdf[m - 1] < f) return m;
            r = m - 1;
        } else {
            l = m + 1;
        }
    }
    return l;
}

// -------------------------------------------------------------- state
// Where a task's request state comes from. RenderSrc: a render, one request
// for every task. PoolSrc: a server pool, the task's own request slot; tasks
// of other integrators share the queues and are skipped.
struct RenderSrc {
    RlReqState s;
    int depth_;
    __device__ bool take(int64_t) const { return true; }
    __device__ const RlReqState& get(int64_t) const { return s; }
    // the state a task reads; the same for every lane, so it stays in the
    // kernel argument (selecting between it and another address would make
    // the compiler copy the argument to scratch memory in every thread)
    __device__ const RlReqState& state(int64_t, bool) const { return s; }
    __device__ int depth(int64_t) const { return depth_; }
    __device__ int slot(int64_t) const { return 0; }
    __device__ void tag_shadow(uint32_t, int64_t) const {}
};
struct PoolSrc {
    const RlReqState* rl;  // per request slot
    const int32_t* rid;    // request slot of each task of the input queue
    const uint8_t* dep;    // bounce index of each task
    int32_t* sh_rid;       // request slot of each shadow task
    uint8_t* live;         // per request slot: it had a hit this bounce
    __device__ bool take(int64_t i) const { return rl[rid[i]].t.Q != nullptr; }
    __device__ const RlReqState& get(int64_t i) const { return rl[rid[i]]; }
    // the task's request slot (slot 0 for a lane without a task: never read)
    __device__ const RlReqState& state(int64_t i, bool live) const { return rl[live ? rid[i] : 0]; }
    __device__ int depth(int64_t i) const { return dep[i]; }
    __device__ int slot(int64_t i) const { return rid[i]; }
    __device__ void tag_shadow(uint32_t ss, int64_t i) const { sh_rid[ss] = rid[i]; }
};
__global__ __launch_bounds__(kBlock) void k_init_tables(RlTables t, int64_t ncells, int64_t nent, int L) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < ncells) {  // QEntry() (:36-41) and prepareForUpdateQ
        t.Q[i] = 1.f;
        t.newQ[i] = 0.f;
        t.cnt[i] = 0;
        t.total[i] = 0;
    }
    if (i < nent * L) {  // initialiseLightSamplingDist (:724-735)
        const int k = (int)(i % L);
        t.ldist[i] = ((float)k + 1.f) / (float)L;
        t.lres[i] = 0.f;
        t.lnew[i] = 0.f;
        t.lcount[i] = 0;
        t.lnewc[i] = 0;
    }
}

// genInitialRays' QEntryInfo {-1, -1} and the camera sample's two dimensions
__global__ __launch_bounds__(kBlock) void k_init_tasks(int64_t nsamp, RlQueue q) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nsamp) return;
    q.prev_entry[i] = -1;
    q.prev_cell[i] = -1;
    q.dim[i] = 2;
}

__device__ inline int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// syncDimension before computeLighting: maximum over the tasks that hit, one
// atomic per 1024-thread block at most (almost every block finds the value
// already stored and skips it).
constexpr int kMaxBlock = 1024;
__global__ __launch_bounds__(kMaxBlock) void k_dim_max(const uint32_t* counts, int qin, RlQueue q,
                                                       const int32_t* __restrict__ hit_prim, int32_t* dims) {
    __shared__ int wmax[kMaxBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kMaxBlock + threadIdx.x;
    const bool live = i < (int64_t)counts[qin] && hit_prim[i] >= 0;
    const int m = wave_max(live ? q.dim[i] : 0);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = 0;
        for (int w = 0; w < kMaxBlock / 64; ++w) b = max(b, wmax[w]);
        if (b > __hip_atomic_load(&dims[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&dims[0], b);
    }
}

// The same per pool task into its own request's word; marks the request as
// having hits this bounce.
__global__ __launch_bounds__(kBlock) void k_pool_dim_max(PoolSrc src, const uint32_t* counts, int qin, RlQueue q,
                                                         const int32_t* __restrict__ hit_prim) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)counts[qin] || hit_prim[i] < 0 || !src.take(i)) return;
    int32_t* d = src.get(i).dims;
    const int v = q.dim[i];
    if (v > __hip_atomic_load(&d[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(&d[0], v);
    const int r = src.rid[i];
    if (!src.live[r]) src.live[r] = 1;
}

// computeLighting (:388-460) up to the visibility test: emission with MIS,
// learnt light choice, point on the light. The shadow task carries the slot.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_lighting(Src src, SceneDev sc, HaltonDev h,
                                                     Queue qi, RlSlots sl, Shadow sh,
                                                     const int32_t* __restrict__ hit_prim,
                                                     const float* __restrict__ hit_t,
                                                     const float* __restrict__ hit_u,
                                                     const float* __restrict__ hit_v,
                                                     uint32_t* counts, int qin,
                                                     const uint32_t* __restrict__ perm) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    bool live = j < (int64_t)counts[qin];
    const int64_t i = live ? task_slot(perm, j) : 0;
    live = live && src.take(i);
    const int prim = live ? hit_prim[i] : -1;
    const RlReqState& R = src.state(i, live);
    if (live && prim < 0 && sc.env_miss >= 0) {  // intersectScene's miss (:292-296)
        float* acc = R.res + (int64_t)qi.sid[i] * 3;
        const v3 e = mul3(miss_radiance(sc, mk3(qi.ray[i * 6 + 3], qi.ray[i * 6 + 4], qi.ray[i * 6 + 5])),
                          mk3(qi.mult[i * 3], qi.mult[i * 3 + 1], qi.mult[i * 3 + 2]));
        acc[0] = acc[0] + e.x; acc[1] = acc[1] + e.y; acc[2] = acc[2] + e.z;
    }
    live = live && prim >= 0;
    v3 pos = mk3(0, 0, 0), dir = pos;
    float lim = 0.f;
    int mesh = 0, dst = 0;
    if (live) {
        const int D = R.dims[0];
        const int depth = src.depth(i);
        float* res = R.res;
        const uint32_t sid = qi.sid[i];
        const float* r = qi.ray + i * 6;
        const v3 ro = mk3(r[0], r[1], r[2]), rd = mk3(r[3], r[4], r[5]);
        const v3 mult = mk3(qi.mult[i * 3], qi.mult[i * 3 + 1], qi.mult[i * 3 + 2]);
        mesh = sc.tri_mesh[prim];
        const cav_mesh M = sc.meshes[mesh];
        const v3 nn = shading_normal(sc, prim, hit_u ? hit_u[i] : 0.f, hit_u ? hit_v[i] : 0.f);
        pos = add3(ro, scale3(rd, hit_t[i]));
        float emis = 0.f;
        if (M.is_light) {
            const v3 Le = dot3(rd, nn) >= 0 ? mk3(0, 0, 0) : mk3(M.emission[0], M.emission[1], M.emission[2]);
            v3 acc = mk3(res[sid * 3], res[sid * 3 + 1], res[sid * 3 + 2]);
            if (qi.delta[i]) {
                acc = add3(acc, mul3(Le, mult));
            } else if (depth > 0) {
                const float lpdf = light_pdf(sc, mesh, ro, pos, nn);
                const float w = power_heuristic(qi.pdf[i], lpdf);
                if (isfinite(w)) {
                    emis = luminance(Le);
                    acc = add3(acc, scale3(mul3(Le, mult), w));
                }
            }
            res[sid * 3] = acc.x; res[sid * 3 + 1] = acc.y; res[sid * 3 + 2] = acc.z;
        }
        const int entry = q_entry(R.gr, pos);
        const float* cdf = R.t.ldist + (int64_t)entry * sc.nlights;
        const int li = dist_index(cdf, sc.nlights, halton(h, sid, D));
        const float sel = li == 0 ? cdf[0] : cdf[li] - cdf[li - 1];
        const LightPick L = sample_light(sc, h, sid, D + 5, pos, sc.lights[li]);
        const float lpdf = L.prob;
        const v3 inc = L.incident;
        dir = L.dir;
        lim = L.lim;
        dst = L.dst;
        sl.entry[i] = entry;
        sl.light[i] = li;
        sl.sel[i] = sel;
        sl.lpdf[i] = lpdf;
        sl.emis[i] = emis;
        sl.ldir[i * 3] = dir.x; sl.ldir[i * 3 + 1] = dir.y; sl.ldir[i * 3 + 2] = dir.z;
        sl.inc[i * 3] = inc.x; sl.inc[i * 3 + 1] = inc.y; sl.inc[i * 3 + 2] = inc.z;
    }
    const uint32_t ss = block_push<kBlock>(live, &counts[2]);
    if (live) {
        float* r = sh.ray + (int64_t)ss * 6;
        r[0] = pos.x; r[1] = pos.y; r[2] = pos.z; r[3] = dir.x; r[4] = dir.y; r[5] = dir.z;
        sh.limit[ss] = lim; sh.src[ss] = mesh; sh.dst[ss] = dst;
        sl.sh_idx[i] = (int32_t)ss;
        src.tag_shadow(ss, i);
    }
}

// Unit direction inside bin `cell` at offset (rx, ry):
// QEntry::sampleDirectionInCell (:131-147).
__device__ inline v3 cell_dir(float rx, float ry, int cell) {
    const int ti = cell / kNX, pi = cell % kNX;
    float u = ((float)ti + rx) * (1.f / (float)kNY);
    u = u * 2 - 1.f;
    const float v = ((float)pi + ry) * (1.f / (float)kNX);
    const float xy = sqrtf(1.0f - u * u);
    const float phi = (float)(2 * kPi * (double)v);
    return mk3(xy * cos_f(phi), xy * sin_f(phi), u);
}

// Bins of the Q-weighted distribution: qv = |cos| * luminance(f) * Q for
// the bin directions at offset (px, py). Each thread keeps the cos / sin of
// its 16 column angles in LDS (`cs`, stride B: the workgroup size) so the
// bin loops need no register arrays.
template <int B>
struct BinEval {
    Bsdf bsdf;
    v3 t0, t1, nn, ol;
    const float* Qe;
    const float* cs;
    float px;
    bool matte;   // Lambertian: luminance(f) is one constant on the exitant side
    float lum_m;  // luminance(kd / pi)

    __device__ inline void row(int ti, float& u, float& xy) const {
        u = ((float)ti + px) * (1.f / (float)kNY);
        u = u * 2 - 1.f;
        xy = sqrtf(1.0f - u * u);
    }
    // bin p of a row: |cos| * luminance(f) * q, q = the bin's Q value
    __device__ inline float bin(float u, float xy, int p, float q) const {
        const v3 d = mk3(xy * cs[p * B], xy * cs[(kNX + p) * B], u);
        if (matte) {  // the same bits as the general case: the local z is dot3(nn, d)
            const float z = dot3(nn, d);
            return (fabsf(z) * (z * ol.z > 0 ? lum_m : 0.f)) * q;
        }
        const float scat = luminance(bsdf.eval(to_local(d, t0, t1, nn), ol));
        return (fabsf(dot3(d, nn)) * scat) * q;
    }
    // the Q values of bins 4 g .. 4 g + 3 of row ti (one 16-byte load)
    __device__ inline float4 q4(int ti, int g) const {
        return reinterpret_cast<const float4*>(Qe)[ti * (kNX / 4) + g];
    }
    // prefix[m] and prefix[m - 1] (0 for m = 0), summed bin by bin
    __device__ inline void prefix(int m, float& pm, float& pm1) const {
        float s = 0.f;
        pm1 = 0.f;
        for (int c = 0; c <= m; ++c) {
            pm1 = s;
            float u, xy;
            row(c / kNX, u, xy);
            s += bin(u, xy, c % kNX, Qe[c]);
        }
        pm = s;
    }
};

// For every lane with `on`: fsum[k] += v and cnt[k] += 1, as atomics. Lanes of
// the wave that share k are summed first, so one atomic pair per distinct k
// reaches L2 (neighbouring paths share Q entries, cells and lights, and
// same-address atomics serialise there). Each group costs a dependent chain
// of cross-lane reductions, so once the leader's k is a singleton (a wave of
// incoherent paths) the remaining lanes add one by one instead (rlpath at
// 320x180x256 spp on 2M triangles: 265 -> 256 ms). Float sums are
// order-dependent either way, as in the reference. Call with the whole wave
// converged.
// fsum / cnt point at this lane's words; lanes with the same words (the same
// table and index) are summed first.
__device__ inline void wave_add(float* fsum, int* cnt, float v, bool on) {
    const int lane = threadIdx.x & 63;
    uint64_t todo = __ballot(on);
    const unsigned long long key = reinterpret_cast<unsigned long long>(fsum);
    while (todo) {
        const int lead = __ffsll((unsigned long long)todo) - 1;
        const unsigned long long lk = __shfl(key, lead, 64);
        const bool mine = ((todo >> lane) & 1ull) && key == lk;
        const uint64_t grp = __ballot(mine);
        if (grp == (1ull << lead)) break;
        float sv = mine ? v : 0.f;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) sv += __shfl_xor(sv, o, 64);
        if (lane == lead) {
            atomicAdd(fsum, sv);
            atomicAdd(cnt, (int)__popcll(grp));
        }
        todo &= ~grp;
    }
    if ((todo >> lane) & 1ull) {
        atomicAdd(fsum, v);
        atomicAdd(cnt, 1);
    }
}

// Deterministic mode: the table updates of a bounce as records. Key bits:
// 58 = light statistics (else Q), 33..57 = the table index, 1..32 = the
// sample, 0 = the kind (0: the previous bin's proposal, 1: the direct-light
// proposal), the order in which the oracle, which runs the tasks one by one
// in sample order, adds them.
constexpr int kRecKeyShift = 33;
constexpr int kRecKeyBits = 25;

// bits 59..63: the request's index (rlx; 0 in a render, its slot in a pool)
constexpr int kRecReqShift = 59;
constexpr int kRecMaxReqs = 32;
__device__ inline uint64_t rec_key(uint32_t rlx, bool light, int64_t k, uint32_t sid, int kind) {
    return ((uint64_t)rlx << kRecReqShift) | (light ? (1ull << 58) : 0ull) | ((uint64_t)k << kRecKeyShift) |
           ((uint64_t)sid << 1) | (uint64_t)kind;
}
// append (key, v) for every lane with `on`; call with the whole wave converged
__device__ inline void rec_push(const RlRecords& r, uint64_t key, float v, bool on) {
    const uint64_t m = __ballot(on);
    if (m == 0) return;
    const int lane = threadIdx.x & 63;
    const int first = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == first) base = atomicAdd(r.count, (uint32_t)__popcll(m));
    base = __shfl(base, first, 64);
    if (on) {
        const uint32_t slot = base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        r.keys[slot] = key;
        r.val[slot] = v;
        r.idx[slot] = slot;
    }
}

// The end of computeLighting (after the visibility test), then
// computeQDistributions (:517-606) and accumulateLightingContribs (:609-647).
template <class Src, int B>
__global__ __launch_bounds__(B) void k_qdist(Src src, SceneDev sc, HaltonDev h, Queue qi, RlQueue rq,
                                                  RlSlots sl, const uint8_t* __restrict__ vis,
                                                  const int32_t* __restrict__ hit_prim,
                                                  const float* __restrict__ hit_t,
                                                  const float* __restrict__ hit_u,
                                                  const float* __restrict__ hit_v,
                                                  const uint32_t* counts, int qin,
                                                  const uint32_t* __restrict__ perm, RlRecords rec) {
    __shared__ float trig[2 * kNX][B];
    const int64_t j = (int64_t)blockIdx.x * B + threadIdx.x;
    bool live = j < (int64_t)counts[qin];
    const int64_t i = live ? task_slot(perm, j) : 0;
    live = live && src.take(i);
    const int prim = live ? hit_prim[i] : -1;
    live = live && prim >= 0;
    const RlReqState& R = src.state(i, live);
    float *up_fl = nullptr, *up_fp = nullptr, *up_fd = nullptr;  // the words the updates go to (atomics)
    int32_t *up_cl = nullptr, *up_cp = nullptr, *up_cd = nullptr;
    uint32_t rlx = 0;
    bool use_rec = false;  // this task's request sums its updates from records
    int32_t* dims = nullptr;
    bool valid = false;
    // the table updates of this task, made after the divergent part
    bool up_l = false, up_p = false, up_d = false;
    int64_t k_l = 0, k_p = 0, k_d = 0;
    float v_l = 0.f, v_p = 0.f, v_d = 0.f;
    uint32_t sid_up = 0;
    if (live) {
        const RlTables& t = R.t;
        float* res = R.res;
        dims = R.dims;
        rlx = R.rlx;
        use_rec = R.det != 0;
        const int D = dims[0] + 5 + sc.light_dims;
        const uint32_t sid = qi.sid[i];
        sid_up = sid;
        const float* r = qi.ray + i * 6;
        const v3 ro = mk3(r[0], r[1], r[2]), rd = mk3(r[3], r[4], r[5]);
        const v3 mult = mk3(qi.mult[i * 3], qi.mult[i * 3 + 1], qi.mult[i * 3 + 2]);
        const int mesh = sc.tri_mesh[prim];
        const float hu = hit_u ? hit_u[i] : 0.f, hv = hit_u ? hit_v[i] : 0.f;
        const Bsdf bsdf = surface_bsdf(sc, mesh, prim, hu, hv);
        const v3 nn = shading_normal(sc, prim, hu, hv);
        v3 t0, t1;
        frame(nn, t0, t1);
        const v3 ol = to_local(scale3(rd, -1.f), t0, t1, nn);
        const int entry = sl.entry[i];
        const int li = sl.light[i];
        const v3 dir = mk3(sl.ldir[i * 3], sl.ldir[i * 3 + 1], sl.ldir[i * 3 + 2]);
        v3 inc = mk3(sl.inc[i * 3], sl.inc[i * 3 + 1], sl.inc[i * 3 + 2]);
        float lpdf = sl.lpdf[i];
        if (!(vis[sl.sh_idx[i]] && isfinite(lpdf))) {
            inc = mk3(0, 0, 0);
            lpdf = 1;
        }
        const v3 il = to_local(dir, t0, t1, nn);
        // Material::eval (Core/Material.h:20-26)
        v3 contrib = scale3(mul3(inc, bsdf.eval(il, ol)), fabsf(dot3(dir, nn)));
        contrib = scale3(contrib, 1.f / (lpdf * sl.sel[i]));
        const float sampled = luminance(inc);
        up_l = true;
        k_l = (int64_t)entry * sc.nlights + li;
        v_l = sampled;
        up_fl = t.lnew + k_l;
        up_cl = t.lnewc + k_l;

        // 128-bin distribution of |cos| * luminance(f) * Q at one shared offset.
        // Bin c = 16 * theta_row + phi_column. The running (unnormalised) sum
        // is kept only at the end of each row; the row holding the sampled
        // bin is evaluated again, which reproduces the same partial sums.
        const float* Qe = t.Q + (int64_t)entry * kXY;
        const float px = halton(h, sid, D), py = halton(h, sid, D + 1);
        const int lc = dir_to_cell(dir);
        float* cs = &trig[0][threadIdx.x];
        for (int p = 0; p < kNX; ++p) {
            const float v = ((float)p + py) * (1.f / (float)kNX);
            const float phi = (float)(2 * kPi * (double)v);
            float sp, cp;
            sincos_f(phi, sp, cp);
            cs[p * B] = cp;
            cs[(kNX + p) * B] = sp;
        }
        const bool matte = bsdf.type == CAV_MAT_MATTE;
        const BinEval<B> be{bsdf, t0, t1, nn, ol, Qe, cs, px, matte,
                         matte ? luminance(divs3(bsdf.kd, (float)kPi)) : 0.f};
        float rows[kNY];
        float sum = 0.f, lc0 = 0.f, lc1 = 0.f;
        bool mono = true;
        for (int ti = 0; ti < kNY; ++ti) {
            float u, xy;
            be.row(ti, u, xy);
            for (int g = 0; g < kNX / 4; ++g) {
                const float4 qg = be.q4(ti, g);
                const float qs[4] = {qg.x, qg.y, qg.z, qg.w};
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int p = 4 * g + j;
                    const int c = ti * kNX + p;
                    const float q = be.bin(u, xy, p, qs[j]);
                    mono = mono && q >= 0.f;
                    sum += q;
                    lc0 = c == lc ? sum : lc0;
                    lc1 = c == lc - 1 ? sum : lc1;
                }
            }
#pragma unroll
            for (int k = 0; k < kNY; ++k) rows[k] = k == ti ? sum : rows[k];
        }
        mono = mono && isfinite(sum) && sum > 0.f;
        const float cell_sa = (float)(4.f * kPi / (double)(float)kXY);
        const float sumw = sum * cell_sa;
        const int prev = rq.prev_entry[i];
        if (prev != -1) {
            up_p = true;
            k_p = (int64_t)prev * kXY + rq.prev_cell[i];
            v_p = sumw + sl.emis[i];
            up_fp = t.newQ + k_p;
            up_cp = t.cnt + k_p;
        }
        // sample a bin: Distribution1D::sample on cdf[c] = prefix[c] / sum
        int cell = 0;
        float b0 = 0.f, b1 = 0.f;  // cdf[cell], cdf[cell - 1]
        if (sumw != 0) {
            const float f = halton(h, sid, D + 2);
            if (mono) {
                // non-decreasing, NaN-free CDF: the binary search returns the
                // first bin with cdf >= f (the last bin if none)
                int r = 0;
#pragma unroll
                for (int k = 0; k < kNY - 1; ++k) r += rows[k] / sum < f ? 1 : 0;
                float pre = 0.f;
#pragma unroll
                for (int k = 0; k < kNY - 1; ++k) pre = k == r - 1 ? rows[k] : pre;
                float u, xy;
                be.row(r, u, xy);
                float prevc = r == 0 ? 0.f : pre / sum;
                int found = kNX - 1;
                bool hit = false;
                for (int g = 0; g < kNX / 4 && !hit; ++g) {
                    const float4 qg = be.q4(r, g);
                    const float qs[4] = {qg.x, qg.y, qg.z, qg.w};
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const int p = 4 * g + j;
                        if (hit) break;
                        pre += be.bin(u, xy, p, qs[j]);
                        const float cm = pre / sum;
                        if (cm >= f || p == kNX - 1) {
                            found = p;
                            b0 = cm;
                            b1 = prevc;
                            hit = true;
                        } else {
                            prevc = cm;
                        }
                    }
                }
                cell = r * kNX + found;
            } else {
                // the reference's binary search, recomputing the prefix sums
                int l = 0, rr = kXY - 1;
                cell = -1;
                while (l < rr) {
                    const int m = (l + rr) / 2;
                    float pm, pm1;
                    be.prefix(m, pm, pm1);
                    const float cm = pm / sum, cm1 = pm1 / sum;
                    if (cm >= f) {
                        if (m == 0 || cm1 < f) { cell = m; break; }
                        rr = m - 1;
                    } else {
                        l = m + 1;
                    }
                }
                if (cell < 0) cell = l;
                float pm, pm1;
                be.prefix(cell, pm, pm1);
                b0 = pm / sum;
                b1 = pm1 / sum;
            }
        }
        const float a0 = lc0 / sum, a1 = lc1 / sum;
        float pdfq = 0.f;
        if (sumw != 0) pdfq = (lc == 0 ? a0 : a0 - a1) / cell_sa;
        if (sampled > 0) {  // direct-lighting proposal for this entry (:578-585)
            const float sa = 1.f / lpdf;
            const float ratio = fmn(1.f, 0.5f * sa / cell_sa);
            const float prop = sampled * ratio;
            if (prop > Qe[lc]) {
                up_d = true;
                k_d = (int64_t)entry * kXY + lc;
                v_d = prop;
                up_fd = t.newQ + k_d;
                up_cd = t.cnt + k_d;
            }
        }
        if (sumw != 0) {
            float p = cell == 0 ? b0 : b0 - b1;
            p = (float)((double)((float)kXY * p) / (4 * kPi));
            const v3 nd = cell_dir(px, py, cell);
            sl.ndir[i * 3] = nd.x; sl.ndir[i * 3 + 1] = nd.y; sl.ndir[i * 3 + 2] = nd.z;
            sl.npdf[i] = p;
            sl.ncell[i] = cell;
            valid = true;
        } else {
            sl.ncell[i] = -1;
        }
        // accumulateLightingContribs
        const float mp = bsdf.pdf(il, ol);
        float spdf = 0;
        if (sumw != 0) spdf = pdfq;
        if (sumw == 0 || spdf == 0 || bsdf.delta() || bsdf.almost_delta()) spdf = mp;
        const float w = power_heuristic(lpdf, spdf);
        const v3 add = scale3(mul3(contrib, mult), w);
        res[sid * 3] = res[sid * 3] + add.x;
        res[sid * 3 + 1] = res[sid * 3 + 1] + add.y;
        res[sid * 3 + 2] = res[sid * 3 + 2] + add.z;
    }
    // (both branches run for the whole wave: a pool wave may hold tasks of
    // deterministic and of atomic requests)
    if (rec.keys) {
        rec_push(rec, rec_key(rlx, true, k_l, sid_up, 0), v_l, up_l && use_rec);
        rec_push(rec, rec_key(rlx, false, k_p, sid_up, 0), v_p, up_p && use_rec);
        rec_push(rec, rec_key(rlx, false, k_d, sid_up, 1), v_d, up_d && use_rec);
    }
    wave_add(up_fl, up_cl, v_l, up_l && !use_rec);
    wave_add(up_fp, up_cp, v_p, up_p && !use_rec);
    wave_add(up_fd, up_cd, v_d, up_d && !use_rec);
    if (valid && __hip_atomic_load(&dims[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
        atomicMax(&dims[1], 1);
}

// finishUpdateQ (:97-110, 753-762) and averageQ (:52-72, 765-772) for one
// entry per wave, two bins per lane; also clears newQ / proposalCount for the
// next bounce (prepareForUpdateQ). Entries without proposals are unchanged by
// both steps and are skipped.
__device__ inline float lane_f(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

__device__ inline void finish_q_entry(const RlTables& t, int64_t e, float alpha_scale) {
    const int lane = threadIdx.x & 63;
    const int64_t base = e * kXY;
    const int c0 = lane, c1 = lane + 64;
    const int n0 = t.cnt[base + c0], n1 = t.cnt[base + c1];
    if (__ballot(n0 != 0 || n1 != 0) == 0) return;
    float q0 = t.Q[base + c0], q1 = t.Q[base + c1];
    int tot0 = t.total[base + c0], tot1 = t.total[base + c1];
    if (n0 != 0) {
        float a = (float)n0 / (float)(tot0 + n0);
        a = a * alpha_scale;
        if (a != 0) {
            q0 = q0 * (1.f - a) + a * t.newQ[base + c0] / (float)n0;
            tot0 += n0;
        }
        t.newQ[base + c0] = 0.f;
        t.cnt[base + c0] = 0;
    }
    if (n1 != 0) {
        float a = (float)n1 / (float)(tot1 + n1);
        a = a * alpha_scale;
        if (a != 0) {
            q1 = q1 * (1.f - a) + a * t.newQ[base + c1] / (float)n1;
            tot1 += n1;
        }
        t.newQ[base + c1] = 0.f;
        t.cnt[base + c1] = 0;
    }
    t.total[base + c0] = tot0;
    t.total[base + c1] = tot1;
    const uint64_t m0 = __ballot(tot0 > 0), m1 = __ballot(tot1 > 0);
    const int count = __popcll(m0) + __popcll(m1);
    if (count > 0) {
        float s = 0.f;  // in bin order, as averageQ sums (the masks are wave-uniform)
        for (int k = 0; k < 64; ++k)
            if ((m0 >> k) & 1) s += lane_f(q0, k);
        for (int k = 0; k < 64; ++k)
            if ((m1 >> k) & 1) s += lane_f(q1, k);
        const float avg = s / (float)count;
        if (tot0 == 0) q0 = avg;
        if (tot1 == 0) q1 = avg;
    }
    t.Q[base + c0] = q0;
    t.Q[base + c1] = q1;
}

__global__ __launch_bounds__(64 * kEntryWaves) void k_finish_q(RlTables t, int64_t nent, float alpha_scale) {
    const int64_t e = (int64_t)blockIdx.x * kEntryWaves + (threadIdx.x >> 6);
    if (e < nent) finish_q_entry(t, e, alpha_scale);
}

// The same over the tables of every pool request with hits this bounce
// (live[slot]), one entry per wave on a fixed grid.
__global__ __launch_bounds__(64 * kEntryWaves) void k_pool_finish_q(const RlReqState* __restrict__ rl,
                                                                    const uint8_t* __restrict__ live, int nslots) {
    const int64_t w0 = (int64_t)blockIdx.x * kEntryWaves + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * kEntryWaves;
    for (int sl = 0; sl < nslots; ++sl) {
        if (!live[sl]) continue;
        const RlReqState& R = rl[sl];
        const int64_t nent = (int64_t)R.gr.g * R.gr.g * R.gr.g;
        for (int64_t e = w0; e < nent; e += nw) finish_q_entry(R.t, e, R.alpha_scale);
    }
}

// finishLightingTraining (:463-514) per entry, clearing the new statistics
// (prepareLightingTraining).
__device__ inline void finish_light_entry(const RlTables& t, int64_t e, int L) {
    float* res = t.lres + e * L;
    float* nres = t.lnew + e * L;
    int32_t* cnt = t.lcount + e * L;
    int32_t* ncnt = t.lnewc + e * L;
    float* cdf = t.ldist + e * L;
    float sum_r = 0, valid = 0;
    bool any = false;
    for (int k = 0; k < L; ++k) {
        const int nc = ncnt[k];
        if (nc > 0) {
            const int c = cnt[k];
            const float tc = (float)(c + nc);
            const float r = ((float)c / tc) * res[k] + ((float)nc / tc) * (nres[k] / (float)nc);
            res[k] = r;
            cnt[k] = c + nc;
            sum_r += r;
            valid += 1;
            ncnt[k] = 0;
            nres[k] = 0.f;
            any = true;
        }
    }
    if (!any || !(sum_r > 0)) return;
    const float avg = sum_r / valid;
    float dens = 0;
    for (int k = 0; k < L; ++k) {
        const float r = res[k];
        if (r > 0) dens += r;
        else if (cnt[k] > 0) dens += avg / (float)cnt[k];
        else dens += avg;
        cdf[k] = dens;
    }
    for (int k = 0; k < L; ++k) cdf[k] = cdf[k] / dens;
}

// k_qdist over `count` task slots in workgroups of `block` threads (64, 128
// or 256): each task's distribution and updates are its own, so the
// workgroup size changes only the speed.
template <class Src, typename... A>
void launch_qdist(const char* name, int block, int64_t count, LaunchSink& ls, A... args) {
    const auto grid = [count](int b) { return dim3((unsigned)((count + b - 1) / b)); };
    if (block == 64) ls.launch(name, &k_qdist<Src, 64>, grid(64), dim3(64), args...);
    else if (block == 128) ls.launch(name, &k_qdist<Src, 128>, grid(128), dim3(128), args...);
    else ls.launch(name, &k_qdist<Src, 256>, grid(256), dim3(256), args...);
}

__global__ __launch_bounds__(kBlock) void k_finish_light(RlTables t, int64_t nent, int L) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < nent) finish_light_entry(t, e, L);
}

__global__ __launch_bounds__(kBlock) void k_pool_finish_light(const RlReqState* __restrict__ rl,
                                                              const uint8_t* __restrict__ live, int nslots, int L) {
    const int64_t t0 = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nt = (int64_t)gridDim.x * kBlock;
    for (int sl = 0; sl < nslots; ++sl) {
        if (!live[sl]) continue;
        const RlReqState& R = rl[sl];
        const int64_t nent = (int64_t)R.gr.g * R.gr.g * R.gr.g;
        for (int64_t e = t0; e < nent; e += nt) finish_light_entry(R.t, e, L);
    }
}

// genNextRay (:305-375) after the second syncDimension.
template <class Src>
__global__ __launch_bounds__(kBlock) void k_next(Src src, SceneDev sc, HaltonDev h, Queue qi, Queue qo,
                                                 RlQueue rqo, RlSlots sl, const int32_t* __restrict__ hit_prim,
                                                 const float* __restrict__ hit_t,
                                                 const float* __restrict__ hit_u,
                                                 const float* __restrict__ hit_v, uint32_t* counts,
                                                 int qin, int qout, const uint32_t* __restrict__ perm,
                                                 int32_t* __restrict__ rid_out, uint8_t* __restrict__ dep_out) {
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    bool live = j < (int64_t)counts[qin];
    const int64_t i = live ? task_slot(perm, j) : 0;
    live = live && src.take(i);
    const int prim = live ? hit_prim[i] : -1;
    live = live && prim >= 0;
    int depth = 0, rid = 0;
    bool push = false;
    uint32_t sid = 0;
    v3 nro = mk3(0, 0, 0), nrd = nro, mult = nro;
    float npdf = 0.f;
    int cell = 0, entry = 0, dim = 0;
    bool ndelta = false;
    if (live) {
        const RlReqState& R = src.get(i);
        const int32_t* dims = R.dims;
        depth = src.depth(i);
        rid = src.slot(i);
        dim = dims[0] + 7 + sc.light_dims + dims[1];
        sid = qi.sid[i];
        const float* r = qi.ray + i * 6;
        const v3 ro = mk3(r[0], r[1], r[2]), rd = mk3(r[3], r[4], r[5]);
        mult = mk3(qi.mult[i * 3], qi.mult[i * 3 + 1], qi.mult[i * 3 + 2]);
        bool alive = true;
        if (depth > 3) {
            float q = 1;
            q = fmn(q, 1 - mult.x);
            q = fmn(q, 1 - mult.y);
            q = fmn(q, 1 - mult.z);
            q = fmx(q, 0.05f);
            if (halton(h, sid, dim) < q) alive = false;
            else mult = divs3(mult, 1.f - q);
            ++dim;
        }
        if (alive) {
            const int mesh = sc.tri_mesh[prim];
            const float hu = hit_u ? hit_u[i] : 0.f, hv = hit_u ? hit_v[i] : 0.f;
            const Bsdf bsdf = surface_bsdf(sc, mesh, prim, hu, hv);
            const v3 nn = shading_normal(sc, prim, hu, hv);
            v3 t0, t1;
            frame(nn, t0, t1);
            const v3 pos = add3(ro, scale3(rd, hit_t[i]));
            const v3 wo = scale3(rd, -1.f);
            entry = sl.entry[i];
            const int qc = sl.ncell[i];
            v3 f;
            if (bsdf.delta() || bsdf.almost_delta() || qc < 0) {
                const float ux = halton(h, sid, dim), uy = halton(h, sid, dim + 1);
                dim += 2;
                v3 wl;
                float p = 0.f;
                f = bsdf.sample(ux, uy, wl, to_local(wo, t0, t1, nn), p);
                nrd = to_world(wl, t0, t1, nn);
                npdf = p;
                cell = dir_to_cell(nrd);
            } else {
                cell = qc;
                npdf = sl.npdf[i];
                nrd = mk3(sl.ndir[i * 3], sl.ndir[i * 3 + 1], sl.ndir[i * 3 + 2]);
                f = bsdf.eval(to_local(nrd, t0, t1, nn), to_local(wo, t0, t1, nn));
            }
            nro = add3(pos, scale3(nrd, 0.001f));
            if (zero3(f)) mult = mk3(0, 0, 0);
            else mult = divs3(scale3(mul3(mult, f), fabsf(dot3(nrd, nn))), npdf);
            ndelta = bsdf.delta();
            push = depth + 1 < R.max_depth;  // the last bounce's rays are never traced
        }
    }
    const uint32_t ns = block_push<kBlock>(push, &counts[qout]);
    if (push) {
        float* r = qo.ray + (int64_t)ns * 6;
        r[0] = nro.x; r[1] = nro.y; r[2] = nro.z; r[3] = nrd.x; r[4] = nrd.y; r[5] = nrd.z;
        qo.sid[ns] = sid;
        qo.mult[(int64_t)ns * 3] = mult.x;
        qo.mult[(int64_t)ns * 3 + 1] = mult.y;
        qo.mult[(int64_t)ns * 3 + 2] = mult.z;
        qo.pdf[ns] = npdf;
        qo.delta[ns] = ndelta ? 1 : 0;
        rqo.prev_entry[ns] = entry;
        rqo.prev_cell[ns] = cell;
        rqo.dim[ns] = dim;
        if (rid_out) {
            rid_out[ns] = rid;
            dep_out[ns] = (uint8_t)(depth + 1);
        }
    }
}

__global__ void k_step(uint32_t* counts, unsigned long long* rays, int32_t* dims, int qin) {
    if (threadIdx.x == 0) {
        rays[0] += counts[qin];
        rays[1] += counts[2];
        counts[qin] = 0;
        counts[2] = 0;
        dims[0] = 0;
        dims[1] = 0;
    }
}

// A pool bounce's end for its rlpath requests: clear the syncDimension words
// of the requests that had hits, and their marks.
__global__ void k_pool_rl_step(const RlReqState* __restrict__ rl, uint8_t* live, int nslots) {
    for (int sl = threadIdx.x; sl < nslots; sl += blockDim.x)
        if (live[sl]) {
            rl[sl].dims[0] = 0;
            rl[sl].dims[1] = 0;
            live[sl] = 0;
        }
}

// sortLightingQueue's key (:226-272): material * entries + Q entry of the hit
// (misses after every material), value = queue slot, for the first *count
// slots; the stable radix sort of these pairs is the shading order.
__global__ __launch_bounds__(kBlock) void k_entry_keys(SceneDev sc, Grid gr, Queue qi, const int32_t* __restrict__ hit_prim,
                                                       const float* __restrict__ hit_t, const uint32_t* count,
                                                       int64_t cap, uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= cap || i >= (int64_t)*count) return;
    const uint32_t nent = (uint32_t)gr.g * (uint32_t)gr.g * (uint32_t)gr.g;
    const int prim = hit_prim[i];
    uint32_t key = 7u * nent;
    if (prim >= 0) {
        const int m = sc.meshes[sc.tri_mesh[prim]].material;
        const float* r = qi.ray + i * 6;
        const v3 pos = add3(mk3(r[0], r[1], r[2]), scale3(mk3(r[3], r[4], r[5]), hit_t[i]));
        key = (uint32_t)(m >= 0 && m < 6 ? m : 6) * nent + (uint32_t)q_entry(gr, pos);
    }
    keys[i] = key;
    vals[i] = (uint32_t)i;
}

// Deterministic mode, before k_qdist: no records yet (k_qdist appends them;
// the sort takes only the appended ones).
__global__ void k_rec_reset(uint32_t* count) {
    if (threadIdx.x == 0) *count = 0u;
}

// The tables a record's request index selects.
struct RenderTables {
    RlTables t;
    __device__ const RlTables& of(uint32_t) const { return t; }
};
struct PoolTables {
    const RlReqState* rl;
    __device__ const RlTables& of(uint32_t x) const { return rl[x].t; }
};

// After the sort: the first record of each key sums its run in order, from 0
// as the oracle's += does, and sets the table word and its count (the tables
// hold zeros there: k_finish_q / k_finish_light cleared them).
template <class Tab>
__global__ __launch_bounds__(kBlock) void k_rec_sum(const uint64_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ idx,
                                                    const float* __restrict__ val, const uint32_t* count,
                                                    int64_t cap, Tab tab) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t n = (int64_t)*count < cap ? (int64_t)*count : cap;
    if (i >= n) return;
    const uint64_t k = keys[i];
    const uint64_t seg = k >> kRecKeyShift;
    if (i > 0 && (keys[i - 1] >> kRecKeyShift) == seg) return;
    float sum = 0.f;
    int32_t c = 0;
    for (int64_t j = i; j < n && (keys[j] >> kRecKeyShift) == seg; ++j) {
        sum += val[idx[j]];
        ++c;
    }
    const RlTables& t = tab.of((uint32_t)(k >> kRecReqShift));
    const int64_t key = (int64_t)(seg & ((1ull << kRecKeyBits) - 1ull));
    if ((seg >> kRecKeyBits) & 1ull) {
        t.lnew[key] = sum;
        t.lnewc[key] = c;
    } else {
        t.newQ[key] = sum;
        t.cnt[key] = c;
    }
}

// Sorts a bounce's records by key (only the appended ones) and sums them.
template <class Tab>
hipError_t rec_sort_sum(uint64_t* const keys[2], uint32_t* const idx[2], float* val, uint32_t* count,
                        uint32_t* hist, uint32_t* counts, uint32_t* offsets, int64_t cap, Tab tab,
                        LaunchSink& ls) {
    SortBuffers<uint64_t> sb;
    sb.keys[0] = keys[0]; sb.keys[1] = keys[1];
    sb.vals[0] = idx[0]; sb.vals[1] = idx[1];
    sb.hist = hist; sb.counts = counts; sb.offsets = offsets;
    sb.n_dev = count;
    hipError_t e = radix_sort_pairs(sb, cap, 64, ls);
    if (e != hipSuccess) return e;
    ls.launch("k_rec_sum", &k_rec_sum<Tab>, dim3((unsigned)((cap + kBlock - 1) / kBlock)), dim3(kBlock),
              sb.keys[sb.result], sb.vals[sb.result], val, count, cap, tab);
    return ls.err;
}

}  // namespace

hipError_t render_rlpath(const RenderArgs& a, const RlArgs& rl, bool init_tables, hipStream_t s) {
    const int64_t nsamp = a.nq;
    float* hu = a.scene.needs_bary() ? a.hit_u : nullptr;
    float* hv = a.scene.needs_bary() ? a.hit_v : nullptr;
    const unsigned nb = (unsigned)((nsamp + kBlock - 1) / kBlock);
    const int L = a.scene.nlights;
    const int64_t nent = (int64_t)rl.grid * rl.grid * rl.grid;
    const int64_t ncells = nent * kXY;
    Grid gr;
    for (int k = 0; k < 3; ++k) {
        gr.lo[k] = rl.bounds[k];
        gr.size[k] = rl.bounds[3 + k] - rl.bounds[k];
    }
    gr.g = rl.grid;
    RlTables t{rl.Q, rl.newQ, rl.cnt, rl.total, rl.ldist, rl.lres, rl.lnew, rl.lcount, rl.lnewc};
    const RlReqState st{t, gr, a.res, rl.dims, rl.alpha_scale, 0u, a.max_depth, rl.rec_keys[0] ? 1 : 0};
    hipError_t e = hipMemsetAsync(a.counts, 0, 64, s);
    if (e != hipSuccess) return e;
    if ((e = hipMemsetAsync(rl.dims, 0, 16, s)) != hipSuccess) return e;
    const int64_t ninit = ncells > nent * L ? ncells : nent * L;
    if (init_tables)
        hipLaunchKernelGGL(k_init_tables, dim3((unsigned)((ninit + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, t,
                       ncells, nent, L);
    render_raygen(a, s);
    hipLaunchKernelGGL(k_init_tasks, dim3(nb), dim3(kBlock), 0, s, nsamp, rl.q[0]);
    LaunchSink ls(s);
    int qin = 0;
    for (int depth = 0; depth < a.max_depth; ++depth) {
        const int qout = qin ^ 1;
        const RenderSrc src{st, depth};
        if ((e = trace_closest_queue(a.scene.rec, a.scene.ntris, a.scene.root, a.q[qin].ray,
                                     a.counts + qin, nsamp, a.hit_prim, a.hit_t, hu, hv,
                                     a.trace_counters, a.launch, s, &a.coop)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_dim_max, dim3((unsigned)((nsamp + kMaxBlock - 1) / kMaxBlock)), dim3(kMaxBlock), 0, s,
                           a.counts, qin, rl.q[qin], a.hit_prim, rl.dims);
        const uint32_t* perm = nullptr;
        if (a.sort_perm) {  // shading order: stable sort by (material, Q entry)
            const unsigned kb = (unsigned)((nsamp + kBlock - 1) / kBlock);
            hipLaunchKernelGGL(k_entry_keys, dim3(kb), dim3(kBlock), 0, s, a.scene, gr, a.q[qin], a.hit_prim,
                               a.hit_t, a.counts + qin, nsamp, rl.skeys[0], rl.svals[0]);
            SortBuffers<uint32_t> sb;
            sb.keys[0] = rl.skeys[0]; sb.keys[1] = rl.skeys[1];
            sb.vals[0] = rl.svals[0]; sb.vals[1] = rl.svals[1];
            sb.hist = rl.shist; sb.counts = rl.scounts; sb.offsets = rl.soffsets;
            sb.n_dev = a.counts + qin;
            int bits = 1;
            while (((int64_t)1 << bits) < 8 * nent) ++bits;
            if ((e = radix_sort_pairs(sb, nsamp, bits, s, nullptr)) != hipSuccess) return e;
            perm = sb.vals[sb.result];
        }
        ls.launch("k_lighting", &k_lighting<RenderSrc>, dim3(nb), dim3(kBlock), src, a.scene, a.halton, a.q[qin],
                  rl.slots, a.sh, a.hit_prim, a.hit_t, hu, hv, a.counts, qin, perm);
        if ((e = trace_occluded_queue(a.scene.rec, a.scene.ntris, a.sh.ray, a.counts + 2, nsamp,
                                      a.sh.limit, a.sh.src, a.sh.dst, a.vis, a.trace_counters, a.launch, s, &a.coop)) != hipSuccess)
            return e;
        const RlRecords rec{rl.rec_keys[0], rl.rec_val, rl.rec_count, rl.rec_idx[0]};
        if (rec.keys) ls.launch("k_rec_reset", &k_rec_reset, dim3(1), dim3(64), rl.rec_count);
        launch_qdist<RenderSrc>("k_qdist", a.launch.shade_block, nsamp, ls, src, a.scene, a.halton, a.q[qin],
                                rl.q[qin], rl.slots, a.vis, a.hit_prim, a.hit_t, hu, hv, a.counts, qin, perm, rec);
        if (rec.keys &&
            (e = rec_sort_sum(rl.rec_keys, rl.rec_idx, rl.rec_val, rl.rec_count, rl.rec_hist, rl.rec_counts,
                              rl.rec_offsets, rl.rec_cap, RenderTables{t}, ls)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_finish_q, dim3((unsigned)((nent + kEntryWaves - 1) / kEntryWaves)),
                           dim3(64 * kEntryWaves), 0, s, t, nent, rl.alpha_scale);
        hipLaunchKernelGGL(k_finish_light, dim3((unsigned)((nent + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, t,
                           nent, L);
        ls.launch("k_next", &k_next<RenderSrc>, dim3(nb), dim3(kBlock), src, a.scene, a.halton, a.q[qin], a.q[qout],
                  rl.q[qout], rl.slots, a.hit_prim, a.hit_t, hu, hv, a.counts, qin, qout, perm, (int32_t*)nullptr,
                  (uint8_t*)nullptr);
        hipLaunchKernelGGL(k_step, dim3(1), dim3(64), 0, s, a.counts, a.rays, rl.dims, qin);
        qin = qout;
        if (ls.err != hipSuccess) return ls.err;
    }
    return hipGetLastError();
}

// ------------------------------------------------------------ server pools
namespace {

PoolSrc pool_src(const PoolDev& p, int qin) {
    return PoolSrc{p.rl.req, p.rid[qin], p.depth[qin], p.sh_rid, p.rl.live};
}

__global__ void k_clear_dims(int32_t* dims) {
    if (threadIdx.x < 2) dims[threadIdx.x] = 0;
}

}  // namespace

void pool_rl_lighting(const PoolDev& p, int qin, LaunchSink& ls) {
    const unsigned nb = (unsigned)((p.cap + kBlock - 1) / kBlock);
    const bool bary = p.scene.needs_bary();
    const PoolSrc src = pool_src(p, qin);
    ls.launch("k_pool_dim_max", &k_pool_dim_max, dim3(nb), dim3(kBlock), src, p.counts, qin, p.rl.q[qin], p.hit_prim);
    ls.launch("k_lighting<pool>", &k_lighting<PoolSrc>, dim3(nb), dim3(kBlock), src, p.scene, p.halton, p.q[qin],
              p.rl.sl, p.sh, p.hit_prim, p.hit_t, bary ? p.hit_u : nullptr, bary ? p.hit_v : nullptr, p.counts, qin,
              (const uint32_t*)nullptr);
}

void pool_rl_learn(const PoolDev& p, int qin, bool det, LaunchSink& ls) {
    const int qout = qin ^ 1;
    const unsigned nb = (unsigned)((p.cap + kBlock - 1) / kBlock);
    const bool bary = p.scene.needs_bary();
    float* hu = bary ? p.hit_u : nullptr;
    float* hv = bary ? p.hit_v : nullptr;
    const PoolSrc src = pool_src(p, qin);
    RlRecords rec{nullptr, nullptr, nullptr, nullptr};
    if (det) {
        rec = RlRecords{p.rl.rec_keys[0], p.rl.rec_val, p.rl.rec_count, p.rl.rec_idx[0]};
        ls.launch("k_rec_reset", &k_rec_reset, dim3(1), dim3(64), p.rl.rec_count);
    }
    launch_qdist<PoolSrc>("k_qdist<pool>", p.launch.shade_block, p.cap, ls, src, p.scene, p.halton, p.q[qin],
                          p.rl.q[qin], p.rl.sl, p.vis, p.hit_prim, p.hit_t, hu, hv, p.counts, qin,
                          (const uint32_t*)nullptr, rec);
    if (det)
        (void)rec_sort_sum(p.rl.rec_keys, p.rl.rec_idx, p.rl.rec_val, p.rl.rec_count, p.rl.rec_hist, p.rl.rec_counts,
                           p.rl.rec_offsets, p.rl.rec_cap, PoolTables{p.rl.req}, ls);
    // the updates of every request with hits; a fixed grid, so that the
    // launch does not depend on which requests are in the pool
    ls.launch("k_pool_finish_q", &k_pool_finish_q, dim3(2048), dim3(64 * kEntryWaves), (const RlReqState*)p.rl.req,
              (const uint8_t*)p.rl.live, kPoolSlots);
    ls.launch("k_pool_finish_light", &k_pool_finish_light, dim3(512), dim3(kBlock), (const RlReqState*)p.rl.req,
              (const uint8_t*)p.rl.live, kPoolSlots, p.scene.nlights);
    ls.launch("k_next<pool>", &k_next<PoolSrc>, dim3(nb), dim3(kBlock), src, p.scene, p.halton, p.q[qin], p.q[qout],
              p.rl.q[qout], p.rl.sl, p.hit_prim, p.hit_t, hu, hv, p.counts, qin, qout, (const uint32_t*)nullptr,
              p.rid[qout], p.depth[qout]);
    ls.launch("k_pool_rl_step", &k_pool_rl_step, dim3(1), dim3(kPoolSlots), (const RlReqState*)p.rl.req, p.rl.live,
              kPoolSlots);
}

hipError_t pool_rl_init(const PoolDev& p, const RlReqState& st, int nlights, hipStream_t s) {
    (void)p;
    const int64_t nent = (int64_t)st.gr.g * st.gr.g * st.gr.g;
    const int64_t ncells = nent * kXY;
    const int64_t ninit = ncells > nent * nlights ? ncells : nent * nlights;
    LaunchSink ls(s);
    ls.launch("k_init_tables", &k_init_tables, dim3((unsigned)((ninit + kBlock - 1) / kBlock)), dim3(kBlock), st.t,
              ncells, nent, nlights);
    ls.launch("k_clear_dims", &k_clear_dims, dim3(1), dim3(64), st.dims);
    return ls.err;
}

}  // namespace cav
