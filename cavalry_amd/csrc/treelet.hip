// Aila-Karras 7-leaf treelet restructuring, one treelet per 32-lane group.
//
// Reference: optimizeTreelet / optimizeBVHImpl / optimizeBVH
// (BVH/Optimization.cu:49-525). The reference runs one 32-lane warp per
// treelet; here each wave64 carries two independent 32-lane groups, and every
// cross-lane step uses 32-wide shuffles so the decisions follow the reference
// lane for lane: formation by largest surface area (lowest lane on ties), the
// scheduled partition enumeration for subsets of 2..5 leaves, the
// min(l+d,31) shuffle-reduction tree for 6 and 7 leaves, and reconstruction
// that hands the two halves of the first proposing lane to the next two lanes.
//
// Race-free by construction: a lane reads the treelet data it needs into
// registers before anything is written, and box/area/cost are recomputed
// bottom-up through shuffles instead of in-place sweeps over global memory.
#include "dev_common.h"
#include "lbvh.h"

namespace cav {
namespace {

constexpr int kBlock = 256;
constexpr int kGroups = kBlock / 32;
constexpr int kT = 13;  // 2*7-1 treelet nodes

// BVH/Optimization.cu:20-47
__constant__ uint8_t kSchedule[160] = {
    0b00011, 0b00110, 0b00101, 0b01001, 0b01010, 0b01100, 0b10001, 0b10010, 0b10100, 0b11000,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    0b00111, 0b01011, 0b01101, 0b01110, 0b10011, 0b10101, 0b10110, 0b11001, 0b11010, 0b11100,
    0b100001, 0b100010, 0b100100, 0b101000, 0b110000,
    0b1000001, 0b1000010, 0b1000100, 0b1001000, 0b1010000, 0b1100000,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    0b100011, 0b100101, 0b100110, 0b101001, 0b101010, 0b101100, 0b110001, 0b110010, 0b110100, 0b111000,
    0b1000011, 0b1000101, 0b1000110, 0b1001001, 0b1001010, 0b1001100, 0b1010001, 0b1010010,
    0b1010100, 0b1011000, 0b1100001, 0b1100010, 0b1100100, 0b1101000, 0b1110000,
    0b0001111, 0b0010111, 0b0011110,
    0, 0, 0, 0,
    0b1000111, 0b0011011, 0b0101011, 0b1001011, 0b0110011, 0b1010011, 0b1100011, 0b0011101,
    0b0101101, 0b1001101, 0b0110101, 0b1010101, 0b1100101, 0b0111001, 0b1011001, 0b1101001,
    0b1110001, 0b0100111, 0b0101110, 0b1001110, 0b0110110, 0b1010110, 0b1100110, 0b0111010,
    0b1011010, 0b1101010, 0b1110010, 0b0111100, 0b1011100, 0b1101100, 0b1110100, 0b1111000,
    0b0011111, 0b0101111, 0b1001111, 0b0110111, 0b1010111, 0b1100111, 0b0111011, 0b1011011,
    0b1101011, 0b1110011, 0b0111101, 0b1011101, 0b1101101, 0b1110101, 0b1111001, 0b0111110,
    0b1011110, 0b1101110, 0b1110110, 0b1111010, 0b1111100,
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

struct GroupLds {
    float area[128];
    float cost[128];
    float lbox[7][6];
    int part[128];
};

__device__ __forceinline__ float shf(float v, int src) { return __shfl(v, src, 32); }
__device__ __forceinline__ int shi(int v, int src) { return __shfl(v, src, 32); }
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

struct LaneNode {
    int node;
    float lx, ly, lz, hx, hy, hz, sa, cost;
    int left, right, parent, prim;
};

__device__ __forceinline__ void load_node(LaneNode& n, const NodeBox* nbox, const int4* topo, int id) {
    n.node = id;
    const NodeBox b = nbox[id];
    const int4 t = topo[id];
    n.lx = b.a.x; n.ly = b.a.y; n.lz = b.a.z; n.sa = b.a.w;
    n.hx = b.b.x; n.hy = b.b.y; n.hz = b.b.z; n.cost = b.b.w;
    n.left = t.x; n.right = t.y; n.parent = t.z; n.prim = t.w;
}

// BVH/Optimization.cu:49-462 for the treelet rooted at `root` (an internal
// node). Every lane of the group calls it; control flow is group-uniform.
__device__ void optimize_treelet(int root, NodeBox* nbox, int4* topo, GroupLds& L) {
    const int lane = threadIdx.x & 31;
    LaneNode me;
    me.node = -1;
    if (lane == 0) load_node(me, nbox, topo, root);
    bool expanded = false;

    // formation
    int size = 1;
    while (size < kT) {
        float my_area = -1.f;
        if (me.node != -1 && !(me.prim >= 0 || expanded)) my_area = me.sa;
        int max_lane = 0;
        float max_area = -1.f;
        for (int i = 0; i < size; ++i) {
            const float a = shf(my_area, i);
            if (a > max_area) { max_area = a; max_lane = i; }
        }
        int cand = 0;
        if (lane == max_lane) {
            if (expanded || me.prim >= 0) cand = -1;
            else { expanded = true; cand = me.node; }
        }
        const int nx = shi(cand, max_lane);
        if (nx == -1) return;
        const int cl = shi(me.left, max_lane), cr = shi(me.right, max_lane);
        if (lane == size) load_node(me, nbox, topo, cl);
        if (lane == size + 1) load_node(me, nbox, topo, cr);
        size += 2;
    }
    // treelet leaves in lane order
    int leaf_lane[7];
    {
        int k = 0;
        for (int i = 0; i < kT; ++i) {
            const int ne = shi(expanded ? 0 : 1, i);
            if (ne) { leaf_lane[k] = i; ++k; }
        }
    }
    for (int k = 0; k < 7; ++k)
        if (lane == leaf_lane[k]) {
            L.lbox[k][0] = me.lx; L.lbox[k][1] = me.ly; L.lbox[k][2] = me.lz;
            L.lbox[k][3] = me.hx; L.lbox[k][4] = me.hy; L.lbox[k][5] = me.hz;
        }
    wave_sync();

    // subset areas: lane l covers subsets (i << 5) | l, leaves unioned in bit order
    {
        float bx[6];
        bool has = false;
        for (int bit = 0; bit < 5; ++bit)
            if ((lane >> bit) & 1) {
                if (!has) { for (int c = 0; c < 6; ++c) bx[c] = L.lbox[bit][c]; has = true; }
                else {
                    for (int c = 0; c < 3; ++c) bx[c] = fmn(bx[c], L.lbox[bit][c]);
                    for (int c = 3; c < 6; ++c) bx[c] = fmx(bx[c], L.lbox[bit][c]);
                }
            }
        for (int i = 0; i < 4; ++i) {
            const int s = (i << 5) | lane;
            float b2[6];
            bool ne = has;
            for (int c = 0; c < 6; ++c) b2[c] = bx[c];
            for (int bit = 5; bit < 7; ++bit)
                if ((s >> bit) & 1) {
                    if (!ne) { for (int c = 0; c < 6; ++c) b2[c] = L.lbox[bit][c]; ne = true; }
                    else {
                        for (int c = 0; c < 3; ++c) b2[c] = fmn(b2[c], L.lbox[bit][c]);
                        for (int c = 3; c < 6; ++c) b2[c] = fmx(b2[c], L.lbox[bit][c]);
                    }
                }
            L.area[s] = ne ? surface_area(mk3(b2[0], b2[1], b2[2]), mk3(b2[3], b2[4], b2[5])) : 0.f;
            L.cost[s] = -1.f;
        }
    }
    wave_sync();
    for (int k = 0; k < 7; ++k)
        if (lane == leaf_lane[k]) L.cost[1 << k] = me.cost;
    wave_sync();

    // sizes 2..5 on the fixed schedule
    for (int round = 0; round < 5; ++round) {
        const int s = kSchedule[round * 32 + lane];
        float best = -1.f;
        int bestp = 0;
        if (s != 0) {
            const int delta = (s - 1) & s;
            int p = (0 - delta) & s;
            do {
                const float c = L.cost[p] + L.cost[s ^ p];
                if (best == -1.f || c < best) { best = c; bestp = p; }
                p = (p - delta) & s;
            } while (p != 0);
        }
        wave_sync();
        if (s != 0) {
            L.cost[s] = internal_cost2(L.area[s], best);
            L.part[s] = bestp;
        }
        wave_sync();
    }

    // sizes 6 and 7: one partition per lane (two for size 7), reduced over
    // the 32 lanes with the reference's min(l+d,31) tree
    for (int e = 0; e <= 7; ++e) {
        const int s = e < 7 ? (0x7f ^ (1 << e)) : 0x7f;
        int part;
        float tc = -1.f;
        if (e < 7) {
            // lane bits map onto the subset's bits below its top bit; for e = 6
            // there are no bits above e to fill
            const int hm = 5 - e > 0 ? (((1 << (5 - e)) - 1) << (e + 1)) : 0;
            part = (lane & ((1 << e) - 1)) | ((lane << 1) & hm);
            if (part != 0) tc = L.cost[part] + L.cost[s ^ part];
        } else {
            part = lane;
            if (part != 0) tc = L.cost[part] + L.cost[s ^ part];
            const int q = lane | 32;
            const float oc = L.cost[q] + L.cost[s ^ q];
            if (tc == -1.f || oc < tc) { tc = oc; part = q; }
        }
        for (int d = 16; d >= 1; d >>= 1) {
            const int ol = min(lane + d, 31);
            const float oc = shf(tc, ol);
            const int op = shi(part, ol);
            if (tc == -1.f || oc < tc) { tc = oc; part = op; }
        }
        wave_sync();
        if (lane == 0) {
            L.cost[s] = internal_cost2(L.area[s], tc);
            L.part[s] = part;
        }
        wave_sync();
    }

    // reconstruction
    int my_subset = lane == 0 ? 0x7f : -1;
    int my_parent = lane == 0 ? me.parent : -1;
    bool rexp = false;
    int new_left = -1, new_right = -1, child_lane_l = -1, child_lane_r = -1;
    int rsize = 1;
    while (rsize < kT) {
        const bool propose = lane < rsize && !rexp && my_subset != -1 && (my_subset & (my_subset - 1)) != 0;
        const uint64_t mask = __ballot(propose);
        const uint32_t gmask = (uint32_t)(mask >> (threadIdx.x & 32));
        if (gmask == 0u) return;  // unreachable for a valid partition table
        const int xl = __ffs((int)gmask) - 1;  // first proposing lane
        const int s = shi(my_subset, xl);
        const int pn = shi(me.node, xl);
        const int p = L.part[s];
        const int nl = shi(me.node, rsize), nr = shi(me.node, rsize + 1);
        if (lane == xl) {
            rexp = true;
            new_left = nl; new_right = nr;
            child_lane_l = rsize; child_lane_r = rsize + 1;
        }
        if (lane == rsize) { my_parent = pn; my_subset = p; }
        if (lane == rsize + 1) { my_parent = pn; my_subset = s ^ p; }
        rsize += 2;
    }

    // treelet leaves take over the content of the original leaf of their bit
    int src_lane = 0;
    if (lane < kT && !rexp) {
        int bit = __ffs(my_subset) - 1;
        for (int k = 0; k < 7; ++k) if (k == bit) src_lane = leaf_lane[k];
    }
    float lx = shf(me.lx, src_lane), ly = shf(me.ly, src_lane), lz = shf(me.lz, src_lane);
    float hx = shf(me.hx, src_lane), hy = shf(me.hy, src_lane), hz = shf(me.hz, src_lane);
    const int oleft = shi(me.left, src_lane), oright = shi(me.right, src_lane);
    const int oprim = shi(me.prim, src_lane);
    float sa = 0.f, cost = 0.f;
    if (lane < kT && !rexp) {
        sa = surface_area(mk3(lx, ly, lz), mk3(hx, hy, hz));
        if (oprim >= 0) cost = leaf_cost(sa);
        else cost = internal_cost3(sa, nbox[oleft].b.w, nbox[oright].b.w);
    }
    // bottom-up recompute of the new internal nodes (treelet depth <= 6)
    for (int it = 0; it < 7; ++it) {
        const int cl = child_lane_l < 0 ? lane : child_lane_l;
        const int cr = child_lane_r < 0 ? lane : child_lane_r;
        const float alx = shf(lx, cl), aly = shf(ly, cl), alz = shf(lz, cl);
        const float ahx = shf(hx, cl), ahy = shf(hy, cl), ahz = shf(hz, cl);
        const float blx = shf(lx, cr), bly = shf(ly, cr), blz = shf(lz, cr);
        const float bhx = shf(hx, cr), bhy = shf(hy, cr), bhz = shf(hz, cr);
        const float ac = shf(cost, cl), bc = shf(cost, cr);
        if (rexp) {
            lx = fmn(alx, blx); ly = fmn(aly, bly); lz = fmn(alz, blz);
            hx = fmx(ahx, bhx); hy = fmx(ahy, bhy); hz = fmx(ahz, bhz);
            sa = surface_area(mk3(lx, ly, lz), mk3(hx, hy, hz));
            cost = internal_cost3(sa, ac, bc);
        }
    }
    if (lane < kT) {
        const int l = rexp ? new_left : oleft;
        const int r = rexp ? new_right : oright;
        const int pr = rexp ? -1 : oprim;
        nbox[me.node] = NodeBox{make_float4(lx, ly, lz, sa), make_float4(hx, hy, hz, cost)};
        topo[me.node] = make_int4(l, r, my_parent, pr);
        if (!rexp && oprim < 0) {
            reinterpret_cast<int*>(&topo[oleft])[2] = me.node;
            reinterpret_cast<int*>(&topo[oright])[2] = me.node;
        }
    }
}

// BVH/Optimization.cu:467-507, scheduled by height: every internal node of
// height h is the root of one treelet; its 32-lane group runs once all nodes of
// lower height are done (kernel boundaries or, for the sparse top levels, the
// barriers of a single workgroup).
__global__ __launch_bounds__(kBlock) void k_treelet_level(NodeBox* nbox, int4* topo,
                                                          const uint32_t* __restrict__ list,
                                                          int count) {
    __shared__ GroupLds lds[kGroups];
    const int group = (int)(((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 5);
    if (group >= count) return;
    optimize_treelet((int)list[group], nbox, topo, lds[threadIdx.x >> 5]);
}

__global__ __launch_bounds__(1024) void k_treelet_tail(NodeBox* nbox, int4* topo,
                                                       const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ off, int h0,
                                                       int h1) {
    __shared__ GroupLds lds[32];
    const int g = threadIdx.x >> 5;
    for (int h = h0; h <= h1; ++h) {
        const uint32_t b = off[h], e = off[h + 1];
        for (uint32_t i = b + g; i < e; i += 32) optimize_treelet((int)list[i], nbox, topo, lds[g]);
        __syncthreads();
    }
}

__global__ void k_leaf_ids(const int32_t* __restrict__ ref, int64_t m, uint32_t* __restrict__ ids) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) { const int32_t r = ref[i]; if (r < 0) ids[~r] = (uint32_t)i; }
}

}  // namespace

// One restructuring round. The first round buckets the nodes by the heights
// the build's climb left (same tree); later rounds list the leaves by a scan
// and climb again.
hipError_t treelet_round_impl(LbvhBuffers& b, int n, bool first, hipStream_t s, Marker* mk) {
    const int64_t m = 2 * (int64_t)n - 1;
    hipError_t e;
    if (!first) {
        if ((e = lbvh_refs(b, n, s, nullptr, nullptr)) != hipSuccess) return e;
        hipLaunchKernelGGL(k_leaf_ids, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, b.ref, m, b.leaf_ids);
        if ((e = levels_from_list(b, n, s, nullptr)) != hipSuccess) return e;
    }
    if ((e = levels_bucket(b, n, s, nullptr)) != hipSuccess) return e;
    mark(mk, "treelet_setup");
    const int top = b.height;
    for (int h = 3; h <= top; ++h) {
        const uint32_t rest = b.h_off[top + 1] - b.h_off[h];
        if (rest <= (uint32_t)kTreeletTailNodes) {
            hipLaunchKernelGGL(k_treelet_tail, dim3(1), dim3(1024), 0, s, b.nbox, b.topo, b.lvl_list,
                               b.lvl_off, h, top);
            break;
        }
        const int cnt = (int)b.h_lvl[h];
        if (cnt == 0) continue;
        const int64_t blocks = ((int64_t)cnt * 32 + kBlock - 1) / kBlock;
        hipLaunchKernelGGL(k_treelet_level, dim3((unsigned)blocks), dim3(kBlock), 0, s, b.nbox,
                           b.topo, b.lvl_list + b.h_off[h], cnt);
    }
    mark(mk, "treelet");
    return hipGetLastError();
}

}  // namespace cav
