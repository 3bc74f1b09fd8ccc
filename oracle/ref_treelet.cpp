// CPU oracle: Aila-Karras 7-leaf treelet restructuring.
//
// TEST INFRASTRUCTURE ONLY: a restatement of BVH/Optimization.cu used to check
// the HIP kernel. Not part of the product.
//
// The reference runs one 32-lane warp per treelet; the lane-level steps that
// decide the result (formation order, partition enumeration order, the
// shuffle-reduction tree for 6- and 7-leaf subsets, reconstruction order) are
// simulated here lane by lane. Nodes use the unified index space described in
// ref_lbvh.cpp.
#include "ref_api.h"
#include "ref_math.h"

#include <cstdio>
#include <vector>

using namespace ref;

namespace {

// BVH/Optimization.cu:20-47: subsets of size 2..5 in five rounds of 32 lanes.
const uint8_t kSchedule[160] = {
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

struct Nodes {
    float* box; float* sa; float* cost;
    int32_t *left, *right, *parent, *prim;
    Box get_box(int i) const {
        const float* b = box + (int64_t)i * 6;
        return Box{mk(b[0], b[1], b[2]), mk(b[3], b[4], b[5])};
    }
    void set_box(int i, const Box& v) {
        float* b = box + (int64_t)i * 6;
        b[0] = v.lo.x; b[1] = v.lo.y; b[2] = v.lo.z; b[3] = v.hi.x; b[4] = v.hi.y; b[5] = v.hi.z;
    }
    bool is_leaf(int i) const { return prim[i] >= 0; }
};

// 32-lane min-reduction used for the 6- and 7-leaf subsets
// (BVH/Optimization.cu:251-261 and 310-319): at each step d = 16..1 every lane
// reads lane min(l+d, 31) and takes its value if its own cost is -1 or the
// other cost is strictly lower. Lane 0's final value is the result.
static void lane_reduce(float cost[32], int part[32]) {
    for (int d = 16; d >= 1; d >>= 1) {
        float oc[32]; int op[32];
        for (int l = 0; l < 32; ++l) {
            int o = l + d < 31 ? l + d : 31;
            oc[l] = cost[o]; op[l] = part[o];
        }
        for (int l = 0; l < 32; ++l)
            if (cost[l] == -1.f || oc[l] < cost[l]) { cost[l] = oc[l]; part[l] = op[l]; }
    }
}

// BVH/Optimization.cu:49-462 for the treelet rooted at `root`. Returns 1 if the
// treelet was restructured, 0 if it has fewer than 7 leaves.
static int optimize_treelet(Nodes& nd, int root) {
    const int T = 13;  // 2*7-1 treelet nodes
    int lane_node[T];
    bool expanded[T];
    for (int l = 0; l < T; ++l) { lane_node[l] = -1; expanded[l] = false; }
    lane_node[0] = root;

    // formation: expand the unexpanded internal lane of largest surface area
    // (strict >, lowest lane wins ties) into the next two lanes.
    int size = 1;
    while (size < T) {
        int max_lane = 0;
        float max_area = -1.f;
        for (int l = 0; l < size; ++l) {
            float a = nd.sa[lane_node[l]];
            if (nd.is_leaf(lane_node[l]) || expanded[l]) a = -1.f;
            if (a > max_area) { max_area = a; max_lane = l; }
        }
        int n = lane_node[max_lane];
        if (expanded[max_lane] || nd.is_leaf(n)) return 0;
        expanded[max_lane] = true;
        lane_node[size] = nd.left[n];
        lane_node[size + 1] = nd.right[n];
        size += 2;
    }
    int leaves[7], nl = 0;
    for (int l = 0; l < T; ++l)
        if (!expanded[l]) leaves[nl++] = lane_node[l];
    if (nl != 7) { std::fprintf(stderr, "treelet: %d leaves\n", nl); return -1; }

    // subset surface areas (leaf boxes unioned in bit order)
    float area[128], opt_cost[128];
    uint8_t opt_part[128] = {0};
    for (int s = 0; s < 128; ++s) {
        bool any = false;
        Box b{};
        for (int bit = 0; bit < 7; ++bit)
            if (s & (1 << bit)) {
                Box lb = nd.get_box(leaves[bit]);
                b = any ? union_box(b, lb) : lb;
                any = true;
            }
        area[s] = any ? surface_area(b) : 0.f;
        opt_cost[s] = -1.f;
    }
    for (int i = 0; i < 7; ++i) opt_cost[1 << i] = nd.cost[leaves[i]];

    // sizes 2..5: each scheduled subset enumerates its partitions in the order
    // p = (-delta)&s, p = (p-delta)&s, ... keeping the first strict minimum.
    for (int round = 0; round < 5; ++round) {
        float rc[32]; uint8_t rp[32]; uint8_t rs[32];
        for (int l = 0; l < 32; ++l) {
            int s = kSchedule[round * 32 + l];
            rs[l] = (uint8_t)s;
            if (!s) continue;
            int delta = (s - 1) & s;
            int p = (0 - delta) & s;
            float best = -1.f; int bestp = 0;
            do {
                float c = opt_cost[p] + opt_cost[s ^ p];
                if (best == -1.f || c < best) { best = c; bestp = p; }
                p = (p - delta) & s;
            } while (p != 0);
            rc[l] = internal_cost2(area[s], best);
            rp[l] = (uint8_t)bestp;
        }
        // a round only reads subsets finished in earlier rounds
        for (int l = 0; l < 32; ++l)
            if (rs[l]) { opt_cost[rs[l]] = rc[l]; opt_part[rs[l]] = rp[l]; }
    }

    // size 6: lane l evaluates one partition that leaves the top bit unset
    for (int e = 0; e < 7; ++e) {
        int s = 0x7f ^ (1 << e);
        float c[32]; int p[32];
        for (int l = 0; l < 32; ++l) {
            // the reference writes ((1<<(5-e))-1) << (e+1); for e = 6 that
            // shift count is negative and the intended term is empty
            const int hm = 5 - e > 0 ? (((1 << (5 - e)) - 1) << (e + 1)) : 0;
            int part = (l & ((1 << e) - 1)) | ((l << 1) & hm);
            p[l] = part;
            c[l] = part ? opt_cost[part] + opt_cost[s ^ part] : -1.f;
        }
        lane_reduce(c, p);
        opt_cost[s] = internal_cost2(area[s], c[0]);
        opt_part[s] = (uint8_t)p[0];
    }
    // size 7: lane l evaluates partitions l and l|32
    {
        const int s = 0x7f;
        float c[32]; int p[32];
        for (int l = 0; l < 32; ++l) {
            p[l] = l;
            c[l] = l ? opt_cost[l] + opt_cost[s ^ l] : -1.f;
            int q = l | 32;
            float oc = opt_cost[q] + opt_cost[s ^ q];
            if (c[l] == -1.f || oc < c[l]) { c[l] = oc; p[l] = q; }
        }
        lane_reduce(c, p);
        opt_cost[s] = internal_cost2(area[s], c[0]);
        opt_part[s] = (uint8_t)p[0];
    }

    // reconstruction: the first proposing lane (non-singleton, not yet
    // expanded) in lane order is split; its two halves go to the next lanes,
    // which keep the node ids they held during formation.
    int lane_subset[T], lane_parent[T], new_left[T], new_right[T];
    bool rexp[T];
    for (int l = 0; l < T; ++l) { lane_subset[l] = -1; lane_parent[l] = -1; rexp[l] = false; new_left[l] = -1; new_right[l] = -1; }
    lane_subset[0] = 0x7f;
    lane_parent[0] = nd.parent[root];
    auto lane_of = [&](int node) { for (int l = 0; l < T; ++l) if (lane_node[l] == node) return l; return -1; };
    int rsize = 1;
    while (rsize < T) {
        int xl = -1;
        for (int l = 0; l < rsize; ++l) {
            int s = lane_subset[l];
            if (!rexp[l] && (s & (s - 1)) != 0) { xl = l; break; }
        }
        if (xl < 0) { std::fprintf(stderr, "treelet: nothing to expand\n"); return -1; }
        rexp[xl] = true;
        int s = lane_subset[xl];
        int pn = lane_node[xl];
        lane_parent[rsize] = pn;
        lane_subset[rsize] = opt_part[s];
        new_left[lane_of(pn)] = lane_node[rsize];
        lane_parent[rsize + 1] = pn;
        lane_subset[rsize + 1] = s ^ opt_part[s];
        new_right[lane_of(pn)] = lane_node[rsize + 1];
        rsize += 2;
    }

    // two phases: every lane reads the original treelet leaves first, then the
    // new node contents are written.
    struct LeafCopy { bool leaf; int prim; Box box; int l, r; };
    LeafCopy cp[T];
    for (int l = 0; l < T; ++l) {
        if (rexp[l]) continue;
        int bit = 0, tmp = lane_subset[l];
        while ((tmp & 1) == 0) { ++bit; tmp >>= 1; }
        int orig = leaves[bit];
        cp[l] = LeafCopy{nd.is_leaf(orig), nd.prim[orig], nd.get_box(orig), nd.left[orig], nd.right[orig]};
    }
    for (int l = 0; l < T; ++l) {
        int me = lane_node[l];
        nd.parent[me] = lane_parent[l];
        if (rexp[l]) {
            nd.left[me] = new_left[l];
            nd.right[me] = new_right[l];
            nd.prim[me] = -1;
        } else {
            nd.left[me] = cp[l].l;
            nd.right[me] = cp[l].r;
            nd.prim[me] = cp[l].leaf ? cp[l].prim : -1;
            nd.set_box(me, cp[l].box);
            nd.sa[me] = surface_area(cp[l].box);
        }
    }
    for (int l = 0; l < T; ++l)
        if (!rexp[l] && !cp[l].leaf) {
            nd.parent[cp[l].l] = lane_node[l];
            nd.parent[cp[l].r] = lane_node[l];
        }
    // the reference's 7 Jacobi sweeps converge to the bottom-up recomputation of
    // box, area and cost over the new treelet (depth <= 6); do that directly.
    std::vector<int> post;
    {
        int st[T * 2], top = 0;
        st[top++] = root;
        std::vector<int> pre;
        while (top) {
            int v = st[--top];
            pre.push_back(v);
            int l = lane_of(v);
            if (l >= 0 && rexp[l]) { st[top++] = nd.left[v]; st[top++] = nd.right[v]; }
        }
        post.assign(pre.rbegin(), pre.rend());
    }
    for (int v : post) {
        int l = lane_of(v);
        if (l < 0) continue;
        if (rexp[l]) {
            Box u = union_box(nd.get_box(nd.left[v]), nd.get_box(nd.right[v]));
            nd.set_box(v, u);
            nd.sa[v] = surface_area(u);
            nd.cost[v] = internal_cost3(nd.sa[v], nd.cost[nd.left[v]], nd.cost[nd.right[v]]);
        } else if (nd.is_leaf(v)) {
            nd.cost[v] = leaf_cost(nd.sa[v], 1);
        } else {
            nd.cost[v] = internal_cost3(nd.sa[v], nd.cost[nd.left[v]], nd.cost[nd.right[v]]);
        }
    }
    return 1;
}

}  // namespace

// BVH/Optimization.cu:467-525: per round, every leaf climbs; the second visitor
// of a node optimises the treelet rooted there. Each node is processed once,
// after both of its subtrees, so a sequential climb gives the parallel result.
extern "C" int ref_treelet_optimize(int n, int rounds, float* box, float* sa, float* cost,
                                    int32_t* left, int32_t* right, int32_t* parent,
                                    int32_t* prim, int64_t* treelets_done) {
    const int64_t m = 2 * (int64_t)n - 1;
    Nodes nd{box, sa, cost, left, right, parent, prim};
    std::vector<uint32_t> visited(m);
    int64_t done = 0;
    for (int r = 0; r < rounds; ++r) {
        std::fill(visited.begin(), visited.end(), 0u);
        // leaf node ids at the start of the round, in id order
        std::vector<int> leaf_ids;
        leaf_ids.reserve(n);
        for (int64_t i = 0; i < m; ++i) if (prim[i] >= 0) leaf_ids.push_back((int)i);
        for (int start : leaf_ids) {
            int cur = start;
            while (true) {
                if (!nd.is_leaf(cur)) {
                    int rc = optimize_treelet(nd, cur);
                    if (rc < 0) return -1;
                    done += rc;
                }
                if (cur == 0) break;
                int p = parent[cur];
                if (visited[p]++ != 1) break;
                cur = p;
            }
        }
    }
    if (treelets_done) *treelets_done = done;
    return 0;
}
