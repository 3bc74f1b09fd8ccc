// Stable LSD radix sort of (key, u32 value) pairs for the LBVH build.
//
// Replaces thrust::stable_sort_by_key at BVH/BVH.cu:335. Keys are 30-bit
// Morton codes in u32 or 63-bit codes in u64; values start as the original
// triangle index, so the output order equals sorting (key, index) pairs.
//
// Per 8-bit digit pass, three kernels:
//   upsweep   : per-tile digit counts (tile = 4096 keys, 256 threads)
//   scan      : one workgroup per digit scans its counts over tiles and adds
//               the digit's global base (from one histogram pass up front)
//   downsweep : wave64 ballot multisplit ranks each key inside its tile (stable
//               in element order), stages the tile in LDS in digit order and
//               writes each digit run contiguously to its global offset.
#include "dev_common.h"
#include "lbvh.h"
#include "radix_sort.h"

namespace cav {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kItems = 16;
constexpr int kTile = kBlock * kItems;
constexpr int kRadix = 256;

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K k, int shift) {
    return (uint32_t)(k >> shift) & (kRadix - 1);
}

// element count: n, or the smaller count held on the device
__device__ __forceinline__ int64_t live_n(int64_t n, const uint32_t* n_dev) {
    if (!n_dev) return n;
    const int64_t d = (int64_t)*n_dev;
    return d < n ? d : n;
}

// One read of the keys: digit histograms for every pass at once.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_hist(const K* __restrict__ keys, int64_t n,
                                                 const uint32_t* __restrict__ n_dev, int passes,
                                                 uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[8 * kRadix];
    n = live_n(n, n_dev);
    for (int i = threadIdx.x; i < passes * kRadix; i += kBlock) h[i] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        K k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&h[p * kRadix + digit_of(k, 8 * p)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * kRadix; i += kBlock)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// counts[d * tiles + t] = number of keys of tile t with digit d.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_upsweep(const K* __restrict__ keys, int64_t n,
                                                    const uint32_t* __restrict__ n_dev, int shift, int tiles,
                                                    uint32_t* __restrict__ counts) {
    __shared__ uint32_t h[kWaves][kRadix];
    n = live_n(n, n_dev);
    const int wave = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&h[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        int64_t i = base + (int64_t)j * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[wave][digit_of(keys[i], shift)], 1u);
    }
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) c += h[w][threadIdx.x];
    counts[(int64_t)threadIdx.x * tiles + blockIdx.x] = c;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive scan of 256 values held one per thread; returns the total too.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* wsum,
                                                       uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t off = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        uint32_t s = wsum[w];
        if (w < wave) off += s;
        total += s;
    }
    __syncthreads();
    return off + inc - v;
}

// Workgroup d: offsets[d * tiles + t] = base(d) + sum_{t' < t} counts[d * tiles + t'].
__global__ __launch_bounds__(kBlock) void k_scan(const uint32_t* __restrict__ counts,
                                                 int tiles, const uint32_t* __restrict__ hist,
                                                 uint32_t* __restrict__ offsets) {
    __shared__ uint32_t wsum[kWaves];
    const int d = blockIdx.x;
    uint32_t total;
    uint32_t below = block_excl_scan256(threadIdx.x < (unsigned)d ? hist[threadIdx.x] : 0u, wsum, total);
    (void)below;
    const uint32_t dbase = total;  // sum of global counts of digits < d
    const int per = (tiles + kBlock - 1) / kBlock;
    const int t0 = threadIdx.x * per;
    const uint32_t* c = counts + (int64_t)d * tiles;
    uint32_t s = 0;
    for (int t = t0; t < t0 + per && t < tiles; ++t) s += c[t];
    uint32_t tot2;
    uint32_t run = block_excl_scan256(s, wsum, tot2) + dbase;
    uint32_t* o = offsets + (int64_t)d * tiles;
    for (int t = t0; t < t0 + per && t < tiles; ++t) {
        o[t] = run;
        run += c[t];
    }
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_downsweep(const K* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      K* __restrict__ kout,
                                                      uint32_t* __restrict__ vout, int64_t n,
                                                      const uint32_t* __restrict__ n_dev, int shift, int tiles,
                                                      const uint32_t* __restrict__ offsets) {
    __shared__ uint32_t wh[kWaves][kRadix];
    __shared__ uint32_t tstart[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t wsum[kWaves];
    __shared__ K skey[kTile];
    __shared__ uint32_t sval[kTile];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t tbase = (int64_t)blockIdx.x * kTile;
    n = live_n(n, n_dev);
    if (tbase >= n) return;  // a tile past the live count (the block is uniform)
    const int tile_n = (int)((n - tbase) < kTile ? (n - tbase) : kTile);

    for (int i = threadIdx.x; i < kWaves * kRadix; i += kBlock) (&wh[0][0])[i] = 0;
    gbase[threadIdx.x] = offsets[(int64_t)threadIdx.x * tiles + blockIdx.x];

    // wave w owns elements [w*64*kItems, (w+1)*64*kItems) of the tile; item j
    // of lane l is element w*64*kItems + j*64 + l, so (w, j, l) is element order.
    K key[kItems];
    uint32_t val[kItems];
    uint32_t rank[kItems];
    const int wbase = wave * 64 * kItems;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        int e = wbase + j * 64 + lane;
        if (e < tile_n) { key[j] = kin[tbase + e]; val[j] = vin[tbase + e]; }
    }
    __syncthreads();

    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int e = wbase + j * 64 + lane;
        const bool valid = e < tile_n;
        const uint32_t d = valid ? digit_of(key[j], shift) : 0u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        uint32_t r = 0;
        if (valid) {
            const uint32_t before = (uint32_t)__popcll(peers & lt);
            const uint32_t base = wh[wave][d];
            if (before == 0) wh[wave][d] = base + (uint32_t)__popcll(peers);
            r = base + before;
        }
        rank[j] = r;
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();

    // per digit: exclusive prefix over waves, then over digits for the tile
    {
        const int d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            uint32_t c = wh[w][d];
            wh[w][d] = run;
            run += c;
        }
        uint32_t tot;
        tstart[d] = block_excl_scan256(run, wsum, tot);
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        const int e = wbase + j * 64 + lane;
        if (e < tile_n) {
            const uint32_t d = digit_of(key[j], shift);
            const uint32_t pos = tstart[d] + wh[wave][d] + rank[j];
            skey[pos] = key[j];
            sval[pos] = val[j];
        }
    }
    __syncthreads();

    for (int i = threadIdx.x; i < tile_n; i += kBlock) {
        const K k = skey[i];
        const uint32_t d = digit_of(k, shift);
        const int64_t o = (int64_t)gbase[d] + (i - (int)tstart[d]);
        kout[o] = k;
        vout[o] = sval[i];
    }
}

__global__ void k_zero(uint32_t* p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// Launches go through ls (a stream, or a graph under construction); mk marks
// stream launches for the build's per-kernel timings.
template <typename K>
hipError_t sort_impl(SortBuffers<K>& b, int64_t n, int key_bits, LaunchSink& ls, Marker* mk) {
    const int passes = (key_bits + 7) / 8;
    const int tiles = (int)((n + kTile - 1) / kTile);
    int hblocks = (int)((n + kBlock * 16 - 1) / (kBlock * 16));
    if (hblocks > 2048) hblocks = 2048;
    if (hblocks < 1) hblocks = 1;
    ls.launch("k_zero", &k_zero, dim3(1), dim3(kBlock), b.hist, passes * kRadix);
    mark(mk, "sort_clear");
    ls.launch("k_hist", &k_hist<K>, dim3(hblocks), dim3(kBlock), b.keys[0], n, b.n_dev, passes, b.hist);
    mark(mk, "sort_hist");
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
        const int shift = 8 * p;
        ls.launch("k_upsweep", &k_upsweep<K>, dim3(tiles), dim3(kBlock), b.keys[cur], n, b.n_dev, shift, tiles,
                  b.counts);
        mark(mk, "sort_upsweep");
        ls.launch("k_scan", &k_scan, dim3(kRadix), dim3(kBlock), b.counts, tiles, b.hist + p * kRadix, b.offsets);
        mark(mk, "sort_scan");
        ls.launch("k_downsweep", &k_downsweep<K>, dim3(tiles), dim3(kBlock), b.keys[cur], b.vals[cur],
                  b.keys[cur ^ 1], b.vals[cur ^ 1], n, b.n_dev, shift, tiles, b.offsets);
        mark(mk, "sort_downsweep");
        cur ^= 1;
    }
    b.result = cur;
    return ls.err;
}

}  // namespace

int64_t sort_tiles(int64_t n) { return (n + kTile - 1) / kTile; }
int sort_radix() { return kRadix; }

hipError_t radix_sort_pairs(SortBuffers<uint32_t>& b, int64_t n, int key_bits, hipStream_t s,
                            Marker* mk) {
    LaunchSink ls(s);
    return sort_impl<uint32_t>(b, n, key_bits, ls, mk);
}
hipError_t radix_sort_pairs(SortBuffers<uint64_t>& b, int64_t n, int key_bits, hipStream_t s,
                            Marker* mk) {
    LaunchSink ls(s);
    return sort_impl<uint64_t>(b, n, key_bits, ls, mk);
}
hipError_t radix_sort_pairs(SortBuffers<uint64_t>& b, int64_t n, int key_bits, LaunchSink& ls) {
    return sort_impl<uint64_t>(b, n, key_bits, ls, nullptr);
}

}  // namespace cav
