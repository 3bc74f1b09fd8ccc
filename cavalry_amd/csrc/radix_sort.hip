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
// The LBVH build uses the single-kernel-per-pass form further down
// (onesweep), whose digit histograms the Morton kernel accumulates.
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

// ---------------------------------------------------------------- onesweep
// One kernel per digit pass (Merrill & Garland's single-pass radix sort with
// decoupled look-back). Workgroups take tiles in launch order from a counter,
// so every tile a workgroup waits on belongs to a workgroup already running.
// Per tile: rank the keys by digit (the ballot multisplit above), publish the
// tile's per-digit counts, add up predecessors' counts back to the first
// published inclusive prefix, publish this tile's inclusive prefix, and
// scatter the tile through LDS so each digit run is written contiguously
// (the keys, then the values through the same LDS buffer).
// Look-back words: flag << 30 | count (flag 1: this tile's count, 2: count
// of this and all earlier tiles), read and written as agent-scope atomics.
// The global start of every digit comes from the pass's digit histogram,
// which the key producer accumulated beforehand.
// The shape (512 threads x 16 keys, 4 predecessors per look-back step) was
// the fastest of a sweep over 256..1024 threads, 16..32 keys per thread and
// windows of 1..32 predecessors on 2^24 Morton codes: larger windows load
// the look-back words faster than they are published.
constexpr int kOsBlock = 512;
constexpr int kOsWaves = kOsBlock / 64;
constexpr int kOsItems = 16;
constexpr int kOsTile = kOsBlock * kOsItems;
constexpr uint32_t kAgg = 1u << 30, kIncl = 2u << 30, kCountMask = (1u << 30) - 1u;
constexpr int kSpinLimit = 1 << 24;  // bounded wait; exceeding it flags *err (never expected)
constexpr int kLookback = 4;          // predecessors read per look-back step

// Exclusive scan over the workgroup of one value per thread (threads past
// the first 256 pass 0); returns the total too.
__device__ __forceinline__ uint32_t os_excl_scan(uint32_t v, uint32_t* wsum, uint32_t& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t off = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < kOsWaves; ++w) {
        const uint32_t s = wsum[w];
        if (w < wave) off += s;
        total += s;
    }
    __syncthreads();
    return off + inc - v;
}

template <typename K>
__global__ __launch_bounds__(kOsBlock) void k_onesweep(const K* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin,
                                                       K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                       int64_t n, int shift,
                                                       const uint32_t* __restrict__ phist,
                                                       uint32_t* status, uint32_t* tile_ctr,
                                                       uint32_t* err) {
    __shared__ uint32_t wh[kOsWaves][kRadix];
    __shared__ uint32_t tstart[kRadix];
    __shared__ uint32_t gbase[kRadix];
    __shared__ uint32_t wsum[kOsWaves];
    __shared__ uint32_t stile;
    __shared__ K stage[kOsTile];  // the tile's keys in digit order, then its values

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) stile = atomicAdd(tile_ctr, 1u);
    for (int i = threadIdx.x; i < kOsWaves * kRadix; i += kOsBlock) (&wh[0][0])[i] = 0;
    __syncthreads();
    const int tile = (int)stile;
    const int64_t tbase = (int64_t)tile * kOsTile;
    const int tile_n = (int)((n - tbase) < kOsTile ? (n - tbase) : kOsTile);

    // wave w owns elements [w*64*kOsItems, (w+1)*64*kOsItems) of the tile;
    // item j of lane l is element w*64*kOsItems + j*64 + l, so (w, j, l) is
    // element order. Loads go through one uniform base per wave.
    K key[kOsItems];
    uint32_t val[kOsItems];
    uint32_t rank[kOsItems];
    const int wbase = wave * 64 * kOsItems;
    const K* kb = kin + tbase + wbase;
    const uint32_t* vb = vin + tbase + wbase;
    if (tile_n == kOsTile) {
#pragma unroll
        for (int j = 0; j < kOsItems; ++j) {
            key[j] = kb[j * 64 + lane];
            val[j] = vb[j * 64 + lane];
        }
    } else {
#pragma unroll
        for (int j = 0; j < kOsItems; ++j) {
            const bool in = wbase + j * 64 + lane < tile_n;
            key[j] = in ? kb[j * 64 + lane] : (K)0;
            val[j] = in ? vb[j * 64 + lane] : 0u;
        }
    }

    // lanes with equal digits: eight ballots; a lane keeps the lanes that
    // agree with it on every bit (m XNOR its own bit, spread over 64 lanes)
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {
        const bool valid = wbase + j * 64 + lane < tile_n;
        const uint32_t d = valid ? digit_of(key[j], shift) : 0u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const int32_t sb = -(int32_t)((d >> b) & 1u);
            const uint64_t m = __ballot(sb != 0);
            const uint64_t sm = (uint64_t)(uint32_t)sb | ((uint64_t)(uint32_t)sb << 32);
            peers &= ~(m ^ sm);
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

    if (threadIdx.x < kRadix) {
        const int d = threadIdx.x;
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < kOsWaves; ++w) {
            const uint32_t c = wh[w][d];
            wh[w][d] = run;
            run += c;
        }
        tstart[d] = run;  // the tile's count of digit d, until the scan below
        // look-back for digit d: publish, then sum predecessors' counts
        uint32_t* my = status + (int64_t)tile * kRadix + d;
        uint32_t excl = 0;
        if (tile == 0) {
            __hip_atomic_store(my, kIncl | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(my, kAgg | run, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            // predecessors are read kLookback at a time (independent loads in
            // flight), then added newest first up to the first inclusive
            // count or the first one not yet published, which is read again
            int t = tile - 1;
            int spins = 0;
            bool done = false;
            while (!done) {
                uint32_t w[kLookback];
#pragma unroll
                for (int q = 0; q < kLookback; ++q)
                    w[q] = t - q >= 0 ? __hip_atomic_load(status + (int64_t)(t - q) * kRadix + d, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT)
                                      : kIncl;
                int used = 0;
#pragma unroll
                for (int q = 0; q < kLookback; ++q) {
                    const uint32_t f = w[q] & ~kCountMask;
                    if (f == 0u) break;
                    excl += w[q] & kCountMask;
                    used = q + 1;
                    if (f == kIncl) { done = true; break; }
                }
                t -= used;
                if (!done && used == 0) {
                    if (++spins > kSpinLimit) { atomicOr(err, 1u); break; }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __hip_atomic_store(my, kIncl | (excl + run), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        gbase[d] = excl;
    }
    __syncthreads();
    {
        // start of digit d in the output: digits below d, then earlier tiles;
        // in the tile: digits below d
        const bool dt = threadIdx.x < kRadix;
        uint32_t tot;
        const uint32_t below = os_excl_scan(dt ? phist[threadIdx.x] : 0u, wsum, tot);
        const uint32_t ts = os_excl_scan(dt ? tstart[threadIdx.x] : 0u, wsum, tot);
        if (dt) {
            gbase[threadIdx.x] += below;
            tstart[threadIdx.x] = ts;
        }
    }
    __syncthreads();

#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {
        if (wbase + j * 64 + lane < tile_n) {
            const uint32_t d = digit_of(key[j], shift);
            rank[j] += tstart[d] + wh[wave][d];  // the item's position in the tile's digit order
            stage[rank[j]] = key[j];
        }
    }
    __syncthreads();
    uint32_t o[kOsItems];
#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {
        const int i = threadIdx.x + j * kOsBlock;
        if (i < tile_n) {
            const K k = stage[i];
            const uint32_t d = digit_of(k, shift);
            o[j] = gbase[d] + (uint32_t)(i - (int)tstart[d]);
            kout[o[j]] = k;
        }
    }
    __syncthreads();
    uint32_t* sval = reinterpret_cast<uint32_t*>(stage);
#pragma unroll
    for (int j = 0; j < kOsItems; ++j)
        if (wbase + j * 64 + lane < tile_n) sval[rank[j]] = val[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kOsItems; ++j) {
        const int i = threadIdx.x + j * kOsBlock;
        if (i < tile_n) vout[o[j]] = sval[i];
    }
}

template <typename K>
hipError_t onesweep_impl(OnesweepBuffers<K>& b, int64_t n, int key_bits, hipStream_t s, Marker* mk) {
    const int passes = (key_bits + 7) / 8;
    const int64_t tiles = onesweep_tiles(n);
    int cur = 0;
    for (int p = 0; p < passes; ++p) {
        hipLaunchKernelGGL(k_onesweep<K>, dim3((unsigned)tiles), dim3(kOsBlock), 0, s, b.keys[cur], b.vals[cur],
                           b.keys[cur ^ 1], b.vals[cur ^ 1], n, 8 * p, b.hist + p * kRadix,
                           b.status + (int64_t)p * kRadix * tiles, b.tile_ctr + p, b.err);
        mark(mk, "sort_onesweep");
        cur ^= 1;
    }
    b.result = cur;
    return hipGetLastError();
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
int64_t onesweep_tiles(int64_t n) { return (n + kOsTile - 1) / kOsTile; }

hipError_t onesweep_sort_pairs(OnesweepBuffers<uint32_t>& b, int64_t n, int key_bits, hipStream_t s,
                               Marker* mk) {
    return onesweep_impl<uint32_t>(b, n, key_bits, s, mk);
}
hipError_t onesweep_sort_pairs(OnesweepBuffers<uint64_t>& b, int64_t n, int key_bits, hipStream_t s,
                               Marker* mk) {
    return onesweep_impl<uint64_t>(b, n, key_bits, s, mk);
}
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
