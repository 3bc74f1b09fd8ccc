// Stable LSD radix sort of (key, u32 value) pairs; see radix_sort.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"

namespace cav {

template <typename K>
struct SortBuffers {
    K* keys[2] = {nullptr, nullptr};        // keys[0] holds the input
    uint32_t* vals[2] = {nullptr, nullptr};  // vals[0] holds the input
    uint32_t* hist = nullptr;                // passes * 256
    uint32_t* counts = nullptr;              // 256 * tiles
    uint32_t* offsets = nullptr;             // 256 * tiles
    const uint32_t* n_dev = nullptr;         // optional: sort only the first min(n, *n_dev) pairs
    int result = 0;                          // index of the buffer holding the output
};

// Onesweep sort (one kernel per 8-bit digit pass, decoupled look-back).
template <typename K>
struct OnesweepBuffers {
    K* keys[2] = {nullptr, nullptr};         // keys[0] holds the input
    uint32_t* vals[2] = {nullptr, nullptr};  // vals[0] holds the input
    const uint32_t* hist = nullptr;          // passes * 256 digit counts of the keys
    uint32_t* status = nullptr;              // passes * 256 * onesweep_tiles look-back words, zero before the sort
    uint32_t* tile_ctr = nullptr;            // passes tile counters, zero before the sort
    uint32_t* err = nullptr;                 // nonzero: a look-back wait exceeded its bound
    int result = 0;                          // index of the buffer holding the output
};

struct Marker;
int64_t sort_tiles(int64_t n);
int64_t onesweep_tiles(int64_t n);  // look-back tiles of the onesweep sort (at most sort_tiles)
int sort_radix();
hipError_t radix_sort_pairs(SortBuffers<uint32_t>& b, int64_t n, int key_bits, hipStream_t s,
                            Marker* mk);
hipError_t radix_sort_pairs(SortBuffers<uint64_t>& b, int64_t n, int key_bits, hipStream_t s,
                            Marker* mk);
hipError_t onesweep_sort_pairs(OnesweepBuffers<uint32_t>& b, int64_t n, int key_bits, hipStream_t s,
                               Marker* mk);
hipError_t onesweep_sort_pairs(OnesweepBuffers<uint64_t>& b, int64_t n, int key_bits, hipStream_t s,
                               Marker* mk);
// launched through a sink (a stream or a graph under construction)
hipError_t radix_sort_pairs(SortBuffers<uint64_t>& b, int64_t n, int key_bits, LaunchSink& ls);

}  // namespace cav
