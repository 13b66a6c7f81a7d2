// Layout of the sort's temp buffer and the wave primitives shared by the LSD
// radix sort (radix_sort.hip) and the bucket sort of the step's cell keys
// (bucket_sort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtps {

constexpr int kMaxRadix = 256;
constexpr int kMaxPasses = 4;

// tiles of `tile` keys
static inline int ntiles_for(int n, int tile) { return (n + tile - 1) / tile; }
constexpr int kMinTile = 1024;   // the smallest tile: sizes the status arrays

// Status word of the look-back: flag in bits 31-30, count in bits 29-0.
constexpr uint32_t kFlagAgg = 1u << 30, kFlagPrefix = 2u << 30, kCountMask = (1u << 30) - 1u;
// Copies of each pass's digit totals: histogram block b adds into copy
// b % kCountCopies, so a total receives ntiles / kCountCopies atomics instead
// of ntiles (same-address atomics from every CU serialise); readers add the
// copies up.
constexpr int kCountCopies = 16;
constexpr int kMaxSortKeys = (1 << 30) - 1;
// Buckets of the bucket sort: at most 2^16 (per-tile histograms of that many
// 16-bit counters in LDS).
constexpr int kMaxBuckets = 1 << 16;

struct SortTemp {
    uint32_t* gcount;   // [kMaxPasses][kCountCopies][kMaxRadix], zero between sorts (offset 0: independent of n)
    uint32_t* ctr;      // [kMaxPasses] tile counters, then the bucket sort's block counter and list length
    uint32_t* bhist;    // [kMaxBuckets] bucket sizes, zero between sorts
    uint32_t* bstart;   // [kMaxBuckets] bucket starts
    uint32_t* blist;    // [kMaxBuckets] buckets left to the large-bucket pass
    uint32_t* status;   // [2][status_words]
    long status_words;  // tiles x kMaxRadix for the smallest tile
    uint32_t* kalt;
    uint32_t* valt;
};

// Bytes from the start of the buffer that must be zero before the first sort
// (digit totals, tile counters and bucket sizes; every sort leaves them zero
// again).
constexpr size_t kSortZeroBytes =
    (size_t)kMaxPasses * kCountCopies * kMaxRadix * sizeof(uint32_t) + 256 + (size_t)kMaxBuckets * sizeof(uint32_t);

static inline size_t sort_layout(int n, SortTemp* t, char* p) {
    const long nn = n > 0 ? n : 1;
    size_t off = 0;
    if (t) t->gcount = (uint32_t*)(p + off);
    off += (size_t)kMaxPasses * kCountCopies * kMaxRadix * sizeof(uint32_t);
    if (t) t->ctr = (uint32_t*)(p + off);
    off += 256;
    if (t) t->bhist = (uint32_t*)(p + off);
    off += (size_t)kMaxBuckets * sizeof(uint32_t);
    if (t) t->bstart = (uint32_t*)(p + off);
    off += (size_t)kMaxBuckets * sizeof(uint32_t);
    if (t) t->blist = (uint32_t*)(p + off);
    off += (size_t)kMaxBuckets * sizeof(uint32_t);
    const long words = (long)ntiles_for((int)nn, kMinTile) * kMaxRadix;
    if (t) {
        t->status_words = words;
        t->status = (uint32_t*)(p + off);
    }
    off += 2 * (size_t)words * sizeof(uint32_t);
    off = (off + 255) & ~(size_t)255;
    if (t) {
        t->kalt = (uint32_t*)(p + off);
        t->valt = t->kalt + nn;
    }
    off += 2 * (size_t)nn * sizeof(uint32_t);   // alternate keys / values
    return off;
}

static inline SortTemp carve(void* temp, int n) {
    SortTemp t;
    sort_layout(n, &t, (char*)temp);
    return t;
}

// Lanes of the calling wave whose digit equals this lane's (restricted to `valid` lanes).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, int bits, bool valid) {
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
    }
    return peers;
}

// Inclusive scan of v across the 64 lanes of a wave.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

}  // namespace rtps
