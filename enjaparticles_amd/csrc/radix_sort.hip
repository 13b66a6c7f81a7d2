// Stable LSD radix sort of (u32 key, u32 value) pairs for gfx950, replacing
// the reference's bitonic network (System::bitonic_sort, System.cpp:407-480;
// Bitonic<T>::Sort, BitonicSort.cpp:62-183).  Only the key bits named by the
// passes are ordered (the cell keys are < 2^key_bits); digits are at most 8
// bits wide and the bits are split evenly over ceil(key_bits / 8) passes.
//
// One sort = 1 + P launches (P passes), single pass over the keys per digit:
//   k_radix_hist_all  one read of the keys: the digit histograms of all P
//                     passes at once (lanes of a wave holding the same digit
//                     are found with ballots and add their count with one LDS
//                     atomic; block totals added to gcount[p][d]).  It also
//                     clears the tile counters and the first look-back status
//                     array of the passes.
//   k_onesweep        one launch per pass, 1024 threads per tile of 1024 x IT
//                     keys (IT = 1, 2, 4 or 8).  A tile takes the next tile
//                     number from an atomic counter (so every smaller tile is
//                     already running), ranks its keys stably (wave w owns
//                     the tile's w-th sixteenth, rounds of 64 keys: rank = same
//                     digit in lower lanes by ballots + the wave's running
//                     count + an exclusive scan over the 16 waves), publishes
//                     its per-digit counts and looks back over the earlier
//                     tiles' published counts / inclusive prefixes (decoupled
//                     look-back: status word = 2-bit flag | 30-bit count) for
//                     its digits' global offsets.  The keys are then placed in
//                     digit order in LDS and written out, so each digit's run
//                     of a tile is one contiguous, coalesced write.  Passes
//                     alternate between two status arrays: a pass clears the
//                     one the next pass uses (the pass before it is finished).
// gcount is zero between sorts: the last tile of each pass clears its pass's
// row after its look-back (every tile has read the row by then).
// Keys and values ping-pong between the caller's arrays and the alternate
// arrays in the temp buffer.

#include <hip/hip_runtime.h>

#include "rtps_internal.h"
#include "sort_layout.h"
#include "sph_device.h"

namespace rtps {

constexpr int kHistThreads = 1024;  // threads of the histogram kernel

// One-sweep tile, process-wide: threads * 100 + keys per thread (0 = by
// size: 1024 x 4 from 2^19 keys up, else 1024 x 2)
int g_sort_items = 0;

bool sort_config_valid(int block, int it) {
    switch (block) {
        case 1024: return it == 1 || it == 2 || it == 4 || it == 8;
        case 512: return it == 4 || it == 8;
        case 256: return it == 4 || it == 8 || it == 16;
        default: return false;
    }
}

int sort_passes(int key_bits) { return key_bits <= 0 ? 0 : (key_bits + 7) / 8; }

constexpr int kLookback = 8;   // status words of earlier tiles loaded together in the look-back

size_t sort_temp_bytes(int n, int key_bits) {
    (void)key_bits;
    return sort_layout(n, nullptr, nullptr);
}

struct PassPlan {
    int shift[kMaxPasses];
    int bits[kMaxPasses];
    int npasses;
};

// Digit histograms of every pass in one read of the keys; clears the tile
// counters and the status array of even passes (status0).
template <int IT>
__global__ __launch_bounds__(kHistThreads) void k_radix_hist_all(const uint32_t* __restrict__ kin, int n, PassPlan plan,
                                                                 uint32_t* __restrict__ gcount,
                                                                 uint32_t* __restrict__ ctr,
                                                                 uint32_t* __restrict__ status0, long status_words) {
    __shared__ uint32_t h[kMaxPasses][kMaxRadix];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int e = threadIdx.x; e < kMaxPasses * kMaxRadix; e += kHistThreads) (&h[0][0])[e] = 0;
    if (blockIdx.x == 0 && threadIdx.x < kMaxPasses) ctr[threadIdx.x] = 0;
    for (long e = (long)blockIdx.x * kHistThreads + threadIdx.x; e < status_words; e += (long)gridDim.x * kHistThreads)
        status0[e] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * (kHistThreads * IT);
    uint32_t kk[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) kk[k] = kin[min(base + k * kHistThreads + threadIdx.x, (long)n - 1)];
    for (int p = 0; p < plan.npasses; ++p) {
        const int shift = plan.shift[p], bits = plan.bits[p];
        const uint32_t mask = (1u << bits) - 1u;
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const bool valid = base + k * kHistThreads + threadIdx.x < n;
            const uint32_t d = valid ? (kk[k] >> shift) & mask : 0u;
            const uint64_t peers = match_digit(d, bits, valid);
            if (valid && (peers & lt) == 0ull) atomicAdd(&h[p][d], (uint32_t)__popcll(peers));
        }
    }
    __syncthreads();
    const int copy = (int)(blockIdx.x % kCountCopies);
    for (int e = threadIdx.x; e < plan.npasses * kMaxRadix; e += kHistThreads) {
        const uint32_t c = (&h[0][0])[e];
        const int p = e / kMaxRadix, d = e % kMaxRadix;
        if (c) atomicAdd(&gcount[((size_t)p * kCountCopies + copy) * kMaxRadix + d], c);
    }
}

// The first kernel of the system's step: the keys of hash.cl computed from
// the positions (hash_key; keys past num get pad_key) and written with
// idx = i as the sort's input, start[0..n_clear) cleared (the per-step host
// clear of CellIndices.cpp:59-62), and the digit histograms of every pass
// taken from the keys still in registers (k_radix_hist_all without reading
// the keys back).
template <int IT>
__global__ __launch_bounds__(kHistThreads) void k_hash_hist_all(const float4* __restrict__ pos_u, int num_host,
                                                                const int* __restrict__ d_num, int n,
                                                                uint32_t pad_key, Grid g, int sub_bits,
                                                                uint32_t* __restrict__ keys,
                                                                uint32_t* __restrict__ idx,
                                                                uint32_t* __restrict__ start_clear, long n_clear,
                                                                PassPlan plan, uint32_t* __restrict__ gcount,
                                                                uint32_t* __restrict__ ctr,
                                                                uint32_t* __restrict__ status0, long status_words) {
    __shared__ uint32_t h[kMaxPasses][kMaxRadix];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int e = threadIdx.x; e < kMaxPasses * kMaxRadix; e += kHistThreads) (&h[0][0])[e] = 0;
    if (blockIdx.x == 0 && threadIdx.x < kMaxPasses) ctr[threadIdx.x] = 0;
    const long stride = (long)gridDim.x * kHistThreads;
    for (long e = (long)blockIdx.x * kHistThreads + threadIdx.x; e < status_words; e += stride) status0[e] = 0;
    for (long e = (long)blockIdx.x * kHistThreads + threadIdx.x; e < n_clear; e += stride) start_clear[e] = kEmpty;
    const int num = d_num ? *d_num : num_host;
    __syncthreads();
    const long base = (long)blockIdx.x * (kHistThreads * IT);
    uint32_t kk[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const long e = base + k * kHistThreads + threadIdx.x;
        kk[k] = e < num ? hash_key(pos_u[e], g, sub_bits) : pad_key;
        if (e < n) {
            keys[e] = kk[k];
            idx[e] = (uint32_t)e;
        }
    }
    for (int p = 0; p < plan.npasses; ++p) {
        const int shift = plan.shift[p], bits = plan.bits[p];
        const uint32_t mask = (1u << bits) - 1u;
#pragma unroll
        for (int k = 0; k < IT; ++k) {
            const bool valid = base + k * kHistThreads + threadIdx.x < n;
            const uint32_t d = valid ? (kk[k] >> shift) & mask : 0u;
            const uint64_t peers = match_digit(d, bits, valid);
            if (valid && (peers & lt) == 0ull) atomicAdd(&h[p][d], (uint32_t)__popcll(peers));
        }
    }
    __syncthreads();
    const int copy = (int)(blockIdx.x % kCountCopies);
    for (int e = threadIdx.x; e < plan.npasses * kMaxRadix; e += kHistThreads) {
        const uint32_t c = (&h[0][0])[e];
        const int p = e / kMaxRadix, d = e % kMaxRadix;
        if (c) atomicAdd(&gcount[((size_t)p * kCountCopies + copy) * kMaxRadix + d], c);
    }
}

// Exclusive scan over the kMaxRadix entries of a[] held by threads
// 0..kMaxRadix-1 (the first four waves); every thread of the block calls it.
__device__ __forceinline__ uint32_t digit_excl_scan(uint32_t v, uint32_t* wsum /* kMaxRadix / 64 entries */) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (threadIdx.x < kMaxRadix && lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t before = 0;
#pragma unroll
    for (int q = 0; q < kMaxRadix / 64; ++q) before += q < w ? wsum[q] : 0u;
    __syncthreads();
    return before + inc - v;
}

template <int BLOCK, int IT>
__global__ __launch_bounds__(BLOCK) void k_onesweep(const uint32_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                           int n, int shift, int bits,
                                                           uint32_t* __restrict__ gcount_p,
                                                           uint32_t* __restrict__ ctr_p, uint32_t* status,
                                                           uint32_t* __restrict__ status_clear, long status_words,
                                                           int ntiles) {
    constexpr int kTile = BLOCK * IT;
    constexpr int kTileWaves = BLOCK / 64;
    constexpr int kPerWave = 64 * IT;
    __shared__ uint32_t cnt[kTileWaves][kMaxRadix];   // per-wave running counts, then wave prefixes
    __shared__ uint32_t loc[kMaxRadix];                // tile-local digit offsets
    __shared__ uint32_t gofs[kMaxRadix];               // global start of the tile's run of each digit
    __shared__ uint32_t wsum[kMaxRadix / 64];
    __shared__ uint32_t sk[kTile], sv[kTile];
    __shared__ int s_tile;
    const int R = 1 << bits;
    const uint32_t mask = (uint32_t)R - 1u;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    if (tid == 0) s_tile = (int)atomicAdd(ctr_p, 1u);
    for (int e = tid; e < kTileWaves * kMaxRadix; e += BLOCK) (&cnt[0][0])[e] = 0;
    // the status array of the next pass (the pass before this one is finished)
    for (long e = (long)blockIdx.x * BLOCK + tid; e < status_words; e += (long)gridDim.x * BLOCK)
        status_clear[e] = 0;
    __syncthreads();
    const int t = s_tile;
    const long tbase = (long)t * kTile;
    const long base = tbase + (long)w * kPerWave;
    uint32_t key[IT], val[IT], pos[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const long e = min(base + k * 64 + lane, (long)n - 1);
        key[k] = kin[e];
        val[k] = vin[e];
    }
    // this pass's digit totals (the sum of the copies), staged in LDS before
    // the barrier below, so every tile has read them before it publishes
    // (the last tile clears them)
    if (tid < R) {
        uint32_t c = 0;
#pragma unroll
        for (int q = 0; q < kCountCopies; ++q) c += gcount_p[q * kMaxRadix + tid];
        gofs[tid] = c;
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const bool valid = base + k * 64 + lane < n;
        const uint32_t d = (key[k] >> shift) & mask;
        const uint64_t peers = match_digit(valid ? d : 0u, bits, valid);
        const uint32_t before = valid ? cnt[w][d] : 0u;
        pos[k] = before + (uint32_t)__popcll(peers & lt);
        if (valid && (peers & lt) == 0ull) cnt[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    const uint32_t gc = tid < R ? gofs[tid] : 0u;
    uint32_t total = 0;
    if (tid < R) {   // exclusive scan over the 16 waves; publish the tile's count
        for (int q = 0; q < kTileWaves; ++q) {
            const uint32_t c = cnt[q][tid];
            cnt[q][tid] = total;
            total += c;
        }
        uint32_t* st = status + (long)t * kMaxRadix + tid;
        __hip_atomic_store(st, (t == 0 ? kFlagPrefix : kFlagAgg) | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const uint32_t gstart = digit_excl_scan(gc, wsum);
    const uint32_t lstart = digit_excl_scan(total, wsum);
    if (tid < R) {   // look back over the earlier tiles for this digit
        // kLookback status words in flight at a time: when every tile reaches
        // its look-back at once (one tile per CU), the prefix travels down the
        // chain kLookback tiles per load latency instead of one
        uint32_t excl = 0;
        int j = t - 1;
        while (j >= 0) {
            uint32_t v[kLookback];
#pragma unroll
            for (int k = 0; k < kLookback; ++k)
                v[k] = j - k >= 0 ? __hip_atomic_load(status + (long)(j - k) * kMaxRadix + tid, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : kFlagPrefix;   // before tile 0 (tile 0 is always an inclusive prefix)
            int used = kLookback;
            bool done = false;
#pragma unroll
            for (int k = 0; k < kLookback; ++k) {
                if (done || used < kLookback) continue;
                if ((v[k] & ~kCountMask) == 0u) {   // tile j - k has not published yet: reload from it
                    used = k;
                    continue;
                }
                excl += v[k] & kCountMask;
                if (v[k] & kFlagPrefix) done = true;
            }
            if (done) break;
            if (used < kLookback) __builtin_amdgcn_s_sleep(1);
            j -= used;
        }
        if (t > 0)
            __hip_atomic_store(status + (long)t * kMaxRadix + tid, kFlagPrefix | (excl + total), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        loc[tid] = lstart;
        gofs[tid] = gstart + excl;
    }
    __syncthreads();
    if (t == ntiles - 1 && tid < R)   // every tile has read them (they all published)
        for (int q = 0; q < kCountCopies; ++q) gcount_p[q * kMaxRadix + tid] = 0;
#pragma unroll
    for (int k = 0; k < IT; ++k) {   // the keys in digit order, in LDS
        if (base + k * 64 + lane < n) {
            const uint32_t d = (key[k] >> shift) & mask;
            const uint32_t q = loc[d] + cnt[w][d] + pos[k];
            sk[q] = key[k];
            sv[q] = val[k];
        }
    }
    __syncthreads();
    const int tn = (int)min((long)kTile, (long)n - tbase);
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int q = k * BLOCK + tid;
        if (q < tn) {
            const uint32_t kq = sk[q];
            const uint32_t d = (kq >> shift) & mask;
            const uint32_t o = gofs[d] + (uint32_t)q - loc[d];
            kout[o] = kq;
            vout[o] = sv[q];
        }
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy2_u32(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint32_t* __restrict__ ao,
                            uint32_t* __restrict__ bo, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        ao[i] = a[i];
        bo[i] = b[i];
    }
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, long n, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (n <= 0) return hipSuccess;
    k_fill_u32<<<(int)((n + 255) / 256), 256, 0, st>>>(p, v, n);
    return hipGetLastError();
}

// Zeroes the digit counts (once per temp buffer; every sort leaves them zero).
hipError_t sort_clear_temp(void* temp, int n, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    (void)n;
    return launch_fill_u32((uint32_t*)temp, 0u, (long)(kSortZeroBytes / sizeof(uint32_t)), st);
}

// Sorts (keys, vals) by the bits [shift[p], shift[p] + width[p]) of each pass
// p (least significant range first, every width <= 8).  The result ends in
// (keys, vals) or in the temp buffer's alternate arrays, returned through
// out_keys / out_vals.  The temp buffer's digit counts must be zero on entry
// (sort_clear_temp once; every sort leaves them zero again).
namespace {

// The passes of sort_pairs_passes after a first kernel (`first`, given the
// plan, the temp layout and the histogram grid) that has written the digit
// counts of every pass, cleared the tile counters and status array 0, and
// left the input keys in `keys`.
template <typename First>
hipError_t sort_passes_after(int n, const int* shifts, const int* widths, int npasses, uint32_t* keys, uint32_t* vals,
                             void* temp, hipStream_t st, uint32_t** out_keys, uint32_t** out_vals, First first) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    *out_keys = keys;
    *out_vals = vals;
    if (npasses > kMaxPasses || n > kMaxSortKeys) return hipErrorInvalidValue;
    SortTemp t = carve(temp, n);
    PassPlan plan{};
    plan.npasses = npasses;
    for (int p = 0; p < npasses; ++p) {
        if (widths[p] < 1 || widths[p] > 8 || shifts[p] < 0 || shifts[p] + widths[p] > 32) return hipErrorInvalidValue;
        plan.shift[p] = shifts[p];
        plan.bits[p] = widths[p];
    }
    // one-sweep tile: `block` threads x `it` keys (g_sort_items = block * 100 + it, or 0 = by size)
    int block = 1024, it = n >= (1 << 19) ? 4 : 2;
    if (g_sort_items > 0) {
        block = g_sort_items / 100;
        it = g_sort_items % 100;
    }
    if (!sort_config_valid(block, it)) return hipErrorInvalidValue;
    const int ntiles = ntiles_for(n, block * it);
    const int hist_blocks = ntiles_for(n, kHistThreads * 4);
    uint32_t* st0 = t.status;
    uint32_t* st1 = t.status + t.status_words;
    first(plan, t, hist_blocks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t *ki = keys, *vi = vals, *ko = t.kalt, *vo = t.valt;
    for (int p = 0; p < npasses; ++p) {
        uint32_t* cur = (p & 1) ? st1 : st0;
        uint32_t* nxt = (p & 1) ? st0 : st1;
        const long clear = p + 1 < npasses ? t.status_words : 0;
#define RTPS_SWEEP(B, I)                                                                                            \
        k_onesweep<B, I><<<ntiles, B, 0, st>>>(ki, vi, ko, vo, n, plan.shift[p], plan.bits[p],                      \
                                               t.gcount + (size_t)p * kCountCopies * kMaxRadix, t.ctr + p, cur, nxt, \
                                               clear, ntiles)
        if (block == 1024) {
            if (it == 1) { RTPS_SWEEP(1024, 1); } else if (it == 2) { RTPS_SWEEP(1024, 2); }
            else if (it == 4) { RTPS_SWEEP(1024, 4); } else { RTPS_SWEEP(1024, 8); }
        } else if (block == 512) {
            if (it == 4) { RTPS_SWEEP(512, 4); } else { RTPS_SWEEP(512, 8); }
        } else {
            if (it == 4) { RTPS_SWEEP(256, 4); } else if (it == 8) { RTPS_SWEEP(256, 8); } else { RTPS_SWEEP(256, 16); }
        }
#undef RTPS_SWEEP
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        uint32_t* tk = ki; ki = ko; ko = tk;
        uint32_t* tv = vi; vi = vo; vo = tv;
    }
    *out_keys = ki;
    *out_vals = vi;
    return hipSuccess;
}

}  // namespace

// Sorts (keys, vals) by the bits [shift[p], shift[p] + width[p]) of each pass
// p (least significant range first, every width <= 8).  The result ends in
// (keys, vals) or in the temp buffer's alternate arrays, returned through
// out_keys / out_vals.  The temp buffer's digit counts must be zero on entry
// (sort_clear_temp once; every sort leaves them zero again).
hipError_t sort_pairs_passes(int n, const int* shifts, const int* widths, int npasses, uint32_t* keys,
                             uint32_t* vals, void* temp, hipStream_t st, uint32_t** out_keys, uint32_t** out_vals) {
    *out_keys = keys;
    *out_vals = vals;
    if (n <= 1 || npasses == 0) return hipSuccess;
    return sort_passes_after(n, shifts, widths, npasses, keys, vals, temp, st, out_keys, out_vals,
                             [&](const PassPlan& plan, const SortTemp& t, int hist_blocks) {
                                 k_radix_hist_all<4><<<hist_blocks, kHistThreads, 0, st>>>(
                                     keys, n, plan, t.gcount, t.ctr, t.status, t.status_words);
                             });
}

hipError_t hash_sort_pairs(int n, int key_bits, const float4* pos_u, int num_host, const int* d_num, uint32_t pad_key,
                           const Grid& g, int sub_bits, uint32_t* start_clear, int n_clear, uint32_t* keys,
                           uint32_t* vals, void* temp, hipStream_t st, uint32_t** out_keys, uint32_t** out_vals) {
    *out_keys = keys;
    *out_vals = vals;
    int shifts[kMaxPasses], widths[kMaxPasses];
    const int np = sort_plan(key_bits, shifts, widths);
    if (np < 0) return hipErrorInvalidValue;
    if (n <= 1 || np == 0) {   // nothing to sort: the hash and the clear alone
        return launch_hash(num_host, d_num, n, pad_key, pos_u, keys, vals, g, start_clear, n_clear, sub_bits, st);
    }
    return sort_passes_after(n, shifts, widths, np, keys, vals, temp, st, out_keys, out_vals,
                             [&](const PassPlan& plan, const SortTemp& t, int hist_blocks) {
                                 k_hash_hist_all<4><<<hist_blocks, kHistThreads, 0, st>>>(
                                     pos_u, num_host, d_num, n, pad_key, g, sub_bits, keys, vals, start_clear,
                                     (long)n_clear, plan, t.gcount, t.ctr, t.status, t.status_words);
                             });
}

// Passes over the low key_bits bits, split evenly into digits of <= 8 bits;
// shifts/widths hold kMaxPasses entries.  -1 for more than 32 bits.
int sort_plan(int key_bits, int* shifts, int* widths) {
    const int passes = sort_passes(key_bits);
    if (passes == 0) return 0;
    if (passes > kMaxPasses || key_bits > 32) return -1;
    const int width = (key_bits + passes - 1) / passes;
    int shift = 0;
    for (int p = 0; p < passes; ++p) {
        widths[p] = min(width, key_bits - shift);
        shifts[p] = shift;
        shift += widths[p];
    }
    return passes;
}

hipError_t sort_pairs_pingpong(int n, int key_bits, uint32_t* keys, uint32_t* vals, void* temp, hipStream_t st,
                               uint32_t** out_keys, uint32_t** out_vals) {
    int shifts[kMaxPasses], widths[kMaxPasses];
    const int np = sort_plan(key_bits, shifts, widths);
    if (np < 0) return hipErrorInvalidValue;
    return sort_pairs_passes(n, shifts, widths, np, keys, vals, temp, st, out_keys, out_vals);
}

// In-place variant for the C ABI: copies back when the pass count is odd.
hipError_t sort_pairs_inplace(int n, int key_bits, uint32_t* keys, uint32_t* vals, void* temp, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (n <= 1 || sort_passes(key_bits) == 0) return hipSuccess;
    hipError_t e = sort_clear_temp(temp, n, st);
    if (e != hipSuccess) return e;
    uint32_t *ok, *ov;
    e = sort_pairs_pingpong(n, key_bits, keys, vals, temp, st, &ok, &ov);
    if (e != hipSuccess) return e;
    if (ok != keys) {
        k_copy2_u32<<<(n + 255) / 256, 256, 0, st>>>(ok, ov, keys, vals, n);
        e = hipGetLastError();
    }
    return e;
}

// (ao, bo) may be (a, b): each element is read, then written, by one thread
__global__ void k_copy2_xor_u32(const uint32_t* a, const uint32_t* b, uint32_t* ao, uint32_t* bo, int n,
                                uint32_t mask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        ao[i] = a[i] ^ mask;
        bo[i] = b[i];
    }
}

__global__ void k_xor_u32(uint32_t* __restrict__ a, int n, uint32_t mask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] ^= mask;
}

// Bitonic<T>::Sort's contract (BitonicSort.h:26-28): `batch` independent
// arrays of `length` pairs, read from the src arrays, written sorted to the
// dst arrays (which may be the src arrays), ascending for dir != 0 and
// descending for dir == 0, on all 32 key bits.  A descending sort orders the
// complemented keys ascending; equal keys keep their source order.
hipError_t sort_segments(int batch, int length, int dir, uint32_t* dkeys, uint32_t* dvals, const uint32_t* skeys,
                         const uint32_t* svals, void* temp, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    const uint32_t mask = dir ? 0u : 0xFFFFFFFFu;
    const int blocks = (length + 255) / 256;
    for (int b = 0; b < batch; ++b) {
        const size_t o = (size_t)b * (size_t)length;
        if (skeys + o != dkeys + o || svals + o != dvals + o || mask)
            k_copy2_xor_u32<<<blocks, 256, 0, st>>>(skeys + o, svals + o, dkeys + o, dvals + o, length, mask);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = sort_pairs_inplace(length, 32, dkeys + o, dvals + o, temp, st);
        if (e == hipSuccess && mask) {
            k_xor_u32<<<blocks, 256, 0, st>>>(dkeys + o, length, mask);
            e = hipGetLastError();
        }
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace rtps
