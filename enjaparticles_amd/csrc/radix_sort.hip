// Stable LSD radix sort of (u32 key, u32 value) pairs for gfx950, replacing
// the reference's bitonic network (System::bitonic_sort, System.cpp:407-480;
// Bitonic<T>::Sort, BitonicSort.cpp:62-183).  Only the low key_bits bits are
// ordered (the cell keys are < 2^key_bits); digits are at most 8 bits wide and
// the bits are split evenly over ceil(key_bits / 8) passes.
//
// Per pass (R = 2^bits buckets, tiles of 256 x ITEMS keys, ITEMS = 4, 8 or
// 16; g_sort_items picks it, by default 16 from 2^19 keys up and 8 below; the
// tile kernels run 1024 threads of ITEMS / 4 keys each, 16 waves per tile for
// latency hiding):
//   k_radix_hist    per-tile digit histogram, digit-major in hist[d * ntiles + t];
//                   lanes of a wave holding the same digit are found with
//                   `bits` ballots and add their count with one LDS atomic;
//                   digit totals accumulated in gtot[d]
//   k_radix_scan    one block per digit: exclusive scan of its row plus the
//                   total of all smaller digits -> global offsets
//   k_radix_scatter stable ranking inside the tile: wave w owns the tile's
//                   w-th sixteenth and walks it in rounds of 64 keys; the rank
//                   of a key = popcount of the lower lanes with the same digit
//                   (ballots) + the wave's running count of that digit (LDS);
//                   a 16-wave exclusive scan per digit then gives each wave's
//                   offset and the key goes to digit offset + rank.  Block 0
//                   clears gtot for the next use.
// Keys and values ping-pong between the caller's arrays and the alternate
// arrays in the temp buffer.

#include <hip/hip_runtime.h>

#include "rtps_internal.h"

namespace rtps {

constexpr int kSortThreads = 256;   // tile = kSortThreads x ITEMS keys; also the scan kernel's block
constexpr int kSortMinItems = 4;    // smallest tile: sizes the tile-histogram buffer
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kTileThreads = 1024;  // threads of the histogram and scatter kernels (ITEMS / 4 keys each)
constexpr int kTileWaves = kTileThreads / 64;

// 256 x g_sort_items keys per sort tile (4, 8 or 16; 0 = by size: 4096-key
// tiles measured 77 us for 1M keys against 84 us with 2048 and 115 us with
// 1024, 65 against 67 us at 655,360 keys; 2048-key tiles are best at 393,216
// and tie with 1024 at 196,608)
int g_sort_items = 0;
constexpr int kMaxRadix = 256;
constexpr int kMaxPasses = 4;

int sort_passes(int key_bits) { return key_bits <= 0 ? 0 : (key_bits + 7) / 8; }

static inline int ntiles_for(int n, int items) { return (n + kSortThreads * items - 1) / (kSortThreads * items); }

size_t sort_temp_bytes(int n, int key_bits) {
    (void)key_bits;
    const size_t nt = (size_t)ntiles_for(n > 0 ? n : 1, kSortMinItems);
    size_t b = (size_t)kMaxPasses * kMaxRadix * sizeof(uint32_t);   // digit totals per pass (fixed offset 0)
    b += 2 * (size_t)n * sizeof(uint32_t);                           // alternate keys / values
    b = (b + 255) & ~(size_t)255;
    b += (size_t)kMaxRadix * nt * sizeof(uint32_t);                  // tile histograms
    return b;
}

struct SortTemp {
    uint32_t* kalt;
    uint32_t* valt;
    uint32_t* hist;
    uint32_t* gtot;
};

// gtot sits at offset 0 so its (kept zero) contents do not depend on n.
static SortTemp carve(void* temp, int n) {
    char* p = (char*)temp;
    SortTemp t;
    t.gtot = (uint32_t*)p;
    size_t off = (size_t)kMaxPasses * kMaxRadix * sizeof(uint32_t);
    t.kalt = (uint32_t*)(p + off);
    t.valt = t.kalt + n;
    off = (off + 2 * (size_t)n * sizeof(uint32_t) + 255) & ~(size_t)255;
    t.hist = (uint32_t*)(p + off);
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

// IT = keys per thread (tile = kTileThreads x IT keys).
template <int IT>
__global__ __launch_bounds__(kTileThreads) void k_radix_hist(const uint32_t* __restrict__ kin, int n, int shift,
                                                             int bits, uint32_t* __restrict__ hist,
                                                             uint32_t* __restrict__ gtot, int ntiles) {
    __shared__ uint32_t h[kMaxRadix];
    const int R = 1 << bits;
    const uint32_t mask = (uint32_t)R - 1u;
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int d = threadIdx.x; d < R; d += kTileThreads) h[d] = 0;
    __syncthreads();
    const long base = (long)blockIdx.x * (kTileThreads * IT);
    // all IT loads first (index clamped to n - 1, n >= 2: no branch, one wait)
    uint32_t kk[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) kk[k] = kin[min(base + k * kTileThreads + threadIdx.x, (long)n - 1)];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const long e = base + k * kTileThreads + threadIdx.x;
        const bool valid = e < n;
        const uint32_t d = valid ? (kk[k] >> shift) & mask : 0u;
        const uint64_t peers = match_digit(d, bits, valid);
        if (valid && (peers & lt) == 0ull) atomicAdd(&h[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    for (int d = threadIdx.x; d < R; d += kTileThreads) {
        const uint32_t c = h[d];
        hist[(long)d * ntiles + blockIdx.x] = c;
        if (c) atomicAdd(&gtot[d], c);
    }
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

// Exclusive block scan (256 threads); returns the exclusive prefix, total in *tot.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* tot) {
    __shared__ uint32_t ws[kSortWaves];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int q = 0; q < kSortWaves; ++q) {
        uint32_t s = ws[q];
        before += q < w ? s : 0u;
        all += s;
    }
    __syncthreads();
    *tot = all;
    return before + inc - v;
}

__global__ __launch_bounds__(kSortThreads) void k_radix_scan(uint32_t* __restrict__ hist,
                                                             const uint32_t* __restrict__ gtot, int ntiles) {
    const int d = blockIdx.x;
    // base = total count of all smaller digits
    uint32_t part = 0;
    for (int e = threadIdx.x; e < d; e += kSortThreads) part += gtot[e];
    uint32_t base;
    block_excl_scan(part, &base);
    // exclusive scan of row d, ntiles entries, in contiguous chunks per thread
    uint32_t* row = hist + (long)d * ntiles;
    const int c = (ntiles + kSortThreads - 1) / kSortThreads;
    const int lo = threadIdx.x * c;
    const int hi = min(lo + c, ntiles);
    uint32_t s = 0;
    for (int e = lo; e < hi; ++e) s += row[e];
    uint32_t tot;
    uint32_t run = base + block_excl_scan(s, &tot);
    for (int e = lo; e < hi; ++e) {
        uint32_t v = row[e];
        row[e] = run;
        run += v;
    }
}

template <int IT>
__global__ __launch_bounds__(kTileThreads) void k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout,
                                                                uint32_t* __restrict__ vout, int n, int shift,
                                                                int bits, const uint32_t* __restrict__ hist,
                                                                int ntiles, uint32_t* __restrict__ gtot_clear) {
    constexpr int kPerWave = 64 * IT;                  // keys per wave, in IT rounds of 64
    __shared__ uint32_t cnt[kTileWaves][kMaxRadix];    // running count per (wave, digit), then wave offsets
    __shared__ uint32_t goff[kMaxRadix];
    const int R = 1 << bits;
    const uint32_t mask = (uint32_t)R - 1u;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    for (int e = tid; e < kTileWaves * kMaxRadix; e += kTileThreads) (&cnt[0][0])[e] = 0;
    for (int d = tid; d < R; d += kTileThreads) goff[d] = hist[(long)d * ntiles + blockIdx.x];
    if (blockIdx.x == 0)
        for (int d = tid; d < R; d += kTileThreads) gtot_clear[d] = 0;
    __syncthreads();

    const long base = (long)blockIdx.x * (kTileThreads * IT) + (long)w * kPerWave;
    uint32_t key[IT], val[IT], pos[IT];
    // all loads first (index clamped to n - 1, n >= 2: no branch, one wait)
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const long e = min(base + k * 64 + lane, (long)n - 1);
        key[k] = kin[e];
        val[k] = vin[e];
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const long e = base + k * 64 + lane;
        const bool valid = e < n;
        if (!valid) key[k] = 0xFFFFFFFFu;
        const uint32_t d = (key[k] >> shift) & mask;
        const uint64_t peers = match_digit(d, bits, valid);
        const uint32_t before = cnt[w][d];                 // this wave's earlier rounds
        pos[k] = before + (uint32_t)__popcll(peers & lt);
        if (valid && (peers & lt) == 0ull) cnt[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    for (int d = tid; d < R; d += kTileThreads) {          // exclusive scan over the 16 waves
        uint32_t run = 0;
#pragma unroll
        for (int q = 0; q < kTileWaves; ++q) {
            const uint32_t c = cnt[q][d];
            cnt[q][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const long e = base + k * 64 + lane;
        if (e < n) {
            const uint32_t d = (key[k] >> shift) & mask;
            const uint32_t p = goff[d] + cnt[w][d] + pos[k];
            kout[p] = key[k];
            vout[p] = val[k];
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

hipError_t sort_clear_temp(void* temp, int n, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    SortTemp t = carve(temp, n);
    return launch_fill_u32(t.gtot, 0u, (long)kMaxPasses * kMaxRadix, st);
}

// Sorts (keys, vals) by the bits [shift[p], shift[p] + width[p]) of each pass
// p (least significant range first, every width <= 8).  The result ends in
// (keys, vals) or in the temp buffer's alternate arrays, returned through
// out_keys / out_vals.  gtot must be zero on entry (sort_clear_temp once;
// every pass leaves it zero again).
hipError_t sort_pairs_passes(int n, const int* shifts, const int* widths, int npasses, uint32_t* keys,
                             uint32_t* vals, void* temp, hipStream_t st, uint32_t** out_keys, uint32_t** out_vals) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    *out_keys = keys;
    *out_vals = vals;
    if (n <= 1 || npasses == 0) return hipSuccess;
    if (npasses > kMaxPasses) return hipErrorInvalidValue;
    SortTemp t = carve(temp, n);
    const int items = g_sort_items == 4 || g_sort_items == 8 || g_sort_items == 16 ? g_sort_items
                      : (n >= (1 << 19) ? 16 : 8);
    const int ntiles = ntiles_for(n, items);
    uint32_t *ki = keys, *vi = vals, *ko = t.kalt, *vo = t.valt;
    for (int p = 0; p < npasses; ++p) {
        const int bits = widths[p], shift = shifts[p];
        if (bits < 1 || bits > 8) return hipErrorInvalidValue;
        uint32_t* gtot = t.gtot + p * kMaxRadix;
#define RTPS_SORT_PASS(I)                                                                                           \
        k_radix_hist<I / 4><<<ntiles, kTileThreads, 0, st>>>(ki, n, shift, bits, t.hist, gtot, ntiles);               \
        k_radix_scan<<<1 << bits, kSortThreads, 0, st>>>(t.hist, gtot, ntiles);                                      \
        k_radix_scatter<I / 4><<<ntiles, kTileThreads, 0, st>>>(ki, vi, ko, vo, n, shift, bits, t.hist, ntiles, gtot)
        if (items == 4) { RTPS_SORT_PASS(4); }
        else if (items == 8) { RTPS_SORT_PASS(8); }
        else { RTPS_SORT_PASS(16); }
#undef RTPS_SORT_PASS
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        uint32_t* tk = ki; ki = ko; ko = tk;
        uint32_t* tv = vi; vi = vo; vo = tv;
    }
    *out_keys = ki;
    *out_vals = vi;
    return hipSuccess;
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

}  // namespace rtps
