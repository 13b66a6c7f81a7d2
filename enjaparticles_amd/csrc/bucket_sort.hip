// Bucket sort of the step's cell keys: the result of hash_sort_pairs (the
// stable sort of (key, particle index) pairs: keys ascending, equal keys in
// index order) in three launches with no look-back between tiles.  It
// replaces the reference's bitonic sort of the hashes (System::bitonic_sort,
// System.cpp:407-480) in the step when the key space has at most kMaxBuckets
// buckets of 2^low keys (a bucket: the keys with key >> low = b; low = key
// bits - 16, at least 8 unless the keys are shorter: buckets of 16 cells of
// 16 sub-cell keys for grids up to 2^20 cells).
//
//   k_bucket_hist   the keys from the positions (hash.cl:50-157, hash_key),
//                   written in particle order, and each tile's histogram of
//                   the bucket numbers in LDS (the lanes of a wave with one
//                   bucket add once); one global atomic per non-empty bucket
//                   of the tile adds the tile's count to the bucket's size and
//                   returns the tile's offset in the bucket, kept in a
//                   [tile][bucket] table.  Clears the cell-start table for
//                   cellindices.
//   k_bucket_place  every block scans the bucket sizes into bucket starts;
//                   each wave takes, per round of 64 consecutive particles
//                   (a "run"), one LDS atomic per bucket the run touches for
//                   the run's place in the tile's part of the bucket and
//                   writes the run's keys of that bucket there in index order
//                   with the particle indices.  No global atomics; the runs of
//                   a bucket land in any order.
//   k_bucket_sort   one wave per bucket of up to kSmallMax keys, then one
//                   block per larger bucket: a counting sort
//                   on the low key bits, then each run of equal keys ordered
//                   by particle index.  A bucket above kBlockMax keys, or with
//                   a run of more than kTieMax equal keys, takes stable digit
//                   passes (particle index, then key) through global memory.
//                   Clears the bucket's size for the next sort.

#include <hip/hip_runtime.h>

#include <algorithm>

#include "rtps_internal.h"
#include "sort_layout.h"
#include "sph_device.h"

namespace rtps {

constexpr int kBkThreads = 1024;           // threads of the histogram and placing kernels
constexpr int kBkItems = 4;                // keys per thread in those kernels
constexpr int kBkTile = kBkThreads * kBkItems;
constexpr int kBucketMaxLow = 8;           // key bits inside a bucket: at most 8 (one counting pass)
constexpr int kBucketBits = 16;            // log2 kMaxBuckets
constexpr int kScanThreads = 256;          // threads of k_bucket_scan
constexpr int kScanTile = 4 * kScanThreads;  // bucket sizes per block of k_bucket_scan
constexpr int kSmallMax = 512;             // buckets of k_bucket_sort: up to 512 keys (8 per lane)
constexpr int kBlockMax = 4096;            // buckets sorted by a block in registers and LDS: up to 4096 keys
constexpr int kSortBlocks = 1024;          // blocks of k_bucket_sort (4 waves per SIMD)
constexpr int kSortWaves = 4;              // buckets (waves) per block of k_bucket_sort
constexpr int kMaxBucketPasses = 7;
constexpr int kTieMax = 256;               // longest run of equal keys the tie fix orders
constexpr int kBitmapDigits = 64;          // digits of a pass of the global-memory path: at most 6 bits

struct BucketPlan {
    int low;                    // key bits inside a bucket
    int nb;                     // buckets
    int npass;                  // digit passes, least significant first: particle index, then key
    int first_key;              // the first pass over key bits
    unsigned long long passes;  // 9 bits per pass: source (bit 0: 0 = particle index, 1 = key), shift (5), width (3)
};

__host__ __device__ __forceinline__ void bucket_pass(unsigned long long passes, int p, int& src, int& sh, int& bits) {
    const uint32_t f = (uint32_t)(passes >> (9 * p)) & 0x1FFu;
    src = (int)(f & 1u);
    sh = (int)((f >> 1) & 31u);
    bits = (int)(f >> 6);
}

// The plan for n keys of key_bits bits; false when the buckets would average
// more than 1024 keys (the LSD sort then runs instead).
static bool bucket_plan(int n, int key_bits, BucketPlan& p) {
    if (n < 2 || key_bits <= 0 || key_bits > 32) return false;
    int low = key_bits - kBucketBits;
    if (low < kBucketMaxLow) low = key_bits < kBucketMaxLow ? key_bits : kBucketMaxLow;
    if (low > kBucketMaxLow) return false;   // more than 2^24 keys: the LSD sort
    const int nb = 1 << (key_bits - low);
    if ((long)nb * 1024 < (long)n) return false;
    p = BucketPlan{};
    p.low = low;
    p.nb = nb;
    int ib = 0;   // bits of the particle index
    while ((1L << ib) < (long)n) ++ib;
    bool ok = true;
    auto split = [&](int src, int total) {
        if (total <= 0) return;
        const int np = (total + 5) / 6, w = (total + np - 1) / np;   // digits of at most 6 bits
        for (int sh = 0; sh < total; sh += w) {
            const int bits = w < total - sh ? w : total - sh;
            if (p.npass >= kMaxBucketPasses) {
                ok = false;
                return;
            }
            p.passes |= (unsigned long long)(src | (sh << 1) | (bits << 6)) << (9 * p.npass);
            ++p.npass;
        }
    };
    split(0, ib);
    p.first_key = p.npass;
    split(1, low);
    return ok;
}

bool bucket_sort_fits(int n, int key_bits) {
    BucketPlan p;
    return bucket_plan(n, key_bits, p);
}

// The sort temp of hash_bucket_sort for every count up to n: the radix
// sort's layout for n, then the [tile][bucket] offset table of the largest
// count the bucket sort takes.
size_t bucket_temp_bytes(int n, int key_bits) {
    const size_t base = sort_layout(n, nullptr, nullptr);
    BucketPlan p;
    if (!bucket_plan(2, key_bits, p)) return base;
    const long m = std::min<long>(n, (long)p.nb * 1024);
    return base + (size_t)ntiles_for((int)m, kBkTile) * (size_t)p.nb * sizeof(uint32_t);
}

__global__ __launch_bounds__(kBkThreads) void k_bucket_hist(const float4* __restrict__ pos_u, int num_host,
                                                            const int* __restrict__ d_num, int n, uint32_t pad_key,
                                                            Grid g, int sub_bits, uint32_t* __restrict__ keys,
                                                            uint32_t* __restrict__ start_clear, long n_clear, int low,
                                                            int nb, uint32_t* __restrict__ bhist,
                                                            uint32_t* __restrict__ btab,
                                                            unsigned long long* __restrict__ scan_status) {
    __shared__ uint32_t h[kMaxBuckets / 2];   // two 16-bit counts per word (a tile has kBkTile keys)
    __shared__ uint32_t touched[kBkTile];     // the tile's buckets, in the order first seen
    __shared__ uint32_t ntouched;
    const int tid = threadIdx.x;
    const int nw = (nb + 1) / 2;
    for (int e = tid; e < nw; e += kBkThreads) h[e] = 0;
    if (tid == 0) ntouched = 0u;
    const long stride = (long)gridDim.x * kBkThreads;
    for (long e = (long)blockIdx.x * kBkThreads + tid; e < n_clear; e += stride) start_clear[e] = kEmpty;
    if (blockIdx.x == 0 && tid < kMaxBuckets / kScanTile) scan_status[tid] = 0ull;
    const int num = d_num ? *d_num : num_host;
    __syncthreads();
    const long base = (long)blockIdx.x * kBkTile;
#pragma unroll
    for (int k = 0; k < kBkItems; ++k) {
        const long e = base + k * kBkThreads + tid;
        const bool valid = e < n;
        const float4 p = pos_u[min(e, (long)n - 1)];   // unconditional: the tile's loads in flight together
        const uint32_t key = valid && e < num ? hash_key(p, g, sub_bits) : pad_key;
        if (valid) keys[e] = key;
        if (valid) {
            const uint32_t b = key >> low, sh = 16u * (b & 1u);
            if (((atomicAdd(&h[b >> 1], 1u << sh) >> sh) & 0xFFFFu) == 0u) touched[atomicAdd(&ntouched, 1u)] = b;
        }
    }
    __syncthreads();
    // one global atomic per bucket the tile touches: the tile's offset in it
    uint32_t* tab = btab + (size_t)blockIdx.x * nb;
    for (uint32_t e = tid; e < ntouched; e += kBkThreads) {
        const uint32_t b = touched[e];
        tab[b] = atomicAdd(&bhist[b], (h[b >> 1] >> (16u * (b & 1u))) & 0xFFFFu);
    }
}

// Bucket starts (exclusive scan of the sizes) and the lists of non-empty
// buckets, the small ones (up to kSmallMax keys) from the front of blist, the
// large ones from the back; counts in nlist[0..1].  Block j scans the sizes
// [j * kScanTile, (j + 1) * kScanTile) (4 consecutive per thread), publishes
// its three totals in one 64-bit status word and adds up the words of every
// earlier block (one load per lane of the first wave: the grid is at most
// 64 blocks, all resident, so the earlier blocks' words come in
// after their local scans, without a chain).  K1 cleared the status words.
__global__ __launch_bounds__(kScanThreads) void k_bucket_scan(int nb, const uint32_t* __restrict__ bhist,
                                                              uint32_t* __restrict__ bstart,
                                                              uint32_t* __restrict__ blist, uint32_t* __restrict__ nlist,
                                                              unsigned long long* status) {
    __shared__ uint32_t wsum[3][kScanThreads / 64];
    __shared__ uint32_t prior[3];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int base = blockIdx.x * kScanTile + tid * 4;
    uint32_t hv[4], sum = 0, nsm = 0, nlg = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t x = bhist[min(base + j, nb - 1)];
        hv[j] = base + j < nb ? x : 0u;
        sum += hv[j];
        nsm += hv[j] > 0u && hv[j] <= (uint32_t)kSmallMax ? 1u : 0u;
        nlg += hv[j] > (uint32_t)kSmallMax ? 1u : 0u;
    }
    const uint32_t i0 = wave_incl_scan(sum), i1 = wave_incl_scan(nsm), i2 = wave_incl_scan(nlg);
    if (lane == 63) {
        wsum[0][w] = i0;
        wsum[1][w] = i1;
        wsum[2][w] = i2;
    }
    __syncthreads();
    uint32_t run = i0 - sum, osm = i1 - nsm, olg = i2 - nlg, t0 = 0, t1 = 0, t2 = 0;
    for (int q = 0; q < kScanThreads / 64; ++q) {
        if (q < w) {
            run += wsum[0][q];
            osm += wsum[1][q];
            olg += wsum[2][q];
        }
        t0 += wsum[0][q];
        t1 += wsum[1][q];
        t2 += wsum[2][q];
    }
    if (w == 0) {
        if (lane == 0)   // totals: sizes (30 bits), small (15), large (15), published flag
            __hip_atomic_store(&status[blockIdx.x],
                               (1ull << 63) | t0 | ((unsigned long long)t1 << 30) | ((unsigned long long)t2 << 45),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        uint64_t sv = 0;
        if (lane < (int)blockIdx.x) {
            do {
                sv = __hip_atomic_load(&status[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } while (!(sv >> 63));
        }
        uint32_t p0 = (uint32_t)(sv & 0x3FFFFFFFu), p1 = (uint32_t)((sv >> 30) & 0x7FFFu),
                 p2 = (uint32_t)((sv >> 45) & 0x7FFFu);
        p0 = wave_incl_scan(p0);
        p1 = wave_incl_scan(p1);
        p2 = wave_incl_scan(p2);
        if (lane == 63) {
            prior[0] = p0;
            prior[1] = p1;
            prior[2] = p2;
        }
    }
    __syncthreads();
    run += prior[0];
    osm += prior[1];
    olg += prior[2];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (base + j < nb) {
            bstart[base + j] = run;
            if (hv[j] > (uint32_t)kSmallMax) blist[nb - 1 - (int)olg++] = (uint32_t)(base + j);
            else if (hv[j] > 0u) blist[osm++] = (uint32_t)(base + j);
        }
        run += hv[j];
    }
    if (blockIdx.x == gridDim.x - 1 && tid == kScanThreads - 1) {
        nlist[0] = osm;
        nlist[1] = olg;
    }
}

__global__ __launch_bounds__(kBkThreads) void k_bucket_place(const uint32_t* __restrict__ keys, int n, int low, int nb,
                                                             const uint32_t* __restrict__ btab,
                                                             const uint32_t* __restrict__ bstart,
                                                             uint32_t* __restrict__ kout,
                                                             uint32_t* __restrict__ vout) {
    __shared__ uint32_t bc[kMaxBuckets / 2];   // the tile's keys placed so far per bucket, two 16-bit counts per word
    const int tid = threadIdx.x;
    const long base = (long)blockIdx.x * kBkTile;
    uint32_t key[kBkItems];
#pragma unroll
    for (int k = 0; k < kBkItems; ++k) {
        const long e = base + k * kBkThreads + tid;
        key[k] = keys[min(e, (long)n - 1)];
    }
    // each key's place in the tile's part of its bucket: an LDS atomic (any
    // order: the bucket sort orders equal keys by particle index)
#pragma unroll
    for (int k = 0; k < kBkItems; ++k)
        if (base + k * kBkThreads + tid < n) bc[(key[k] >> low) >> 1] = 0u;   // the counters of the tile's buckets only
    __syncthreads();
    uint32_t got[kBkItems];
#pragma unroll
    for (int k = 0; k < kBkItems; ++k) {
        got[k] = 0u;
        if (base + k * kBkThreads + tid < n) {
            const uint32_t b = key[k] >> low, sh = 16u * (b & 1u);
            got[k] = (atomicAdd(&bc[b >> 1], 1u << sh) >> sh) & 0xFFFFu;
        }
    }
    const uint32_t* tab = btab + (size_t)blockIdx.x * nb;
#pragma unroll
    for (int k = 0; k < kBkItems; ++k) {
        const long e = base + k * kBkThreads + tid;
        if (e < n) {
            const uint32_t b = key[k] >> low;
            const uint32_t pos = bstart[b] + tab[b] + got[k];
            if (pos < (uint32_t)n) {
                kout[pos] = key[k];
                vout[pos] = (uint32_t)e;
            }
        }
    }
}

// Digit of pass p of an element: of the particle index or of the key.
__device__ __forceinline__ uint32_t bucket_digit(uint32_t k, uint32_t v, int src, int sh, uint32_t mask) {
    return ((src ? k : v) >> sh) & mask;
}

// Per-wave LDS of k_bucket_sort for buckets of up to 64 * SLOTS keys.
template <int SLOTS>
struct BucketLds {
    uint32_t stv[64 * SLOTS];           // particle indices in key order
    uint32_t cnt[1 << kBucketMaxLow];   // per key: count, then start
};

// One wave sorts the S <= 64 * SLOTS elements of a bucket (arriving in any
// order; element i in slot i / 64 of lane i % 64) by (key, particle index):
// a counting sort on the bucket's low key bits (at most 8: the key within the
// bucket) gives each key's run of places; LDS atomics give each element a
// place in its run in any order, through which the particle indices are
// staged by key; each element then takes the place in its run given by the
// run's smaller particle indices and is written out.  A run longer than
// kTieMax sends the bucket through the global-memory passes instead (the
// return value; nothing has been written then).  Every slot is processed (no
// branches on the slot count, so the slots stay in registers).
template <int SLOTS>
__device__ __forceinline__ bool sort_bucket_regs(const uint32_t* __restrict__ ka, const uint32_t* __restrict__ va,
                                                 uint32_t* __restrict__ kb, uint32_t* __restrict__ vb, uint32_t S,
                                                 uint32_t s0, int low, BucketLds<SLOTS>& L) {
    const int lane = threadIdx.x & 63;
    const uint32_t mask = (1u << low) - 1u;
    const int D = 1 << low;
    uint32_t k[SLOTS], v[SLOTS], rk[SLOTS];
#pragma unroll
    for (int r = 0; r < SLOTS; ++r) {
        const uint32_t i = min((uint32_t)(r * 64 + lane), S - 1u);
        k[r] = ka[s0 + i];
        v[r] = va[s0 + i];
    }
    for (int e = lane; e < D; e += 64) L.cnt[e] = 0u;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < SLOTS; ++r) rk[r] = (uint32_t)(r * 64 + lane) < S ? atomicAdd(&L.cnt[k[r] & mask], 1u) : 0u;
    __builtin_amdgcn_wave_barrier();
    {   // counts to starts, 4 keys per lane; the longest run of equal keys
        uint32_t c4[4], sum = 0, longest = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = lane * 4 + j;
            c4[j] = e < D ? L.cnt[e] : 0u;
            sum += c4[j];
            longest = max(longest, c4[j]);
        }
        if (__ballot(longest > (uint32_t)kTieMax)) return true;
        uint32_t run = wave_incl_scan(sum) - sum;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = lane * 4 + j;
            if (e < D) L.cnt[e] = run;
            run += c4[j];
        }
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < SLOTS; ++r)
        if ((uint32_t)(r * 64 + lane) < S) L.stv[L.cnt[k[r] & mask] + rk[r]] = v[r];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int r = 0; r < SLOTS; ++r) {
        if ((uint32_t)(r * 64 + lane) < S) {
            const uint32_t d = k[r] & mask;
            const uint32_t lo = L.cnt[d], hi = (int)d + 1 < D ? L.cnt[d + 1] : S;
            uint32_t to = lo;
            for (uint32_t q = lo; q < hi; ++q) to += L.stv[q] < v[r] ? 1u : 0u;
            kb[s0 + to] = k[r];
            vb[s0 + to] = v[r];
        }
    }
    return false;
}

// Exclusive scan in place of the D <= 64 counts of c (one wave).
__device__ __forceinline__ void wave_excl_scan_counts(uint32_t* c, int D) {
    const int lane = threadIdx.x & 63;
    const uint32_t x = lane < D ? c[lane] : 0u;
    const uint32_t inc = wave_incl_scan(x);
    if (lane < D) c[lane] = inc - x;
    __builtin_amdgcn_wave_barrier();
}

// A bucket of more than 64 * MAXSLOTS keys: the plan's passes through global
// memory, between the bucket's ranges of (ka, va) and (kb, vb), ending in
// (kb, vb).  One wave; c holds at least kBitmapDigits counts.
__device__ void sort_bucket_global(uint32_t* __restrict__ ka, uint32_t* __restrict__ va, uint32_t* __restrict__ kb,
                                   uint32_t* __restrict__ vb, uint32_t S, uint32_t s0, const BucketPlan& plan,
                                   uint32_t* c) {
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t *sk = ka, *sv = va, *dk = kb, *dv = vb;
    if ((plan.npass & 1) == 0) {
        for (uint32_t i = lane; i < S; i += 64) {
            kb[s0 + i] = ka[s0 + i];
            vb[s0 + i] = va[s0 + i];
        }
        __threadfence();   // the wave reads these back below
        sk = kb; sv = vb; dk = ka; dv = va;
    }
    for (int p = 0; p < plan.npass; ++p) {
        int src, sh, bits;
        bucket_pass(plan.passes, p, src, sh, bits);
        const uint32_t mask = (1u << bits) - 1u;
        const int D = 1 << bits;
        if (lane < D) c[lane] = 0u;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i0 = 0; i0 < S; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool valid = i < S;
            const uint32_t d = valid ? bucket_digit(sk[s0 + i], sv[s0 + i], src, sh, mask) : 0u;
            const uint64_t peers = match_digit(d, bits, valid);
            if (valid && (peers & lt) == 0ull) c[d] += (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
        }
        wave_excl_scan_counts(c, D);
        for (uint32_t i0 = 0; i0 < S; i0 += 64) {
            const uint32_t i = i0 + lane;
            const bool valid = i < S;
            const uint32_t kk = valid ? sk[s0 + i] : 0u, vv = valid ? sv[s0 + i] : 0u;
            const uint32_t d = bucket_digit(kk, vv, src, sh, mask);
            const uint64_t peers = match_digit(valid ? d : 0u, bits, valid);
            const uint32_t before = valid ? c[d] : 0u;
            if (valid && (peers & lt) == 0ull) c[d] = before + (uint32_t)__popcll(peers);
            __builtin_amdgcn_wave_barrier();
            const uint32_t pos = before + (uint32_t)__popcll(peers & lt);
            if (valid && pos < S) {
                dk[s0 + pos] = kk;
                dv[s0 + pos] = vv;
            }
        }
        __threadfence();   // this pass's stores before the next pass's loads of them
        uint32_t* t = sk; sk = dk; dk = t;
        t = sv; sv = dv; dv = t;
    }
}

// The counting sort of sort_bucket_regs by the whole block (kSortWaves
// waves) for S <= 64 * kSortWaves * SLOTS keys: element i in slot i / 256 of
// wave (i / 64) % 4, lane i % 64.  Returns true (nothing written) when a run
// of equal keys is longer than kTieMax.
template <int SLOTS>
__device__ __forceinline__ bool sort_bucket_block(const uint32_t* __restrict__ ka, const uint32_t* __restrict__ va,
                                                  uint32_t* __restrict__ kb, uint32_t* __restrict__ vb, uint32_t S,
                                                  uint32_t s0, int low, uint32_t* stv, uint32_t* cnt, int* flag) {
    constexpr int kT = 64 * kSortWaves;
    const int tid = threadIdx.x, lane = tid & 63;
    const uint32_t mask = (1u << low) - 1u;
    const int D = 1 << low;
    uint32_t k[SLOTS], v[SLOTS], rk[SLOTS];
#pragma unroll
    for (int r = 0; r < SLOTS; ++r) {
        const uint32_t i = min((uint32_t)(r * kT + tid), S - 1u);
        k[r] = ka[s0 + i];
        v[r] = va[s0 + i];
    }
    for (int e = tid; e < D; e += kT) cnt[e] = 0u;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SLOTS; ++r) rk[r] = (uint32_t)(r * kT + tid) < S ? atomicAdd(&cnt[k[r] & mask], 1u) : 0u;
    __syncthreads();
    if (tid < 64) {   // counts to starts, 4 keys per lane; the longest run of equal keys
        uint32_t c4[4], sum = 0, longest = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = lane * 4 + j;
            c4[j] = e < D ? cnt[e] : 0u;
            sum += c4[j];
            longest = max(longest, c4[j]);
        }
        uint32_t run = wave_incl_scan(sum) - sum;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int e = lane * 4 + j;
            if (e < D) cnt[e] = run;
            run += c4[j];
        }
        const bool longrun = __ballot(longest > (uint32_t)kTieMax) != 0ull;
        if (lane == 0) *flag = longrun ? 1 : 0;
    }
    __syncthreads();
    if (*flag) return true;
#pragma unroll
    for (int r = 0; r < SLOTS; ++r)
        if ((uint32_t)(r * kT + tid) < S) stv[cnt[k[r] & mask] + rk[r]] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SLOTS; ++r) {
        if ((uint32_t)(r * kT + tid) < S) {
            const uint32_t d = k[r] & mask;
            const uint32_t lo = cnt[d], hi = (int)d + 1 < D ? cnt[d + 1] : S;
            uint32_t to = lo;
            for (uint32_t q = lo; q < hi; ++q) to += stv[q] < v[r] ? 1u : 0u;
            kb[s0 + to] = k[r];
            vb[s0 + to] = v[r];
        }
    }
    return false;
}

// The listed buckets: first the small ones (at most kSmallMax keys), one
// wave each, then the large ones (from the back of the list), one block each:
// up to kBlockMax keys in registers and LDS, above that (or with a run of
// equal keys longer than kTieMax) through global memory by the block's first
// wave.  Grid-stride over the lists; the two phases share the LDS.
__global__ __launch_bounds__(64 * kSortWaves) void k_bucket_sort(uint32_t* __restrict__ ka, uint32_t* __restrict__ va,
                                                                 uint32_t* __restrict__ kb, uint32_t* __restrict__ vb,
                                                                 BucketPlan plan, uint32_t* __restrict__ bhist,
                                                                 const uint32_t* __restrict__ bstart,
                                                                 const uint32_t* __restrict__ blist,
                                                                 const uint32_t* __restrict__ nlist) {
    struct BlockLds {
        uint32_t stv[kBlockMax];
        uint32_t cnt[1 << kBucketMaxLow];
        int flag;
    };
    union Lds {
        BucketLds<kSmallMax / 64> wave[kSortWaves];
        BlockLds block;
    };
    __shared__ Lds lds;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    BucketLds<kSmallMax / 64>& L = lds.wave[w];
    const uint32_t nsmall = nlist[0], nlarge = nlist[1];
    for (uint32_t q = blockIdx.x * kSortWaves + w; q < nsmall; q += gridDim.x * kSortWaves) {
        const uint32_t b = blist[q];
        const uint32_t S = bhist[b], s0 = bstart[b];
        bool again;
        if (S <= 256u) again = sort_bucket_regs<4>(ka, va, kb, vb, S, s0, plan.low, reinterpret_cast<BucketLds<4>&>(L));
        else again = sort_bucket_regs<kSmallMax / 64>(ka, va, kb, vb, S, s0, plan.low, L);
        if (again) sort_bucket_global(ka, va, kb, vb, S, s0, plan, L.cnt);
        if (lane == 0) bhist[b] = 0u;
    }
    if (nlarge == 0u) return;   // uniform: no barrier below is reached
    __syncthreads();   // the waves' LDS before the block's
    BlockLds& B = lds.block;
    for (uint32_t q = blockIdx.x; q < nlarge; q += gridDim.x) {   // block-uniform
        const uint32_t b = blist[plan.nb - 1 - q];
        const uint32_t S = bhist[b], s0 = bstart[b];
        constexpr uint32_t kT = 64 * kSortWaves;
        bool again = true;
        if (S <= 4u * kT) again = sort_bucket_block<4>(ka, va, kb, vb, S, s0, plan.low, B.stv, B.cnt, &B.flag);
        else if (S <= 8u * kT) again = sort_bucket_block<8>(ka, va, kb, vb, S, s0, plan.low, B.stv, B.cnt, &B.flag);
        else if (S <= (uint32_t)kBlockMax)
            again = sort_bucket_block<kBlockMax / kT>(ka, va, kb, vb, S, s0, plan.low, B.stv, B.cnt, &B.flag);
        if (again && threadIdx.x < 64) sort_bucket_global(ka, va, kb, vb, S, s0, plan, B.cnt);
        if (threadIdx.x == 0) bhist[b] = 0u;
        __syncthreads();   // the LDS of this bucket before the next one's
    }
}

hipError_t hash_bucket_sort(int n, int key_bits, const float4* pos_u, int num_host, const int* d_num, uint32_t pad_key,
                            const Grid& g, int sub_bits, uint32_t* start_clear, int n_clear, uint32_t* keys,
                            uint32_t* vals, void* temp, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    BucketPlan plan;
    if (!bucket_plan(n, key_bits, plan)) return hipErrorInvalidValue;
    SortTemp t = carve(temp, n);
    uint32_t* btab = (uint32_t*)((char*)temp + sort_layout(n, nullptr, nullptr));
    uint32_t* nlist = t.ctr + kMaxPasses;   // small / large buckets listed
    // the scan's status words: the radix sort's look-back status array (unused here)
    unsigned long long* scan_status = reinterpret_cast<unsigned long long*>(t.status);
    const int blocks = ntiles_for(n, kBkTile);
    k_bucket_hist<<<blocks, kBkThreads, 0, st>>>(pos_u, num_host, d_num, n, pad_key, g, sub_bits, keys, start_clear,
                                                 (long)n_clear, plan.low, plan.nb, t.bhist, btab, scan_status);
    k_bucket_scan<<<ntiles_for(plan.nb, kScanTile), kScanThreads, 0, st>>>(plan.nb, t.bhist, t.bstart, t.blist, nlist,
                                                                           scan_status);
    k_bucket_place<<<blocks, kBkThreads, 0, st>>>(keys, n, plan.low, plan.nb, btab, t.bstart, t.kalt, t.valt);
    k_bucket_sort<<<kSortBlocks, 64 * kSortWaves, 0, st>>>(t.kalt, t.valt, keys, vals, plan, t.bhist, t.bstart,
                                                           t.blist, nlist);
    return hipGetLastError();
}

}  // namespace rtps
