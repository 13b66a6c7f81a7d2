// Device-side math of the SPH step for gfx950 (wave64).  Each helper states
// the reference code it restates; the kernels in sph_kernels.hip and the
// fused step kernel share these so every path computes identical values.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtps_c.h"

namespace rtps {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;  // empty-cell start marker (CellIndices.cpp:59-62)
constexpr int kMaxGravitySources = RTPS_MAX_GRAVITY_SOURCES;

// Point attractors, by value as a kernel argument (System.cpp:261-269 buffers).
struct GravitySources {
    int n;
    int pad_[3];
    float4 pos[kMaxGravitySources];
    float mass[kMaxGravitySources];
    float alpha[kMaxGravitySources];
};

// The subset of GridParams the device needs, passed by value.
struct Grid {
    float3 min;
    float3 delta;
    float rx, ry;        // grid_res.x, grid_res.y as floats (the hash is evaluated in float)
    int nb_cells;
    int irx, iry;        // (int) rx, (int) ry: kernel arguments, so row offsets are scalar arithmetic
};

inline Grid make_grid(const rtps_grid_params& g) {
    Grid d;
    d.min = make_float3(g.grid_min.x, g.grid_min.y, g.grid_min.z);
    d.delta = make_float3(g.grid_delta.x, g.grid_delta.y, g.grid_delta.z);
    d.rx = g.grid_res.x;
    d.ry = g.grid_res.y;
    d.nb_cells = g.nb_cells;
    d.irx = (int)d.rx;
    d.iry = (int)d.ry;
    return d;
}

// Saturating float -> int (NaN -> 0), the hardware conversion the reference's
// (int) casts rely on (cl_hash.h:39-42, :113).  Written with explicit guards
// because an out-of-range C++ conversion is undefined.
__device__ __forceinline__ int sat_f2i(float v) {
    if (!(v == v)) return 0;
    if (v >= 2147483648.0f) return 0x7FFFFFFF;
    return (int)fmaxf(v, -2147483648.0f);
}

// calcGridCell (cl_hash.h:33-54): truncation toward zero of (p - min) * delta.
__device__ __forceinline__ int3 grid_cell(float x, float y, float z, const Grid& g) {
    return make_int3(sat_f2i((x - g.min.x) * g.delta.x), sat_f2i((y - g.min.y) * g.delta.y),
                     sat_f2i((z - g.min.z) * g.delta.z));
}

// calcGridHash (cl_hash.h:57-114, no wrap): (gz*res.y + gy)*res.x + gx in float.
__device__ __forceinline__ int grid_hash(int gx, int gy, int gz, const Grid& g) {
#pragma clang fp contract(off)
    float v = ((float)gz * g.ry + (float)gy) * g.rx + (float)gx;
    return sat_f2i(v);
}

// hash.cl:50-157: the linear cell of p, clamped to nb_cells (hash.cl:91-92).
// sub_bits > 0 (system path): the key is (cell << sub_bits) | xq with xq the
// x position inside the cell in 1/2^sub_bits (0 for the clamped cell).
__device__ __forceinline__ uint32_t hash_key(const float4& p, const Grid& g, int sub_bits) {
    const float ux = (p.x - g.min.x) * g.delta.x;
    const int cx = sat_f2i(ux);
    const int cy = sat_f2i((p.y - g.min.y) * g.delta.y);
    const int cz = sat_f2i((p.z - g.min.z) * g.delta.z);
    int h = grid_hash(cx, cy, cz, g);
    const bool out = h > g.nb_cells || h < 0;
    h = h > g.nb_cells ? g.nb_cells : h;   // hash.cl:91
    h = h < 0 ? g.nb_cells : h;            // hash.cl:92
    uint32_t key = (uint32_t)h;
    if (sub_bits) {
        int xq = sat_f2i((ux - (float)cx) * (float)(1 << sub_bits));
        xq = out || h == g.nb_cells ? 0 : min(max(xq, 0), (1 << sub_bits) - 1);
        key = (key << sub_bits) | (uint32_t)xq;
    }
    return key;
}

// Particle count for the kernels after cellindices: the out-of-grid deletion
// path (SPH.cpp:220-250) shrinks num to nc = start[nb_cells] when
// 0 <= nc <= num.  Returns true when the deletion path is active.
__device__ __forceinline__ bool effective_num(int num, const uint32_t* start, int nb_cells, int& out) {
    int nc = (int)start[nb_cells];
    bool del = (nc >= 0 && nc <= num);
    out = del ? nc : num;
    return del;
}

// The three consecutive cells (x-1, x, x+1) of one row of the 27-cell
// neighbourhood (cl_sph_neighbors.h:123-139) occupy one contiguous run of the
// sorted arrays; returns [a, b) of that run, skipping cells with
// hash >= nb_cells (unsigned, :68) and empty cells (:79).  Iterating j from a
// to b visits exactly the particles, and in the order, of the reference's
// cell-by-cell loop.
__device__ __forceinline__ void row_range(int cx, int y, int z, const Grid& g, const uint32_t* __restrict__ start,
                                          const uint32_t* __restrict__ end, uint32_t& a, uint32_t& b) {
    a = kEmpty;
    b = 0;
#pragma unroll
    for (int dx = -1; dx <= 1; ++dx) {
        uint32_t h = (uint32_t)grid_hash(cx + dx, y, z, g);
        if (h < (uint32_t)g.nb_cells) {
            uint32_t s = start[h];
            if (s != kEmpty) {
                a = min(a, s);
                b = max(b, end[h]);
            }
        }
    }
    if (a == kEmpty) { a = 0; b = 0; }
}

// ---------------------------------------------------------------------------
// Windowed rows.  The system sorts by an internal key (cell << kSubBits) | xq,
// xq = the particle's x position inside its cell in 1/16ths, so the cells
// keep the reference's order (keys >> kSubBits are the reference keys) and
// every row of three cells is sorted by x.  Only particles with
// |x_j - x_i| <= sqrt(h^2 - d_yz^2) can be neighbours, where d_yz is the
// distance from particle i to the row's y/z column; those form one contiguous
// sub-run, found by two binary searches on the keys.  Rows farther than h in
// y/z are skipped.  The bounds are widened by one sub-cell (and d^2 by 1e-4
// relative) so rounding never drops a candidate the r <= h test would accept.
constexpr int kSubBits = 4;
constexpr int kSubCells = 1 << kSubBits;

// Windowing inputs of a launch: the sorted internal keys (nullptr = plain
// 27-cell rows) and the grid resolution.
struct Win {
    const uint32_t* keys;
    float rx, ry, rz;
    int remap;     // tile order of the neighbour kernels (remap_tile)
    int block;     // threads per block of the neighbour launches (0 = 256)
    int variant;   // force loop (-1 = the process default)
    // sub-cell starts (cellindices with a table, kSubCells entries per cell):
    // the window bounds are looked up instead of searched; nullptr = search
    const uint32_t* sub = nullptr;
    // per-XCD tile queues (claim_tile): kTileQueues counters and a count of
    // finished blocks, all zero between launches; nullptr = one tile per block
    int* tq = nullptr;
    int ntiles = 0;   // tiles of the launch (set by the launcher)
    // Recorded rows (nullptr = none): for particle i and row r = 3 (dz + 1) +
    // (dy + 1), the three words at rows + 3 (r * rows_stride + i) hold its
    // window [a, b) and the rejection bits (rej_bit) of the window's last
    // min(b - a, 32) candidates; after the 27 rows_stride words, two words per
    // row for each wave of 64 particles (18 per wave, from rows + 27 rows_stride
    // + 18 (i >> 6)) hold the wave's union [A, B) of its live lanes' windows.
    // The density kernel writes them; force loop kLoopStagedBits then reads
    // them instead of looking the windows up, reducing them over the wave and
    // testing those candidates again.
    uint32_t* rows = nullptr;
    uint32_t rows_stride = 0;   // particles per row plane (at least the particle count)
};

// Candidates per group in the neighbour loops (loads in flight per lane).
constexpr int kGroup = 4;

struct WindowCtx {
    float ux, uy, uz;        // continuous cell coordinates (p - min) * delta
    float wy, wz;            // cell widths in y and z (1 / delta)
    float hx;                // h in x cell units (h * delta.x)
    int hc;                  // linear hash of cell (c.x - 1, c.y, c.z): row (dy, dz) starts at
                             // hc + (dz * res_y + dy) * res_x
    bool on;                 // windowing allowed for this particle
};

__device__ __forceinline__ WindowCtx window_ctx(float px, float py, float pz, const Grid& g, int3 c, float h,
                                                const uint32_t* keys, float res_x, float res_y, float res_z) {
    WindowCtx w;
    w.ux = (px - g.min.x) * g.delta.x;
    w.uy = (py - g.min.y) * g.delta.y;
    w.uz = (pz - g.min.z) * g.delta.z;
    w.wy = 1.0f / g.delta.y;
    w.wz = 1.0f / g.delta.z;
    w.hx = h * g.delta.x;
    w.hc = (c.z * g.iry + c.y) * g.irx + (c.x - 1);
    // kept opaque: otherwise the compiler folds hc + (dz * res_y + dy) * res_x
    // back into one multiply per row (a quarter-rate integer multiply)
    asm("" : "+v"(w.hc));
    // inside the grid and away from its x edges (no aliased cells in the row)
    w.on = keys != nullptr && w.ux >= 1.0f && w.uy >= 0.0f && w.uz >= 0.0f && c.x + 1 < (int)res_x &&
           w.uy < res_y && w.uz < res_z;
    return w;
}

__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* __restrict__ k, uint32_t lo, uint32_t hi,
                                                    uint32_t v) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (k[mid] < v) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// row_window with the sub-cell start table, for a particle inside the grid
// away from its x edges (WindowCtx::on) and grids of fewer than 2^23 cells.
// The same [a, b) as the general code below, without branches on loaded
// values: there every lookup waits for the one before it (start, then end,
// then the table entry), here the row's loads are issued together.
//  * the cell hashes are computed in integers: for |values| < 2^24 the float
//    arithmetic of grid_hash is exact, so the two agree;
//  * start/end of all three cells are loaded (index nb_cells, the sentinel
//    entry both arrays hold, stands in for hashes outside the grid) and the
//    unoccupied ones are discarded;
//  * the table entries at both window ends are loaded whether or not their
//    cells are occupied (entries of empty cells may be stale; they are
//    discarded the same way).
// Rows farther than h return before loading anything, with an empty [0, 0).
__device__ __forceinline__ void row_window_table(const WindowCtx& w, int3 c, int dy, int dz, float h2,
                                                 const Grid& g, const uint32_t* __restrict__ start,
                                                 const uint32_t* __restrict__ end,
                                                 const uint32_t* __restrict__ sub, uint32_t& a, uint32_t& b) {
    const float ey = dy < 0 ? (w.uy - (float)c.y) * w.wy : (dy > 0 ? ((float)(c.y + 1) - w.uy) * w.wy : 0.0f);
    const float ez = dz < 0 ? (w.uz - (float)c.z) * w.wz : (dz > 0 ? ((float)(c.z + 1) - w.uz) * w.wz : 0.0f);
    const float d2 = ey * ey + ez * ez;
    if (d2 > h2 * 1.0001f) {
        a = 0;
        b = 0;
        return;
    }
    const uint32_t nbc = (uint32_t)g.nb_cells;
    // the row's first cell: the particle's hc plus a wave-uniform offset
    // (scalar arithmetic on the grid size)
    const int hb = w.hc + (dz * g.iry + dy) * g.irx;
    // base[k] through a 32-bit byte offset (k * 4 < 2^29 here): one address
    // register per offset, shared by the start and end loads
    auto ld = [](const uint32_t* __restrict__ base, uint32_t k) {
        return *(const uint32_t*)((const char*)base + k * 4u);
    };
    // hash of cell k of the row, nb_cells when outside the grid
    const uint32_t h0 = min((uint32_t)hb, nbc), h1 = min((uint32_t)(hb + 1), nbc), h2c = min((uint32_t)(hb + 2), nbc);
    // hardware square root (about 1 ulp): the window is widened by a whole
    // sub-cell on each side, so the rounding never drops a candidate
    const float half = __builtin_amdgcn_sqrtf(fmaxf(h2 - d2, 0.0f)) * (w.hx * __builtin_amdgcn_rsqf(h2));
    const float rel = w.ux - (float)(c.x - 1);
    const int r_lo = max((int)floorf((rel - half) * (float)kSubCells) - 1, 0);
    const int r_hi = min((int)floorf((rel + half) * (float)kSubCells) + 1, 3 * kSubCells - 1);
    // table entry of sub-key r: the row's cells are consecutive, so entry q
    // of cell r >> kSubBits is sub[hb * kSubCells + r]; indexes outside the
    // table are clamped into it (their cells are outside the grid, or r is
    // past the row: such entries are discarded below)
    auto entry = [&](int r) { return ld(sub, min((uint32_t)(hb * kSubCells + r), nbc * kSubCells - 1u)); };
    const uint32_t s0 = ld(start, h0), s1 = ld(start, h1), s2 = ld(start, h2c);
    const uint32_t e0 = ld(end, h0), e1 = ld(end, h1), e2 = ld(end, h2c);
    const uint32_t t_lo = entry(r_lo), t_hi = entry(r_hi + 1);
    const bool o0 = h0 < nbc && s0 != kEmpty, o1 = h1 < nbc && s1 != kEmpty, o2 = h2c < nbc && s2 != kEmpty;
    const uint32_t hi = max(o0 ? e0 : 0u, max(o1 ? e1 : 0u, o2 ? e2 : 0u));
    // first index with sub-key >= r: the table entry when r's cell is
    // occupied, else the start of the next occupied cell, else the row end
    const uint32_t occ = (o0 ? 1u : 0u) | (o1 ? 2u : 0u) | (o2 ? 4u : 0u);
    const uint32_t f2 = o2 ? s2 : hi, f1 = o1 ? s1 : f2;
    auto first_at = [&](int r, uint32_t t) {
        const int co = r >> kSubBits;
        const uint32_t fb = co == 0 ? f1 : (co == 1 ? f2 : hi);
        return ((occ >> co) & 1u) ? t : fb;
    };
    // a row without occupied cells comes out as [0, 0) (hi = f1 = f2 = 0)
    // with no branch, so no load waits for the start values
    a = first_at(r_lo, t_lo);
    b = first_at(r_hi + 1, t_hi);
}

// Row (y, z) = (c.y + dy, c.z + dz) of particle i: [a, b) of its candidates.
// key_prefix: high key bits shared by the whole row (the scene id of batched keys).
__device__ __forceinline__ void row_window(const WindowCtx& w, int3 c, int dy, int dz, float h2, const Grid& g,
                                           const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                                           const uint32_t* __restrict__ keys, uint32_t& a, uint32_t& b,
                                           uint32_t key_prefix = 0u, const uint32_t* __restrict__ sub = nullptr) {
    const bool table = sub != nullptr && w.on;
    if (table && g.nb_cells < (1 << 23)) {
        row_window_table(w, c, dy, dz, h2, g, start, end, sub, a, b);
        return;
    }
    // the row's three cells; with the table their starts are kept for the lookups
    const int hb = grid_hash(c.x - 1, c.y + dy, c.z + dz, g);
    uint32_t cs[3] = {kEmpty, kEmpty, kEmpty}, hc[3] = {0u, 0u, 0u};
    if (table) {
        a = kEmpty;
        b = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const uint32_t hk = (uint32_t)grid_hash(c.x - 1 + k, c.y + dy, c.z + dz, g);
            hc[k] = hk;
            if (hk < (uint32_t)g.nb_cells) {
                const uint32_t sk = start[hk];
                if (sk != kEmpty) {
                    cs[k] = sk;
                    a = min(a, sk);
                    b = max(b, end[hk]);
                }
            }
        }
        if (a == kEmpty) { a = 0; b = 0; }
    } else {
        row_range(c.x, c.y + dy, c.z + dz, g, start, end, a, b);
    }
    if (!w.on || a >= b) return;
    const float ey = dy < 0 ? (w.uy - (float)c.y) * w.wy : (dy > 0 ? ((float)(c.y + 1) - w.uy) * w.wy : 0.0f);
    const float ez = dz < 0 ? (w.uz - (float)c.z) * w.wz : (dz > 0 ? ((float)(c.z + 1) - w.uz) * w.wz : 0.0f);
    const float d2 = ey * ey + ez * ez;
    if (d2 > h2 * 1.0001f) {
        b = a;
        return;
    }
    const float half = __builtin_sqrtf(fmaxf(h2 - d2, 0.0f)) * (w.hx * __builtin_amdgcn_rsqf(h2));
    const float rel = w.ux - (float)(c.x - 1);
    const int r_lo = max((int)floorf((rel - half) * (float)kSubCells) - 1, 0);
    const int r_hi = min((int)floorf((rel + half) * (float)kSubCells) + 1, 3 * kSubCells - 1);
    if (table) {
        // first index with sub-key >= r: the table entry of its cell when that
        // cell holds particles, else the start of the next occupied cell
        const uint32_t row_end = b;
        auto first_at = [&](int r) -> uint32_t {
            const int co = r >> kSubBits;
            if (co >= 3) return row_end;
            if (cs[co] != kEmpty) return sub[(size_t)hc[co] * kSubCells + (r & (kSubCells - 1))];
            for (int k = co + 1; k < 3; ++k)
                if (cs[k] != kEmpty) return cs[k];
            return row_end;
        };
        a = first_at(r_lo);
        b = first_at(r_hi + 1);
        return;
    }
    const uint32_t base = key_prefix | ((uint32_t)hb << kSubBits);
    const uint32_t na = lower_bound_u32(keys, a, b, base + (uint32_t)r_lo);
    b = lower_bound_u32(keys, na, b, base + (uint32_t)r_hi + 1u);
    a = na;
}

// XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs, so
// hardware block b runs on XCD b % 8.  Remapping gives each XCD one
// contiguous range of tiles (neighbouring particles share neighbour rows,
// which then hit the same L2).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    constexpr int X = 8;
    const int q = nb / X, r = nb % X;
    const int x = b % X, k = b / X;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Chunked XCD order: the k-th block an XCD receives takes tile
// (chunk * 8 + xcd) * C + (k % C) with chunk = k / C, so every XCD works on
// runs of C consecutive tiles (shared neighbour rows stay in one L2) while the
// runs of the eight XCDs interleave (dense regions stay spread over all of
// them).  Blocks past the last whole group of 8 C tiles keep their index.
__device__ __forceinline__ int xcd_chunk_tile(int b, int nb, int C) {
    constexpr int X = 8;
    const int F = nb / (X * C) * (X * C);
    if (b >= F) return b;
    const int x = b % X, k = b / X;
    return ((k / C) * X + x) * C + k % C;
}

// Tile of block b under the remap option: 0 = hardware order, 1 = one
// contiguous range per XCD, C >= 2 = chunks of C tiles.
__device__ __forceinline__ int remap_tile(int remap, int b, int nb) {
    return remap == 0 ? b : (remap == 1 ? xcd_tile(b, nb) : xcd_chunk_tile(b, nb, remap));
}

// Per-XCD tile queues of the persistent neighbour launches.  The launch has
// a resident grid; block b runs on XCD b % 8 and takes its tiles from queue
// b % 8, which holds the contiguous range [q T / 8, (q + 1) T / 8) of the T
// tiles, so each XCD's L2 serves one contiguous stretch of the sorted
// particles and the neighbour rows its blocks share.  The n_q blocks of
// queue q take the range's first n_q tiles by their rank b / 8 (no atomic);
// the rest are handed out in order by an atomic counter per queue, each on
// its own 256-byte line.  A block whose queue is drained moves on to the
// next queue (drained queues stay drained), which balances XCDs with more
// neighbours per particle against the others.
constexpr int kTileQueues = 8;
constexpr int kTileQueueStride = 64;                                   // ints between two counters
constexpr int kTileQueueWords = kTileQueues * kTileQueueStride + 1;    // + the finished-block count

__device__ __forceinline__ int tile_queue_lo(int ntiles, int q) { return (int)((long)ntiles * q / kTileQueues); }
__device__ __forceinline__ int tile_queue_blocks(int q) {   // blocks whose home queue is q
    return ((int)gridDim.x + kTileQueues - 1 - q) / kTileQueues;
}

// The block's first tile: the tile of its rank in its home queue's range, or
// -1 when the range is shorter than its blocks (block-uniform, no atomics).
__device__ __forceinline__ int first_tile(int ntiles) {
    const int q = (int)(blockIdx.x % kTileQueues), k = (int)(blockIdx.x / kTileQueues);
    const int lo = tile_queue_lo(ntiles, q), hi = tile_queue_lo(ntiles, q + 1);
    return lo + k < hi ? lo + k : -1;
}

// The next tile of the calling block, or -1 once every queue is drained
// (block-uniform).  q: the block's current queue, initially its home queue
// blockIdx.x % 8 (only thread 0's copy is used).
__device__ __forceinline__ int claim_tile(int* __restrict__ tq, int ntiles, int& q) {
    __shared__ int sh_tile;
    if (threadIdx.x == 0) {
        const int q0 = q;
        int t = -1;
        for (;;) {
            const int lo = tile_queue_lo(ntiles, q) + tile_queue_blocks(q);
            const int hi = tile_queue_lo(ntiles, q + 1);
            if (lo < hi) {
                const int k = atomicAdd(&tq[q * kTileQueueStride], 1);
                if (lo + k < hi) {
                    t = lo + k;
                    break;
                }
            }
            q = (q + 1) % kTileQueues;
            if (q == q0) break;   // all drained
        }
        sh_tile = t;
    }
    __syncthreads();
    const int t = __builtin_amdgcn_readfirstlane(sh_tile);
    __syncthreads();   // every wave has read sh_tile before thread 0 writes the next
    return t;
}

// Called once by every block after its last claim: the last block to get
// here clears the counters for the next launch (all other blocks have made
// their last claim by then: thread 0 waits for each claim's result before
// counting itself finished).
__device__ __forceinline__ void finish_tiles(int* __restrict__ tq) {
    if (threadIdx.x == 0 && atomicAdd(&tq[kTileQueues * kTileQueueStride], 1) == (int)gridDim.x - 1)
        for (int k = 0; k <= kTileQueues; ++k) atomicExch(&tq[k * kTileQueueStride], 0);
}

__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// density.cl:44-78 poly6 term of one candidate: sum += (h^2 - r^2)^3 for
// r <= h, else nothing; t = max(h^2 - r^2, 0), with h^2 - |r|^2 formed by three
// fused multiply-adds (one instruction less than |r|^2 first; the rounding stays
// within the FP32 tolerance of the reference's (h^2 - |r|^2)^3).  The density
// kernels (plain and batched) use this form.
__device__ __forceinline__ float h2_minus_r2(float rx, float ry, float rz, float h2) {
    return fmaf(-rz, rz, fmaf(-ry, ry, fmaf(-rx, rx, h2)));
}
__device__ __forceinline__ void poly6_add_r(float& sum, float rx, float ry, float rz, float h2) {
    const float t = fmaxf(h2_minus_r2(rx, ry, rz, h2), 0.0f);
    sum = fmaf(t * t, t, sum);
}

// Neighbour test of every force loop (plain, batched, adjoint): the pair is
// within h when u = h2_minus_r2 has its sign bit clear (u >= 0: u is never
// -0 for h > 0).  force.cl:64 tests rlen <= h; the two differ only for pairs
// within rounding of r = h, where every pair term of force.cl:73-133 is
// (h - r)^k ~ 0.  The density kernel forms the same u for each candidate,
// so it can record the test for the force kernel (rej_bit).
__device__ __forceinline__ bool within_h(float u) { return (int)__float_as_uint(u) >= 0; }
// Shifts the test of one more candidate into a word of rejection bits:
// (rej << 1) | sign(u), one instruction; after a window of L candidates
// its candidate k (k >= L - 32) sits at bit L - 1 - k.
__device__ __forceinline__ uint32_t rej_bit(uint32_t rej, float u) {
    return __builtin_amdgcn_alignbit(rej, __float_as_uint(u), 31u);
}
// Ones in the low n bits, 1 <= n <= 32.
__device__ __forceinline__ uint32_t low_bits(uint32_t n) { return 0xFFFFFFFFu >> (32u - n); }

// Per-particle pair-sum accumulator of force.cl:43-140 (pt.force, pt.viscosity, pt.xsph).
struct ForceAcc {
    float fx, fy, fz;
    float vx, vy, vz;
    float xx, xy, xz;
};

// One accepted pair (rlen <= h, j != i) of force.cl:62-139.  Self pairs add
// exactly zero in the reference (iej = 0 multiplies a finite value), so they
// are skipped.  r = p_i - p_j (scaled), r2 = |r|^2.
// idj = 1 / dj, the hardware reciprocal (per pair here, or kept with a staged row).
__device__ __forceinline__ void force_pair_r(ForceAcc& acc, float rx, float ry, float rz, float r2, float h, float h2,
                                             float eps, float K, float rho0, float Pi, float di, float4 vi, float4 vj,
                                             float dj, float idj) {
    // force.cl:73, rlen = max(|r|, eps), and 1/rlen from one hardware
    // reciprocal square root (~1 ulp; the reference builds OpenCL's length()
    // without correctly rounded sqrt): r2 clamped to eps^2 gives the same max.
    // max on the bit patterns: both are non-negative floats, whose order is
    // their integer order (fmaxf would also canonicalise both operands)
    const float r2c = __int_as_float(max(__float_as_int(r2), __float_as_int(eps * eps)));
    const float rinv = __builtin_amdgcn_rsqf(r2c);
    float rlen = r2c * rinv;
    float hr = h - rlen;
    float dW = hr * hr * rinv;                      // Wspiky_dr, cl_sph_kernels.h:81-97
    float Pj = K * (dj - rho0);                     // force.cl:84-85
    float kern = dW * (Pi + Pj) * idj;              // :97
    float dvx = vj.x - vi.x, dvy = vj.y - vi.y, dvz = vj.z - vi.z;
    float lv = hr * idj;                            // Wvisc_lapl * idj, :109-112
    acc.vx += dvx * lv;
    acc.vy += dvy * lv;
    acc.vz += dvz * lv;
    float t = h2 - r2;                              // Wpoly6, cl_sph_kernels.h:29-45
    float w6 = t * t * t * fast_rcp(di + dj);       // :129-133
    acc.xx += w6 * dvx;
    acc.xy += w6 * dvy;
    acc.xz += w6 * dvz;
    acc.fx += rx * kern;                            // :136
    acc.fy += ry * kern;
    acc.fz += rz * kern;
}

__device__ __forceinline__ void force_pair(ForceAcc& acc, float rx, float ry, float rz, float r2, float h, float h2,
                                           float eps, float K, float rho0, float Pi, float di, float4 vi, float4 vj,
                                           float dj) {
    force_pair_r(acc, rx, ry, rz, r2, h, h2, eps, K, rho0, Pi, di, vi, vj, dj, fast_rcp(dj));
}

// force.cl:181-185 epilogue: force = m/rho_i * (-1/2 wspiky_d F + mu wvisc_dd V),
// xsph = 2 m wpoly6 X.
__device__ __forceinline__ void force_epilogue(const ForceAcc& acc, float di, const rtps_sph_params& p, float4& F,
                                               float4& X) {
    float a = p.mass * (1.0f / di);
    float cp = -0.5f * p.wspiky_d_coef;
    float cv = p.viscosity * p.wvisc_dd_coef;
    F = make_float4(a * (cp * acc.fx + cv * acc.vx), a * (cp * acc.fy + cv * acc.vy), a * (cp * acc.fz + cv * acc.vz),
                    0.0f);
    float cx = 2.0f * p.mass * p.wpoly6_coef;
    X = make_float4(cx * acc.xx, cx * acc.xy, cx * acc.xz, 0.0f);
}

// gravity.cl:9-36 for one particle: accel += mag * fast_normalize(vect).
// A source of magnitude alpha * mass = 0 adds exactly zero to every
// component (0 times a finite unit vector), so it is skipped (a uniform
// branch: the sources are kernel arguments).
__device__ __forceinline__ void gravity_sources(float4& f, float4 pos, const GravitySources& gs, float scale) {
    for (int s = 0; s < gs.n; ++s) {
        if (gs.alpha[s] * gs.mass[s] == 0.0f) continue;
        float vx = scale * (gs.pos[s].x - pos.x), vy = scale * (gs.pos[s].y - pos.y), vz = scale * (gs.pos[s].z - pos.z);
        float d2 = vx * vx + vy * vy + vz * vz;
        float inv = d2 > 0.0f ? __builtin_amdgcn_rsqf(d2) : 0.0f;
        float mag = gs.alpha[s] * gs.mass[s];
        float cut = 0.5f * scale;
        cut *= cut;
        mag = d2 < cut ? mag / -cut : mag / d2;
        f.x += mag * (vx * inv);
        f.y += mag * (vy * inv);
        f.z += mag * (vz * inv);
    }
}

// One wall of collision_wall.cl:44-107 with calculateRepulsionForce /
// calculateFrictionForce (cl_sph_collision.h:31-73); friction coefficients
// are the hard-coded zeros of collision_wall.cl:61-62.
__device__ __forceinline__ void wall(float diff, float nx, float ny, float nz, float4 v, const rtps_sph_params& p,
                                     float3& rf, float3& ff) {
    if (!(diff > p.EPSILON)) return;
    float ndv = nx * v.x + ny * v.y + nz * v.z;
    float m = p.boundary_stiffness * diff - p.boundary_dampening * ndv;
    rf.x += m * nx;
    rf.y += m * ny;
    rf.z += m * nz;
    float vtx = v.x - v.x * ndv, vty = v.y - v.y * ndv, vtz = v.z - v.z * ndv;   // v_n = v * dot(n, v)
    if ((vtx + vty + vtz) / 3.0f > 0.0f) {
        ff.x += -vtx;
        ff.y += -vty;
        ff.z += -vtz;
    } else {
        ff.x += 0.0f * -vtx;
        ff.y += 0.0f * -vty;
        ff.z += 0.0f * -vtz;
    }
}

// collision_wall.cl:31-119 for one particle (p = unscaled sorted position).
__device__ __forceinline__ void collision_wall(float4& f, float4 p, float4 v, const rtps_sph_params& sp,
                                               const rtps_float4& bmin, const rtps_float4& bmax) {
    const float s = sp.simulation_scale;
    const float px = p.x * s, py = p.y * s, pz = p.z * s;
    const float bd = sp.rest_distance * 0.5f;
    float3 rf = make_float3(0.f, 0.f, 0.f), ff = make_float3(0.f, 0.f, 0.f);
    wall(bd - (pz - bmin.z), 0.f, 0.f, 1.f, v, sp, rf, ff);
    wall(bd - (bmax.z - pz), 0.f, 0.f, -1.f, v, sp, rf, ff);
    wall(bd - (py - bmin.y), 0.f, 1.f, 0.f, v, sp, rf, ff);
    wall(bd - (bmax.y - py), 0.f, -1.f, 0.f, v, sp, rf, ff);
    wall(bd - (px - bmin.x), 1.f, 0.f, 0.f, v, sp, rf, ff);
    wall(bd - (bmax.x - px), -1.f, 0.f, 0.f, v, sp, rf, ff);
    f.x += rf.x + ff.x;
    f.y += rf.y + ff.y;
    f.z += rf.z + ff.z;
}

// leapfrog.cl:30-121 for one particle; returns the new unsorted records.
// scaled = false: the caller knows simulation_scale is exactly 1, so the
// multiplications and divisions by it (identities then) are left out.
__device__ __forceinline__ void leapfrog(float4 p, float4 v, float4 f, float4 xs, const rtps_sph_params& sp, float dt,
                                         float4& pos_out, float4& vel_out, float4& veleval_out, bool scaled = true) {
#pragma clang fp contract(off)
    const float s = sp.simulation_scale;
    if (scaled) { p.x *= s; p.y *= s; p.z *= s; }
    f.x += sp.gravity.x; f.y += sp.gravity.y; f.z += sp.gravity.z;
    float speed = __builtin_sqrtf(f.x * f.x + f.y * f.y + f.z * f.z);
    if (speed > sp.velocity_limit) {
        float k = sp.velocity_limit / speed;
        f.x *= k; f.y *= k; f.z *= k;
    }
    float4 vn = make_float4(v.x + dt * f.x, v.y + dt * f.y, v.z + dt * f.z, v.w);
    const float xf = sp.xsph_factor;
    vn.x += xf * xs.x; vn.y += xf * xs.y; vn.z += xf * xs.z; vn.w += xf * xs.w;
    veleval_out = make_float4(0.5f * (v.x + vn.x), 0.5f * (v.y + vn.y), 0.5f * (v.z + vn.z), 0.5f * (v.w + vn.w));
    p.x += dt * vn.x; p.y += dt * vn.y; p.z += dt * vn.z;
    pos_out = scaled ? make_float4(p.x / s, p.y / s, p.z / s, 1.0f) : make_float4(p.x, p.y, p.z, 1.0f);
    vel_out = vn;
}

// euler.cl:33-78 for one particle (integrator "euler"): gravity is the
// kernel's hard-coded -9.8 on z and the clamp its hard-coded 600 on |f.xyz|
// (the settings' gravity and velocity_limit are not read), the clamp scales
// all four components, f.w is kept, there is no XSPH term and veleval is not
// produced.
__device__ __forceinline__ void euler(float4 p, float4 v, float4 f, const rtps_sph_params& sp, float dt,
                                      float4& pos_out, float4& vel_out) {
#pragma clang fp contract(off)
    const float s = sp.simulation_scale;
    p.x *= s; p.y *= s; p.z *= s;
    f.z += -9.8f;
    float speed = __builtin_sqrtf(f.x * f.x + f.y * f.y + f.z * f.z);
    if (speed > 600.0f) {
        float k = 600.0f / speed;
        f.x *= k; f.y *= k; f.z *= k; f.w *= k;
    }
    v.x += dt * f.x; v.y += dt * f.y; v.z += dt * f.z; v.w += dt * f.w;
    p.x += dt * v.x; p.y += dt * v.y; p.z += dt * v.z;
    pos_out = make_float4(p.x / s, p.y / s, p.z / s, 1.0f);
    vel_out = v;
}

}  // namespace rtps
