// Per-stage HIP kernels of the SPH step for gfx950 and their C-ABI launchers
// (hash, cellindices, permute, density, force, gravity, collision, leapfrog),
// plus the fused force+gravity+collision+leapfrog kernel used by the system
// step.  The radix sort lives in radix_sort.hip.
//
// Layout in HBM: one float4 (16 B) per particle for position, velocity and
// veleval; sorted copies are separate arrays; keys/idx are u32; start/end
// are u32[nb_cells + 1].  Kernels run one thread per particle (wave64,
// 256-thread blocks); the neighbour loops walk nine contiguous runs of the
// sorted arrays (one per (y, z) row of the 27-cell block).

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "rtps_internal.h"
#include "sph_device.h"

namespace rtps {

constexpr int kBlock = 256;

static inline int blocks_for(long n, int bs = kBlock) { return (int)((n + bs - 1) / bs); }

// ---------------------------------------------------------------------------
// a2: hash.cl:50-157.  Optionally also clears start[0..n_clear) to kEmpty
// (the per-step host clear of CellIndices.cpp:59-62, moved on device).
// sub_bits > 0 (system path): the stored key is (cell << sub_bits) | xq with
// xq the x position inside the cell in 1/2^sub_bits (see row_window).
__global__ __launch_bounds__(kBlock) void k_hash(int num_host, const int* __restrict__ d_num, int n_pad,
                                                 uint32_t pad_key, const float4* __restrict__ pos_u,
                                                 uint32_t* __restrict__ keys, uint32_t* __restrict__ idx, Grid g,
                                                 uint32_t* __restrict__ start_clear, int n_clear, int sub_bits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_clear) start_clear[i] = kEmpty;
    const int num = d_num ? *d_num : num_host;
    if (i < num) {
        keys[i] = hash_key(pos_u[i], g, sub_bits);
        idx[i] = (uint32_t)i;
    } else if (i < n_pad) {
        keys[i] = pad_key;
        idx[i] = (uint32_t)i;
    }
}

// a4: cell start/end by boundary detection on the sorted keys (semantics of
// cellindices.cl:12-107: start[k] = first i with key k, end[k] = one past the
// last; keys > nb_cells are ignored).
__device__ __forceinline__ void cellindices_one(int i, int num, const uint32_t* __restrict__ keys, uint32_t nb_cells,
                                                uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                int sub_bits, uint32_t* __restrict__ sub) {
    uint32_t h = keys[i] >> sub_bits;
    if (h > nb_cells) return;
    if (sub && h < nb_cells) {
        // sub-cell starts of an occupied cell: entry q = first index of the
        // cell with in-cell x sixteenth >= q (the cell's end when none); each
        // entry is written once, by the particle that starts it
        const uint32_t xq = keys[i] & (kSubCells - 1);
        uint32_t q0 = 0;
        if (i > 0 && (keys[i - 1] >> sub_bits) == h) q0 = (keys[i - 1] & (kSubCells - 1)) + 1;
        uint32_t* row = sub + (size_t)h * kSubCells;
        for (uint32_t q = q0; q <= xq; ++q) row[q] = (uint32_t)i;
        if (i == num - 1 || (keys[i + 1] >> sub_bits) != h)
            for (uint32_t q = xq + 1; q < (uint32_t)kSubCells; ++q) row[q] = (uint32_t)i + 1;
    }
    if (i == 0) {
        start[h] = 0;
    } else {
        uint32_t prev = min(keys[i - 1] >> sub_bits, nb_cells);
        if (prev != h) {
            start[h] = (uint32_t)i;
            end[prev] = (uint32_t)i;
        }
    }
    if (i == num - 1) end[h] = (uint32_t)num;
}

__global__ __launch_bounds__(kBlock) void k_cellindices(int num_host, const int* __restrict__ d_num,
                                                        const uint32_t* __restrict__ keys, uint32_t nb_cells,
                                                        uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                        int sub_bits, uint32_t* __restrict__ sub) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int num = d_num ? *d_num : num_host;
    if (i >= num) return;
    cellindices_one(i, num, keys, nb_cells, start, end, sub_bits, sub);
}

__global__ void k_read_nc(const uint32_t* __restrict__ start, int nb_cells, int* __restrict__ nc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *nc = (int)start[nb_cells];
}

// a5: permute.cl:10-46 (pos, vel, veleval).
__device__ __forceinline__ void permute_one(int i, const uint32_t* __restrict__ idx, const float4* __restrict__ pu,
                                            const float4* __restrict__ vu, const float4* __restrict__ eu,
                                            float4* __restrict__ ps, float4* __restrict__ vs, float4* __restrict__ es,
                                            const float4* __restrict__ cu, float4* __restrict__ cs) {
    const uint32_t s = idx[i];
    ps[i] = pu[s];
    vs[i] = vu[s];
    es[i] = eu[s];
    if (cu) cs[i] = cu[s];   // permute.cl:41 (colours, when the system keeps them)
}

__global__ __launch_bounds__(kBlock) void k_permute(int num_host, const int* __restrict__ d_num,
                                                    const uint32_t* __restrict__ idx, const float4* __restrict__ pu,
                                                    const float4* __restrict__ vu, const float4* __restrict__ eu,
                                                    float4* __restrict__ ps, float4* __restrict__ vs,
                                                    float4* __restrict__ es, const float4* __restrict__ cu,
                                                    float4* __restrict__ cs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int num = d_num ? *d_num : num_host;
    if (i >= num) return;
    permute_one(i, idx, pu, vu, eu, ps, vs, es, cu, cs);
}

// a4 + a5 in one pass over the sorted order (the system's step): both read
// the sorted arrays at i and neither reads what the other writes.
__global__ __launch_bounds__(kBlock) void k_cells_permute(int num_host, const int* __restrict__ d_num,
                                                          const uint32_t* __restrict__ keys, uint32_t nb_cells,
                                                          uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                          int sub_bits, uint32_t* __restrict__ sub,
                                                          const uint32_t* __restrict__ idx,
                                                          const float4* __restrict__ pu, const float4* __restrict__ vu,
                                                          const float4* __restrict__ eu, float4* __restrict__ ps,
                                                          float4* __restrict__ vs, float4* __restrict__ es,
                                                          const float4* __restrict__ cu, float4* __restrict__ cs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int num = d_num ? *d_num : num_host;
    if (i >= num) return;
    permute_one(i, idx, pu, vu, eu, ps, vs, es, cu, cs);
    cellindices_one(i, num, keys, nb_cells, start, end, sub_bits, sub);
}

// ---------------------------------------------------------------------------
// a7: density.cl:86-126.  PACKED: the density is written to veleval_s[i].w
// (the system layout, so the force loop loads one float4 per neighbour).
// candidates per density group, their loads issued together
constexpr int kGroupDensity = 4;
// STAGE: each wave copies the union of its lanes' windows of a row (when at
// most kStageCap candidates) into its own LDS slice with full-line loads, and
// the lanes read their candidates from there; wider unions (waves straddling
// two rows of cells) read global memory as before.  Same candidates in the
// same order: the same sums.
constexpr int kStageCap = 256;
int g_density_stage = 1;

// At most 56 VGPRs asked for (nine waves per SIMD): the row-by-row walk fits
// (55; 57 and eight waves without the cap); the staged walk takes 63 anyway
// (eight waves, whose LDS slices take 128 of the 160 KB).
// EMIT: each row's window and the rejection bits of its last 32 candidates
// go to win.rows (one rej_bit per candidate, the same u as the poly6 term),
// for every live particle and all nine rows (empty rows as [a, a)).
template <bool SCALED, bool PACKED, bool STAGE, bool EMIT>
__device__ __forceinline__ void density_tile(int tile, int num_host, const int* __restrict__ d_num,
                                             const float4* __restrict__ pos_s, const uint32_t* __restrict__ start,
                                             const uint32_t* __restrict__ end, const rtps_sph_params& sp,
                                             const Grid& g, const Win& win, float* __restrict__ dens,
                                             float4* __restrict__ packed) {
    const int i = tile * (int)blockDim.x + (int)threadIdx.x;
    int num = d_num ? *d_num : num_host;
    if (d_num) effective_num(num, start, g.nb_cells, num);
    if (i >= num) return;
    const float scale = sp.simulation_scale;
    const float h = sp.smoothing_distance, h2 = h * h;
    float4 pi = pos_s[i];
    if (SCALED) { pi.x *= scale; pi.y *= scale; pi.z *= scale; }
    const int3 c = grid_cell(pi.x, pi.y, pi.z, g);
    const WindowCtx wc = window_ctx(pi.x, pi.y, pi.z, g, c, h, win.keys, win.rx, win.ry, win.rz);
    float sum = 0.0f;
    uint32_t rej = 0u;   // EMIT: the current row's rejection bits
    auto accum = [&](float4 pj) {   // the poly6 term of one candidate
        if (SCALED) { pj.x *= scale; pj.y *= scale; pj.z *= scale; }
        const float u = h2_minus_r2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, h2);
        const float t = fmaxf(u, 0.0f);
        sum = fmaf(t * t, t, sum);
        if (EMIT) rej = rej_bit(rej, u);
    };
    // candidates [a, b) of one row at P[j - off], kGroupDensity loads at a
    // time; one pointer steps through the window (immediate offsets for the
    // group's loads) under a count of the groups left
    auto walk = [&](const float4* __restrict__ P, uint32_t a, uint32_t b, uint32_t off) {
        const float4* __restrict__ pg = P + (a - off);
        const uint32_t n = a < b ? b - a : 0u;
        for (uint32_t g = n / kGroupDensity; g != 0u; --g, pg += kGroupDensity) {
            float4 p[kGroupDensity];
#pragma unroll
            for (int k = 0; k < kGroupDensity; ++k) p[k] = pg[k];
#pragma unroll
            for (int k = 0; k < kGroupDensity; ++k) accum(p[k]);
        }
        for (uint32_t t = n % kGroupDensity; t != 0u; --t, ++pg) accum(*pg);
    };
    uint32_t* __restrict__ rec = EMIT ? win.rows + 3 * (size_t)i : nullptr;
    // the wave's row unions [A, B) follow the particles' records, two words
    // per row for each wave of 64 particles (written by its first live lane)
    uint32_t* __restrict__ wrec = EMIT ? win.rows + 27 * (size_t)win.rows_stride + 18 * (size_t)(i >> 6) : nullptr;
    const uint64_t live = __ballot(true);   // lanes past num have returned
    const uint32_t nlive = (uint32_t)__popcll(live);
    const uint32_t rank = (uint32_t)__popcll(live & ((1ull << (threadIdx.x & 63)) - 1ull));
    auto record = [&](int dz, int dy, uint32_t a, uint32_t b, uint32_t A, uint32_t B) {
        if (EMIT) {
            const int r = 3 * (dz + 1) + dy + 1;
            uint32_t* __restrict__ o = rec + 3 * ((size_t)r * win.rows_stride);
            o[0] = a;
            o[1] = b;
            o[2] = rej;
            rej = 0u;
            if (rank == 0) {
                wrec[2 * r] = A;
                wrec[2 * r + 1] = B;
            }
        }
    };
    if constexpr (STAGE) {
        extern __shared__ float4 stage_rows[];   // kStageCap per wave
        float4* __restrict__ my = stage_rows + (threadIdx.x >> 6) * kStageCap;
        // the wave's live lanes (rank of nlive) share the copies
        for (int dz = -1; dz <= 1; ++dz) {
            for (int dy = -1; dy <= 1; ++dy) {
                uint32_t a, b;
                row_window(wc, c, dy, dz, h2, g, start, end, win.keys, a, b, 0u, win.sub);
                const uint32_t A = __ockl_wfred_min_u32(a < b ? a : 0xFFFFFFFFu);
                const uint32_t B = __ockl_wfred_max_u32(a < b ? b : 0u);
                if (B > A) {
                    if (B - A <= (uint32_t)kStageCap) {
                        for (uint32_t t = rank; t < B - A; t += nlive) my[t] = pos_s[A + t];
                        __builtin_amdgcn_wave_barrier();   // the copies before any lane reads them
                        walk(my, a, b, A);
                        __builtin_amdgcn_wave_barrier();   // every read before the next row's copies
                    } else {
                        walk(pos_s, a, b, 0u);
                    }
                }
                record(dz, dy, a, b, A, B);
            }
        }
    } else {
        for (int dz = -1; dz <= 1; ++dz) {
            for (int dy = -1; dy <= 1; ++dy) {
                uint32_t a, b;
                row_window(wc, c, dy, dz, h2, g, start, end, win.keys, a, b, 0u, win.sub);
                walk(pos_s, a, b, 0u);
                if (EMIT)
                    record(dz, dy, a, b, __ockl_wfred_min_u32(a < b ? a : 0xFFFFFFFFu),
                           __ockl_wfred_max_u32(a < b ? b : 0u));
            }
        }
    }
    const float rho = sp.wpoly6_coef * sp.mass * sum;
    if (PACKED) packed[i].w = rho; else dens[i] = rho;
}

// One tile of blockDim.x sorted particles per block, or (QUEUED) a resident
// grid working through the per-XCD tile queues of win.tq.  The queued
// kernels are separate instances: the loop around the body raises its
// register demand, which must not cost the one-tile kernels.
template <bool SCALED, bool PACKED, bool STAGE, bool QUEUED, bool EMIT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_num_vgpr(56))) void k_density(int num_host, const int* __restrict__ d_num,
                                                    const float4* __restrict__ pos_s,
                                                    const uint32_t* __restrict__ start,
                                                    const uint32_t* __restrict__ end, rtps_sph_params sp, Grid g,
                                                    Win win, float* __restrict__ dens, float4* __restrict__ packed) {
    if constexpr (QUEUED) {
        int q = (int)(blockIdx.x % kTileQueues);
        for (int t = first_tile(win.ntiles);;) {
            if (t < 0) t = claim_tile(win.tq, win.ntiles, q);
            if (t < 0) break;
            density_tile<SCALED, PACKED, STAGE, EMIT>(t, num_host, d_num, pos_s, start, end, sp, g, win, dens, packed);
            t = -1;
        }
        finish_tiles(win.tq);
    } else {
        density_tile<SCALED, PACKED, STAGE, EMIT>(remap_tile(win.remap, (int)blockIdx.x, (int)gridDim.x), num_host, d_num,
                                            pos_s, start, end, sp, g, win, dens, packed);
    }
}

// ---------------------------------------------------------------------------
// a8 (+a9, a10, a11 when FUSED): force.cl:148-186.  PACKED: veleval and
// density come from one float4 (xyz = veleval, w = density).  FUSED: the
// epilogue goes on to the point sources (gravity.cl), the walls
// (collision_wall.cl) and the leapfrog step (leapfrog.cl), scattering to the
// unsorted arrays; force/xsph are then stored only if force_out != nullptr.
//
// LOOP (the force-loop variant):
// kLoopDirect: each lane walks its own window of every row in global memory,
// kGroup candidates per iteration (position loads, then the veleval loads of
// the accepted ones, issued together) with the pair terms under a branch on
// r <= h (within_h).
// kLoopStaged: the direct loop, with each row's candidates (positions, their
// w replaced by 1 / density, and veleval|density) copied once per wave into
// LDS as for the density kernel's STAGE (at most kForceStageCap candidates;
// wider unions read global memory).  Same pairs, same order, same
// reciprocals: the direct loop's bits.
// kLoopStagedBits (the default): the staged loop with the window taken in
// chunks of 32 candidates, each chunk first tested into a mask of accepted
// candidates and the pair terms then run once per set bit.  About a third of
// the windowed candidates pass r <= h, so a branch on the test keeps the wave
// in the pair terms for nearly every candidate; with the bits the wave's
// pair iterations are its lanes' largest accepted count.  Ascending j again:
// the direct loop's bits.
// kLoopShared (kLoopStagedBits when the density kernel recorded the rows,
// Win::rows): the windows are read instead of looked up, and the last chunk
// of each window (the whole window up to 32 candidates) takes its bits from
// the record instead of testing: the density kernel formed the same u for
// the same candidates.  The direct loop's bits once more.
constexpr int kLoopDirect = 0, kLoopStaged = 4, kLoopStagedBits = 5, kLoopShared = 6;
constexpr int kForceStageCap = 192;   // 6 KB per wave: the 20 waves per CU of kStagedWavesPerSimd take 120 KB

// At least six waves per SIMD: the compiler keeps the force kernel within 80
// VGPRs (from 85, with a few spilled bytes), which measured 3-8% faster than
// five waves on the 1M dam break; eight waves spill too much.
constexpr int kForceWavesPerSimd = 6;
// The staged loop (kLoopStaged) at five waves: 96 VGPRs instead of 80 and no
// spills in the row loop, 2-11% faster than six on the 1M dam.
constexpr int kStagedWavesPerSimd = 5;
// The tile-queue instances at four (128 VGPRs): the loop keeps more values live.
constexpr int kQueuedWavesPerSimd = 4;

template <bool SCALED, bool PACKED, bool FUSED, int LOOP>
__device__ __forceinline__ void force_tile(int tile, int num_host, int* d_num,
                                           const float4* __restrict__ pos_s, const float* __restrict__ dens_in,
                                           const float4* __restrict__ ve_in, const uint32_t* __restrict__ start,
                                           const uint32_t* __restrict__ end, const rtps_sph_params& sp, const Grid& g,
                                           const Win& win, float4* __restrict__ force_out,
                                           float4* __restrict__ xsph_out,
                                           // FUSED only:
                                           const float4* __restrict__ vel_s, const uint32_t* __restrict__ idx,
                                           const GravitySources& gsrc, int collision, const rtps_float4& bmin,
                                           const rtps_float4& bmax, float dt, float4* __restrict__ pos_u,
                                           float4* __restrict__ vel_u, float4* __restrict__ veleval_u,
                                           const float4* __restrict__ color_s, float4* __restrict__ color_u) {
    const int i = tile * (int)blockDim.x + (int)threadIdx.x;
    int num = d_num ? *d_num : num_host;
    bool deleted = false;
    if (d_num) deleted = effective_num(num, start, g.nb_cells, num);
    // FUSED: the step's last kernel commits the count of a deletion step for
    // the next step (SPH.cpp:220-250).  Blocks that read d_num before or
    // after this store get the same count: effective_num of the committed
    // count is that count.  (d_num is not __restrict__: it is read and
    // written here.)
    if (FUSED && deleted && tile == 0 && threadIdx.x == 0) *d_num = num;
    if (i >= num) return;
    const float scale = sp.simulation_scale;
    const float h = sp.smoothing_distance, h2 = h * h, eps = sp.EPSILON;
    const float K = sp.K, rho0 = sp.rest_density;
    float4 pi = pos_s[i];
    const float4 p_unscaled = pi;
    if (SCALED) { pi.x *= scale; pi.y *= scale; pi.z *= scale; }
    float4 vi;
    float di;
    if (PACKED) { vi = ve_in[i]; di = vi.w; } else { vi = ve_in[i]; di = dens_in[i]; }
    const float Pi = K * (di - rho0);
    ForceAcc acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};

    auto pair = [&](uint32_t j, float rx, float ry, float rz, float r2) {
        float4 vj = ve_in[j];
        float dj = PACKED ? vj.w : dens_in[j];
        force_pair(acc, rx, ry, rz, r2, h, h2, eps, K, rho0, Pi, di, vi, vj, dj);
    };
    auto sep = [&](uint32_t j, float& rx, float& ry, float& rz) {
        float4 pj = pos_s[j];
        if (SCALED) { pj.x *= scale; pj.y *= scale; pj.z *= scale; }
        rx = pi.x - pj.x; ry = pi.y - pj.y; rz = pi.z - pj.z;
        return rx * rx + ry * ry + rz * rz;
    };

    if constexpr (LOOP == kLoopDirect) {
        const int3 c = grid_cell(pi.x, pi.y, pi.z, g);
        const WindowCtx wc = window_ctx(pi.x, pi.y, pi.z, g, c, h, win.keys, win.rx, win.ry, win.rz);
        for (int dz = -1; dz <= 1; ++dz) {
            for (int dy = -1; dy <= 1; ++dy) {
                uint32_t a, b;
                row_window(wc, c, dy, dz, h2, g, start, end, win.keys, a, b, 0u, win.sub);
                // groups of kGroup candidates: the position loads, then the
                // veleval loads of the accepted ones, are issued together
                uint32_t j = a;
                for (; j + kGroup <= b; j += kGroup) {
                    // one base address per group: the loads take immediate offsets
                    const float4* __restrict__ pg = pos_s + j;
                    const float4* __restrict__ vg = ve_in + j;
                    float4 q[kGroup], v[kGroup];
#pragma unroll
                    for (int k = 0; k < kGroup; ++k) q[k] = pg[k];
                    float rx[kGroup], ry[kGroup], rz[kGroup], r2[kGroup];
                    bool ok[kGroup];
#pragma unroll
                    for (int k = 0; k < kGroup; ++k) {
                        float4 pj = q[k];
                        if (SCALED) { pj.x *= scale; pj.y *= scale; pj.z *= scale; }
                        rx[k] = pi.x - pj.x; ry[k] = pi.y - pj.y; rz[k] = pi.z - pj.z;
                        r2[k] = rx[k] * rx[k] + ry[k] * ry[k] + rz[k] * rz[k];
                        ok[k] = within_h(h2_minus_r2(rx[k], ry[k], rz[k], h2)) && j + k != (uint32_t)i;
                    }
#pragma unroll
                    for (int k = 0; k < kGroup; ++k)
                        if (ok[k]) v[k] = vg[k];
#pragma unroll
                    for (int k = 0; k < kGroup; ++k)
                        if (ok[k]) {
                            const float dj = PACKED ? v[k].w : dens_in[j + k];
                            force_pair(acc, rx[k], ry[k], rz[k], r2[k], h, h2, eps, K, rho0, Pi, di, vi, v[k], dj);
                        }
                }
                for (; j < b; ++j) {
                    float rx, ry, rz;
                    float r2 = sep(j, rx, ry, rz);
                    if (within_h(h2_minus_r2(rx, ry, rz, h2)) && j != (uint32_t)i) pair(j, rx, ry, rz, r2);
                }
            }
        }
    }
    if constexpr (LOOP == kLoopStaged || LOOP == kLoopStagedBits || LOOP == kLoopShared) {
        const int3 c = grid_cell(pi.x, pi.y, pi.z, g);
        const WindowCtx wc = window_ctx(pi.x, pi.y, pi.z, g, c, h, win.keys, win.rx, win.ry, win.rz);
        extern __shared__ float4 force_rows[];   // 2 x kForceStageCap per wave: positions, then veleval|density
        float4* __restrict__ my_p = force_rows + (threadIdx.x >> 6) * (2 * kForceStageCap);
        float4* __restrict__ my_v = my_p + kForceStageCap;
        const uint64_t live = __ballot(true);   // lanes past num have returned
        const uint32_t nlive = (uint32_t)__popcll(live);
        const uint32_t rank = (uint32_t)__popcll(live & ((1ull << (threadIdx.x & 63)) - 1ull));
        // candidates [a, b) of one row, candidate j at P[j - off] / V[j - off]
        // inv_w: P[.].w holds 1 / d_j (the staged copies), else one reciprocal per pair
        auto walk = [&](const float4* __restrict__ P, const float4* __restrict__ V, uint32_t a, uint32_t b,
                        uint32_t off, bool inv_w) {
            uint32_t j = a;
            for (; j + kGroup <= b; j += kGroup) {
                const float4* __restrict__ pg = P + (j - off);
                const float4* __restrict__ vg = V + (j - off);
                float4 q[kGroup], v[kGroup];
#pragma unroll
                for (int k = 0; k < kGroup; ++k) q[k] = pg[k];
                float rx[kGroup], ry[kGroup], rz[kGroup], r2[kGroup];
                bool ok[kGroup];
#pragma unroll
                for (int k = 0; k < kGroup; ++k) {
                    float4 pj = q[k];
                    if (SCALED) { pj.x *= scale; pj.y *= scale; pj.z *= scale; }
                    rx[k] = pi.x - pj.x; ry[k] = pi.y - pj.y; rz[k] = pi.z - pj.z;
                    r2[k] = rx[k] * rx[k] + ry[k] * ry[k] + rz[k] * rz[k];
                    ok[k] = within_h(h2_minus_r2(rx[k], ry[k], rz[k], h2)) && j + k != (uint32_t)i;
                }
#pragma unroll
                for (int k = 0; k < kGroup; ++k)
                    if (ok[k]) v[k] = vg[k];
#pragma unroll
                for (int k = 0; k < kGroup; ++k)
                    if (ok[k]) {
                        const float dj = PACKED ? v[k].w : dens_in[j + k];
                        force_pair_r(acc, rx[k], ry[k], rz[k], r2[k], h, h2, eps, K, rho0, Pi, di, vi, v[k], dj,
                                     inv_w ? q[k].w : fast_rcp(dj));
                    }
            }
            for (; j < b; ++j) {
                const float4 pq = P[j - off];
                float4 pj = pq;
                if (SCALED) { pj.x *= scale; pj.y *= scale; pj.z *= scale; }
                const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (within_h(h2_minus_r2(rx, ry, rz, h2)) && j != (uint32_t)i) {
                    const float4 vj = V[j - off];
                    const float dj = PACKED ? vj.w : dens_in[j];
                    force_pair_r(acc, rx, ry, rz, r2, h, h2, eps, K, rho0, Pi, di, vi, vj, dj,
                                 inv_w ? pq.w : fast_rcp(dj));
                }
            }
        };
        // Loops 5 and 6 on the staged copies: the lane's window is taken in
        // chunks, the first one ragged (1-32 candidates) and the rest 32
        // long, so that the last chunk is the window's last min(b - a, 32)
        // candidates.  A chunk's tests shift one rejection bit per candidate
        // into a word (rej_bit: the chunk's first candidate ends in its
        // highest bit); loop 6 takes the last chunk's word from the rows the
        // density kernel recorded (rej_last) instead.  The particle's own bit
        // is cleared, then the pair terms run once per accepted candidate, in
        // ascending order.  The wave's pair iterations are its lanes' largest accepted
        // count instead of their longest window, and the pairs are summed in
        // the same ascending order as walk (the same bits).
        auto walk_bits = [&](const float4* __restrict__ P, const float4* __restrict__ V, uint32_t a, uint32_t b,
                             uint32_t off, uint32_t rej_last) {
            uint32_t n = ((b - a - 1u) & 31u) + 1u;
            for (uint32_t j0 = a; j0 < b; j0 += n, n = 32u) {
                const float4* __restrict__ pc = P + (j0 - off);
                uint32_t rej = rej_last;
                if (LOOP != kLoopShared || j0 + n != b) {
                    rej = 0u;
                    uint32_t k = 0u;
                    auto test = [&](float4 pj) {
                        if (SCALED) { pj.x *= scale; pj.y *= scale; pj.z *= scale; }
                        rej = rej_bit(rej, h2_minus_r2(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z, h2));
                    };
                    for (; k + kGroup <= n; k += kGroup) {
                        float4 q[kGroup];
#pragma unroll
                        for (int t = 0; t < kGroup; ++t) q[t] = pc[k + t];
#pragma unroll
                        for (int t = 0; t < kGroup; ++t) test(q[t]);
                    }
                    for (; k < n; ++k) test(pc[k]);
                }
                // accepted bits reversed once, so that bit t is candidate t: the
                // pair loop takes the lowest set bit (ascending j as before)
                uint32_t m = __builtin_bitreverse32(~rej) >> (32u - n);
                const uint32_t self = (uint32_t)i - j0;   // the particle itself passes (r = 0): drop it
                if (self < n) m &= ~(1u << self);
                while (m != 0u) {
                    const uint32_t t = (uint32_t)__builtin_ctz(m);
                    m &= m - 1u;
                    const float4 pq = pc[t];
                    float4 pj = pq;
                    if (SCALED) { pj.x *= scale; pj.y *= scale; pj.z *= scale; }
                    const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
                    const float r2 = rx * rx + ry * ry + rz * rz;
                    const float4 vj = V[j0 + t - off];
                    const float dj = PACKED ? vj.w : dens_in[j0 + t];
                    force_pair_r(acc, rx, ry, rz, r2, h, h2, eps, K, rho0, Pi, di, vi, vj, dj, pq.w);
                }
            }
        };
        // the row's window and last-chunk rejection bits: recorded (loop 6) or looked up
        const uint32_t* __restrict__ rec = LOOP == kLoopShared ? win.rows + 3 * (size_t)i : nullptr;
        for (int dz = -1; dz <= 1; ++dz) {
            for (int dy = -1; dy <= 1; ++dy) {
                uint32_t a, b, A, B, rej_last = 0u;
                if constexpr (LOOP == kLoopShared) {
                    const int r = 3 * (dz + 1) + dy + 1;
                    const uint32_t* __restrict__ o = rec + 3 * ((size_t)r * win.rows_stride);
                    a = o[0];
                    b = o[1];
                    rej_last = o[2];
                    // the wave's union, recorded by the density kernel for the same 64 particles
                    const uint32_t* __restrict__ u = win.rows + 27 * (size_t)win.rows_stride + 18 * (size_t)(i >> 6);
                    A = __builtin_amdgcn_readfirstlane(u[2 * r]);
                    B = __builtin_amdgcn_readfirstlane(u[2 * r + 1]);
                } else {
                    row_window(wc, c, dy, dz, h2, g, start, end, win.keys, a, b, 0u, win.sub);
                    A = __ockl_wfred_min_u32(a < b ? a : 0xFFFFFFFFu);
                    B = __ockl_wfred_max_u32(a < b ? b : 0u);
                }
                if (B <= A) continue;
                if (B - A <= (uint32_t)kForceStageCap) {
                    // the copy keeps 1 / d_j in the position's unused w: one
                    // reciprocal per copied candidate instead of one per pair
                    for (uint32_t t = rank; t < B - A; t += nlive) {
                        const float4 p = pos_s[A + t], e = ve_in[A + t];
                        my_p[t] = make_float4(p.x, p.y, p.z, fast_rcp(PACKED ? e.w : dens_in[A + t]));
                        my_v[t] = e;
                    }
                    __builtin_amdgcn_wave_barrier();   // the copies before any lane reads them
                    if constexpr (LOOP == kLoopStaged) walk(my_p, my_v, a, b, A, true);
                    else walk_bits(my_p, my_v, a, b, A, rej_last);
                    __builtin_amdgcn_wave_barrier();   // every read before the next row's copies
                } else {
                    walk(pos_s, ve_in, a, b, 0u, false);
                }
            }
        }
    }
    float4 F, X;
    force_epilogue(acc, di, sp, F, X);
    if (force_out) { force_out[i] = F; xsph_out[i] = X; }
    if (FUSED) {
        gravity_sources(F, p_unscaled, gsrc, scale);
        const float4 v = vel_s[i];
        if (collision) collision_wall(F, p_unscaled, v, sp, bmin, bmax);
        float4 po, vo, eo;
        leapfrog(p_unscaled, v, F, X, sp, dt, po, vo, eo, SCALED);
        const uint32_t o = deleted ? (uint32_t)i : idx[i];
        pos_u[o] = po;
        vel_u[o] = vo;
        veleval_u[o] = eo;
        // leapfrog.cl:107 writes color_u[o] = color_s[i] every step; it changes
        // something only when the survivors are renumbered
        if (deleted && color_u) color_u[i] = color_s[i];
    }
}

// One tile of blockDim.x sorted particles per block, or (QUEUED) a resident
// grid working through the per-XCD tile queues of win.tq (separate instances,
// at four waves per SIMD: the loop around the body needs the registers).
template <bool SCALED, bool PACKED, bool FUSED, int LOOP, bool QUEUED>
__global__ __launch_bounds__(kBlock, QUEUED ? kQueuedWavesPerSimd : (LOOP >= kLoopStaged ? kStagedWavesPerSimd : kForceWavesPerSimd)) void k_force(int num_host, int* d_num,
                                                  const float4* __restrict__ pos_s,
                                                  const float* __restrict__ dens_in,
                                                  const float4* __restrict__ ve_in,
                                                  const uint32_t* __restrict__ start,
                                                  const uint32_t* __restrict__ end, rtps_sph_params sp, Grid g,
                                                  Win win, float4* __restrict__ force_out,
                                                  float4* __restrict__ xsph_out,
                                                  // FUSED only:
                                                  const float4* __restrict__ vel_s, const uint32_t* __restrict__ idx,
                                                  GravitySources gsrc, int collision, rtps_float4 bmin,
                                                  rtps_float4 bmax, float dt, float4* __restrict__ pos_u,
                                                  float4* __restrict__ vel_u, float4* __restrict__ veleval_u,
                                                  const float4* __restrict__ color_s,
                                                  float4* __restrict__ color_u) {
    if constexpr (QUEUED) {
        int q = (int)(blockIdx.x % kTileQueues);
        for (int t = first_tile(win.ntiles);;) {
            if (t < 0) t = claim_tile(win.tq, win.ntiles, q);
            if (t < 0) break;
            force_tile<SCALED, PACKED, FUSED, LOOP>(t, num_host, d_num, pos_s, dens_in, ve_in, start, end, sp, g, win,
                                                    force_out, xsph_out, vel_s, idx, gsrc, collision, bmin, bmax, dt,
                                                    pos_u, vel_u, veleval_u, color_s, color_u);
            t = -1;
        }
        finish_tiles(win.tq);
    } else {
        force_tile<SCALED, PACKED, FUSED, LOOP>(remap_tile(win.remap, (int)blockIdx.x, (int)gridDim.x), num_host,
                                                d_num, pos_s, dens_in, ve_in, start, end, sp, g, win, force_out,
                                                xsph_out, vel_s, idx, gsrc, collision, bmin, bmax, dt, pos_u, vel_u,
                                                veleval_u, color_s, color_u);
    }
}

// a9 + a10 on stored forces (the unfused update() path): point sources then
// walls, in the reference's order (SPH.cpp:289-297).
__global__ __launch_bounds__(kBlock) void k_post_force(int num_host, const int* __restrict__ d_num,
                                                       const uint32_t* __restrict__ start, int nb_cells,
                                                       const float4* __restrict__ pos_s,
                                                       const float4* __restrict__ vel_s, float4* __restrict__ force,
                                                       rtps_sph_params sp, GravitySources gsrc, int collision,
                                                       rtps_float4 bmin, rtps_float4 bmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int num = d_num ? *d_num : num_host;
    if (d_num) effective_num(num, start, nb_cells, num);
    if (i >= num) return;
    float4 f = force[i];
    const float4 p = pos_s[i];
    gravity_sources(f, p, gsrc, sp.simulation_scale);
    if (collision) collision_wall(f, p, vel_s[i], sp, bmin, bmax);
    force[i] = f;
}

// a11: leapfrog.cl:30-121.
__global__ __launch_bounds__(kBlock) void k_leapfrog(int num_host, const int* __restrict__ d_num,
                                                     const uint32_t* __restrict__ start, int nb_cells,
                                                     const float4* __restrict__ pos_s,
                                                     const float4* __restrict__ vel_s,
                                                     const float4* __restrict__ force,
                                                     const float4* __restrict__ xsph,
                                                     const uint32_t* __restrict__ idx, rtps_sph_params sp, float dt,
                                                     float4* __restrict__ pos_u, float4* __restrict__ vel_u,
                                                     float4* __restrict__ veleval_u, const float4* __restrict__ color_s,
                                                     float4* __restrict__ color_u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int num = d_num ? *d_num : num_host;
    bool deleted = false;
    if (d_num) deleted = effective_num(num, start, nb_cells, num);
    if (i >= num) return;
    float4 po, vo, eo;
    leapfrog(pos_s[i], vel_s[i], force[i], xsph[i], sp, dt, po, vo, eo);
    const uint32_t o = deleted ? (uint32_t)i : idx[i];
    pos_u[o] = po;
    vel_u[o] = vo;
    veleval_u[o] = eo;
    if (deleted && color_u) color_u[i] = color_s[i];   // leapfrog.cl:107
}

// a11 alternative: euler.cl:33-78 (integrator "euler").  veleval is not
// written (euler.cl leaves it alone); on a deletion step the survivors'
// sorted records become the unsorted ones (call_prep, SPH.cpp:445-457), so
// veleval and the colours are copied from the sorted arrays there.  The
// sorted veleval array holds veleval.xyz | density; its copy gets w = 0 (no
// kernel reads veleval.w).
__global__ __launch_bounds__(kBlock) void k_euler(int num_host, const int* __restrict__ d_num,
                                                  const uint32_t* __restrict__ start, int nb_cells,
                                                  const float4* __restrict__ pos_s, const float4* __restrict__ vel_s,
                                                  const float4* __restrict__ force, const float4* __restrict__ ve_s,
                                                  const uint32_t* __restrict__ idx, rtps_sph_params sp, float dt,
                                                  float4* __restrict__ pos_u, float4* __restrict__ vel_u,
                                                  float4* __restrict__ veleval_u,
                                                  const float4* __restrict__ color_s, float4* __restrict__ color_u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int num = d_num ? *d_num : num_host;
    bool deleted = false;
    if (d_num) deleted = effective_num(num, start, nb_cells, num);
    if (i >= num) return;
    float4 po, vo;
    euler(pos_s[i], vel_s[i], force[i], sp, dt, po, vo);
    const uint32_t o = deleted ? (uint32_t)i : idx[i];
    pos_u[o] = po;
    vel_u[o] = vo;
    if (deleted) {
        const float4 e = ve_s[i];
        veleval_u[i] = make_float4(e.x, e.y, e.z, 0.0f);
        if (color_u) color_u[i] = color_s[i];
    }
}

// Pair totals of the neighbour loop on the current sorted state: out[0] =
// candidates of the reference's 27-cell loop (K), out[1] = accepted pairs
// (M, self included), out[2] = candidates the windowed rows actually test.
__global__ __launch_bounds__(kBlock) void k_count_pairs(int num_host, const int* __restrict__ d_num,
                                                        const float4* __restrict__ pos_s,
                                                        const uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ end, rtps_sph_params sp, Grid g,
                                                        Win win, unsigned long long* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int num = d_num ? *d_num : num_host;
    if (d_num) effective_num(num, start, g.nb_cells, num);
    unsigned long long k = 0, m = 0, kw = 0;
    if (i < num) {
        const float scale = sp.simulation_scale, h = sp.smoothing_distance, h2 = h * h;
        float4 pi = pos_s[i];
        pi.x *= scale; pi.y *= scale; pi.z *= scale;
        const int3 c = grid_cell(pi.x, pi.y, pi.z, g);
        const WindowCtx wc = window_ctx(pi.x, pi.y, pi.z, g, c, h, win.keys, win.rx, win.ry, win.rz);
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy) {
                uint32_t a, b, wa, wb;
                row_range(c.x, c.y + dy, c.z + dz, g, start, end, a, b);
                row_window(wc, c, dy, dz, h2, g, start, end, win.keys, wa, wb, 0u, win.sub);
                k += b - a;
                kw += wb - wa;
                for (uint32_t j = a; j < b; ++j) {
                    float4 pj = pos_s[j];
                    float rx = pi.x - pj.x * scale, ry = pi.y - pj.y * scale, rz = pi.z - pj.z * scale;
                    m += (rx * rx + ry * ry + rz * rz <= h2) ? 1u : 0u;
                }
            }
    }
    for (int o = 32; o > 0; o >>= 1) {
        k += __shfl_down(k, o, 64);
        m += __shfl_down(m, o, 64);
        kw += __shfl_down(kw, o, 64);
    }
    if ((threadIdx.x & 63) == 0 && (k | m | kw)) {
        atomicAdd(&out[0], k);
        atomicAdd(&out[1], m);
        atomicAdd(&out[2], kw);
    }
}

hipError_t launch_count_pairs(int num, const int* d_num, const float4* pos_s, const uint32_t* start,
                              const uint32_t* end, const rtps_sph_params& sp, const Grid& g, const Win& win,
                              unsigned long long* out, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_count_pairs<<<blocks_for(num), kBlock, 0, st>>>(num, d_num, pos_s, start, end, sp, g, win, out);
    return hipGetLastError();
}

// Commits the deletion path's particle count for the next step.
__global__ void k_commit_num(int* __restrict__ d_num, const uint32_t* __restrict__ start, int nb_cells) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        int num = *d_num;
        int nc = (int)start[nb_cells];
        if (nc >= 0 && nc <= num) *d_num = nc;
    }
}

// ---------------------------------------------------------------------------
// Launchers shared by the C ABI and the system.

// Force-loop variant: 0 = one branch per candidate, 4 = 0 with the rows
// staged through LDS, 5 = 4 with the accepted candidates of each 32-candidate
// chunk marked in a bit mask first (the default).
int g_force_variant = kLoopStagedBits;

// Threads per block of a neighbour launch (a multiple of 64, at most kBlock)
// and its force-loop variant.
static inline int nbr_block(const Win& w) { return w.block > 0 && w.block <= kBlock && w.block % 64 == 0 ? w.block : kBlock; }
static inline int nbr_variant(const Win& w) {
    const int v = w.variant >= 0 ? w.variant : g_force_variant;
    return v == kLoopStaged || v == kLoopStagedBits ? v : kLoopDirect;
}

hipError_t launch_hash(int num, const int* d_num, int n_pad, uint32_t pad_key, const float4* pos_u, uint32_t* keys,
                       uint32_t* idx, const Grid& g, uint32_t* start_clear, int n_clear, int sub_bits,
                       hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    const int n = n_pad > n_clear ? n_pad : n_clear;
    if (n <= 0) return hipSuccess;
    k_hash<<<blocks_for(n), kBlock, 0, st>>>(num, d_num, n_pad, pad_key, pos_u, keys, idx, g, start_clear, n_clear,
                                             sub_bits);
    return hipGetLastError();
}

hipError_t launch_cellindices(int num, const int* d_num, const uint32_t* keys, int nb_cells, uint32_t* start,
                              uint32_t* end, int sub_bits, hipStream_t st, uint32_t* sub) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_cellindices<<<blocks_for(num), kBlock, 0, st>>>(num, d_num, keys, (uint32_t)nb_cells, start, end, sub_bits,
                                                       sub_bits == kSubBits ? sub : nullptr);
    return hipGetLastError();
}

hipError_t launch_cells_permute(int num, const int* d_num, const uint32_t* keys, int nb_cells, uint32_t* start,
                                uint32_t* end, int sub_bits, uint32_t* sub, const uint32_t* idx, const float4* pu,
                                const float4* vu, const float4* eu, float4* ps, float4* vs, float4* es,
                                hipStream_t st, const float4* cu, float4* cs) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_cells_permute<<<blocks_for(num), kBlock, 0, st>>>(num, d_num, keys, (uint32_t)nb_cells, start, end, sub_bits,
                                                        sub_bits == kSubBits ? sub : nullptr, idx, pu, vu, eu, ps,
                                                        vs, es, cu, cs);
    return hipGetLastError();
}

hipError_t launch_read_nc(const uint32_t* start, int nb_cells, int* nc, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    k_read_nc<<<1, 64, 0, st>>>(start, nb_cells, nc);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, long n, hipStream_t st);

hipError_t launch_permute(int num, const int* d_num, const uint32_t* idx, const float4* pu, const float4* vu,
                          const float4* eu, float4* ps, float4* vs, float4* es, hipStream_t st, const float4* cu,
                          float4* cs) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_permute<<<blocks_for(num), kBlock, 0, st>>>(num, d_num, idx, pu, vu, eu, ps, vs, es, cu, cs);
    return hipGetLastError();
}

// Blocks of a resident grid of `kernel` (every block on the device at once),
// at most `ntiles`: the launch shape of the tile-queue neighbour kernels.
// Looked up once per kernel, block size and LDS size (host queries only, so
// it is safe inside a stream capture).
static int resident_grid(const void* kernel, int bs, size_t lds, int ntiles) {
    static std::mutex mu;
    static std::map<std::pair<const void*, long>, int> cache;
    const long key = (long)bs << 32 | (long)lds;
    int total = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find({kernel, key});
        if (it != cache.end()) total = it->second;
    }
    if (total == 0) {
        int dev = 0, cus = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, bs, lds) != hipSuccess || per_cu <= 0)
            per_cu = 1;
        (void)hipGetLastError();
        total = cus * per_cu;
        std::lock_guard<std::mutex> lk(mu);
        cache[{kernel, key}] = total;
    }
    return total < ntiles ? total : ntiles;
}

hipError_t launch_density(int num, const int* d_num, const float4* pos_s, const uint32_t* start, const uint32_t* end,
                          const rtps_sph_params& sp, const Grid& g, const Win& win_in, float* dens, float4* packed,
                          hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    const bool scaled = sp.simulation_scale != 1.0f;
    const int bs = nbr_block(win_in);
    const int nb = blocks_for(num, bs);
    const bool stage = g_density_stage != 0;
    const size_t lds = stage ? (size_t)(bs / 64) * kStageCap * sizeof(float4) : 0;
    if (win_in.rows && win_in.rows_stride < (uint32_t)num) return hipErrorInvalidValue;
    Win win = win_in;
    win.ntiles = nb;
#define RTPS_DENS(S, P, T, E)                                                                                       \
    if (win.tq && P)                                                                                                \
        k_density<S, P, T, true, E><<<resident_grid((const void*)&k_density<S, P, T, true, E>, bs, lds, nb), bs,     \
                                      lds, st>>>(num, d_num, pos_s, start, end, sp, g, win, dens, packed);             \
    else                                                                                                            \
        k_density<S, P, T, false, E><<<nb, bs, lds, st>>>(num, d_num, pos_s, start, end, sp, g, win, dens, packed)
#define RTPS_DENS_T(T, E)                                                                                           \
    if (packed) { if (scaled) { RTPS_DENS(true, true, T, E); } else { RTPS_DENS(false, true, T, E); } }               \
    else { if (scaled) { RTPS_DENS(true, false, T, E); } else { RTPS_DENS(false, false, T, E); } }
    if (win.rows) {
        if (stage) { RTPS_DENS_T(true, true) } else { RTPS_DENS_T(false, true) }
    } else {
        if (stage) { RTPS_DENS_T(true, false) } else { RTPS_DENS_T(false, false) }
    }
#undef RTPS_DENS_T
#undef RTPS_DENS
    return hipGetLastError();
}

static inline size_t force_lds(int loop, int bs) {
    return loop == kLoopDirect ? 0 : (size_t)(bs / 64) * 2 * kForceStageCap * sizeof(float4);
}

// The force loop of a launch: loop 5 becomes 6 when the rows are recorded.
static inline int force_loop(const Win& w) {
    const int v = nbr_variant(w);
    return v == kLoopStagedBits && w.rows ? kLoopShared : v;
}

hipError_t launch_force(int num, int* d_num, const float4* pos_s, const float* dens, const float4* ve,
                        bool packed, const uint32_t* start, const uint32_t* end, const rtps_sph_params& sp,
                        const Grid& g, const Win& win_in, float4* force, float4* xsph, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    const bool scaled = sp.simulation_scale != 1.0f;
    const int bs = nbr_block(win_in);
    const int nb = blocks_for(num, bs);
    const int loop = force_loop(win_in);
    if (win_in.rows && win_in.rows_stride < (uint32_t)num) return hipErrorInvalidValue;
    const size_t lds = force_lds(loop, bs);
    GravitySources none{};
    rtps_float4 z{0, 0, 0, 0};
    Win win = win_in;
    win.ntiles = nb;
#define RTPS_FORCE_Q(S, P, L, Q, GRID)                                                                              \
    k_force<S, P, false, L, Q><<<GRID, bs, lds, st>>>(num, d_num, pos_s, dens, ve, start, end, sp, g, win, force, xsph, \
                                                     nullptr, nullptr, none, 0, z, z, 0.f, nullptr, nullptr, nullptr, \
                                                     nullptr, nullptr)
#define RTPS_FORCE_L(S, P, L)                                                                                       \
    if (win.tq && P) {                                                                                              \
        RTPS_FORCE_Q(S, P, L, true, resident_grid((const void*)&k_force<S, P, false, L, true>, bs, lds, nb));        \
    } else {                                                                                                        \
        RTPS_FORCE_Q(S, P, L, false, nb);                                                                           \
    }
#define RTPS_FORCE(S, P)                                                                                            \
    if (loop == kLoopStaged) { RTPS_FORCE_L(S, P, kLoopStaged) }                                                    \
    else if (loop == kLoopStagedBits) { RTPS_FORCE_L(S, P, kLoopStagedBits) }                                       \
    else if (loop == kLoopShared) { RTPS_FORCE_L(S, P, kLoopShared) }                                               \
    else { RTPS_FORCE_L(S, P, kLoopDirect) }
    if (packed) { if (scaled) { RTPS_FORCE(true, true); } else { RTPS_FORCE(false, true); } }
    else { if (scaled) { RTPS_FORCE(true, false); } else { RTPS_FORCE(false, false); } }
#undef RTPS_FORCE
#undef RTPS_FORCE_L
#undef RTPS_FORCE_Q
    return hipGetLastError();
}

hipError_t launch_force_fused(int num, int* d_num, const float4* pos_s, const float4* ve_packed,
                              const uint32_t* start, const uint32_t* end, const rtps_sph_params& sp, const Grid& g,
                              const Win& win_in, float4* force, float4* xsph, const float4* vel_s,
                              const uint32_t* idx, const GravitySources& gsrc, int collision,
                              const rtps_grid_params& gps, float dt, float4* pos_u, float4* vel_u, float4* veleval_u,
                              hipStream_t st, const float4* color_s, float4* color_u) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    const bool scaled = sp.simulation_scale != 1.0f;
    const int bs = nbr_block(win_in);
    const int nb = blocks_for(num, bs);
    const int loop = force_loop(win_in);
    if (win_in.rows && win_in.rows_stride < (uint32_t)num) return hipErrorInvalidValue;
    const size_t lds = force_lds(loop, bs);
    Win win = win_in;
    win.ntiles = nb;
#define RTPS_FFUSED_Q(S, L, Q, GRID)                                                                                \
    k_force<S, true, true, L, Q><<<GRID, bs, lds, st>>>(num, d_num, pos_s, nullptr, ve_packed, start, end, sp, g, win, \
                                                       force, xsph, vel_s, idx, gsrc, collision, gps.bnd_min,        \
                                                       gps.bnd_max, dt, pos_u, vel_u, veleval_u, color_s, color_u)
#define RTPS_FFUSED_L(S, L)                                                                                         \
    if (win.tq) {                                                                                                   \
        RTPS_FFUSED_Q(S, L, true, resident_grid((const void*)&k_force<S, true, true, L, true>, bs, lds, nb));        \
    } else {                                                                                                        \
        RTPS_FFUSED_Q(S, L, false, nb);                                                                             \
    }
#define RTPS_FFUSED(S)                                                                                              \
    if (loop == kLoopStaged) { RTPS_FFUSED_L(S, kLoopStaged) }                                                      \
    else if (loop == kLoopStagedBits) { RTPS_FFUSED_L(S, kLoopStagedBits) }                                         \
    else if (loop == kLoopShared) { RTPS_FFUSED_L(S, kLoopShared) }                                                 \
    else { RTPS_FFUSED_L(S, kLoopDirect) }
    if (scaled) { RTPS_FFUSED(true) } else { RTPS_FFUSED(false) }
#undef RTPS_FFUSED
#undef RTPS_FFUSED_L
#undef RTPS_FFUSED_Q
    return hipGetLastError();
}

hipError_t launch_post_force(int num, const int* d_num, const uint32_t* start, int nb_cells, const float4* pos_s,
                             const float4* vel_s, float4* force, const rtps_sph_params& sp,
                             const GravitySources& gsrc, int collision, const rtps_grid_params& gps, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    if (gsrc.n == 0 && !collision) return hipSuccess;
    k_post_force<<<blocks_for(num), kBlock, 0, st>>>(num, d_num, start, nb_cells, pos_s, vel_s, force, sp, gsrc,
                                                      collision, gps.bnd_min, gps.bnd_max);
    return hipGetLastError();
}

hipError_t launch_leapfrog(int num, const int* d_num, const uint32_t* start, int nb_cells, const float4* pos_s,
                           const float4* vel_s, const float4* force, const float4* xsph, const uint32_t* idx,
                           const rtps_sph_params& sp, float dt, float4* pos_u, float4* vel_u, float4* veleval_u,
                           hipStream_t st, const float4* color_s, float4* color_u) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_leapfrog<<<blocks_for(num), kBlock, 0, st>>>(num, d_num, start, nb_cells, pos_s, vel_s, force, xsph, idx, sp, dt,
                                                    pos_u, vel_u, veleval_u, color_s, color_u);
    return hipGetLastError();
}

hipError_t launch_euler(int num, const int* d_num, const uint32_t* start, int nb_cells, const float4* pos_s,
                        const float4* vel_s, const float4* force, const float4* ve_s, const uint32_t* idx,
                        const rtps_sph_params& sp, float dt, float4* pos_u, float4* vel_u, float4* veleval_u,
                        hipStream_t st, const float4* color_s, float4* color_u) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_euler<<<blocks_for(num), kBlock, 0, st>>>(num, d_num, start, nb_cells, pos_s, vel_s, force, ve_s, idx, sp, dt,
                                                 pos_u, vel_u, veleval_u, color_s, color_u);
    return hipGetLastError();
}

hipError_t launch_commit_num(int* d_num, const uint32_t* start, int nb_cells, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    k_commit_num<<<1, 64, 0, st>>>(d_num, start, nb_cells);
    return hipGetLastError();
}

}  // namespace rtps
