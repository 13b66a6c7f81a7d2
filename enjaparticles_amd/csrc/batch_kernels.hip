// Batched scenes: many independent SPH systems stepped together, one launch
// per stage for all of them (no reference counterpart; each scene follows the
// single-system step of sph_kernels.hip exactly).
//
// Layout: scene s owns the slots [off_s, off_s + cap_s) of every particle
// array, its own parameters (SceneDev) and its own cell table at cell_base_s.
// Keys are segmented, (s << 24) | (cell << 4) | xq, so one sort over all
// slots keeps every scene inside its own slot range; pads (unused slots) get
// (s << 24) | 0xFFFFFF and sort to the end of their scene.  Density and force
// run as persistent kernels over a queue of (scene, 256-slot tile) items: a
// block takes the next item with one atomic, so uneven scene sizes balance
// themselves, and the scene's parameters are wave-uniform (scalar loads).

#include <hip/hip_runtime.h>

#include <type_traits>

#include "batch_types.h"
#include "rtps_internal.h"
#include "sph_device.h"

namespace rtps {

constexpr int kBBlock = 256;
// Candidates per LDS copy of a row in the pair kernels (force: positions and
// veleval|density; density: positions, twice as many).
constexpr int kBStageCap = 192;
constexpr uint32_t kLocalMask = (1u << kSceneShift) - 1u;

static inline int bblocks(long n) { return (int)((n + kBBlock - 1) / kBBlock); }

// Count of scene s after its deletion path (SPH.cpp:220-250): the first
// out-of-grid particle's sorted index, when it lies inside [0, num].
__device__ __forceinline__ bool scene_num(const SceneDev& S, int num, const uint32_t* __restrict__ start, int& out) {
    const uint32_t nc = start[S.cell_base + S.g.nb_cells];
    const int local = nc == kEmpty ? -1 : (int)nc - S.off;
    const bool del = local >= 0 && local <= num;
    out = del ? local : num;
    return del;
}

__global__ __launch_bounds__(kBBlock) void k_hash_b(int T, const uint16_t* __restrict__ sid,
                                                    const SceneDev* __restrict__ scenes, const int* __restrict__ d_num,
                                                    const float4* __restrict__ pos_u, uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ idx, uint32_t* __restrict__ start_clear,
                                                    int n_clear) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_clear) start_clear[i] = kEmpty;
    if (i >= T) return;
    const int s = sid[i];
    const SceneDev& S = scenes[s];
    uint32_t key = ((uint32_t)s << kSceneShift) | kLocalMask;
    if (i - S.off < d_num[s]) {
        const Grid& g = S.g;
        const float4 p = pos_u[i];
        const float ux = (p.x - g.min.x) * g.delta.x;
        const int cx = sat_f2i(ux);
        const int cy = sat_f2i((p.y - g.min.y) * g.delta.y);
        const int cz = sat_f2i((p.z - g.min.z) * g.delta.z);
        int h = grid_hash(cx, cy, cz, g);
        const bool out = h > g.nb_cells || h < 0;
        h = out ? g.nb_cells : h;
        int xq = sat_f2i((ux - (float)cx) * (float)kSubCells);
        xq = out || h == g.nb_cells ? 0 : min(max(xq, 0), kSubCells - 1);
        key = ((uint32_t)s << kSceneShift) | ((uint32_t)h << kSubBits) | (uint32_t)xq;
    }
    keys[i] = key;
    idx[i] = (uint32_t)i;
}

__device__ __forceinline__ bool cell_of(uint32_t key, const SceneDev* __restrict__ scenes, uint32_t& gcell) {
    const int s = (int)(key >> kSceneShift);
    const uint32_t cell = (key & kLocalMask) >> kSubBits;
    const SceneDev& S = scenes[s];
    gcell = (uint32_t)S.cell_base + cell;
    return cell <= (uint32_t)S.g.nb_cells;
}

// Cell start/end on the segmented keys; start/end hold global sorted indices.
__global__ __launch_bounds__(kBBlock) void k_cellindices_b(int T, const uint32_t* __restrict__ keys,
                                                           const SceneDev* __restrict__ scenes,
                                                           uint32_t* __restrict__ start, uint32_t* __restrict__ end,
                                                           uint32_t* __restrict__ sub) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= T) return;
    uint32_t g;
    if (!cell_of(keys[i], scenes, g)) return;
    uint32_t gp = 0, gn = 0;
    const bool first = i == 0 || !cell_of(keys[i - 1], scenes, gp) || gp != g;
    const bool last = i == T - 1 || !cell_of(keys[i + 1], scenes, gn) || gn != g;
    if (first) start[g] = (uint32_t)i;
    if (last) end[g] = (uint32_t)(i + 1);
    const uint32_t key = keys[i];
    const SceneDev& S = scenes[key >> kSceneShift];
    if (sub && ((key & kLocalMask) >> kSubBits) < (uint32_t)S.g.nb_cells) {
        // sub-cell starts of the windowed rows (as in the plain cellindices)
        const uint32_t xq = key & (kSubCells - 1);
        const uint32_t q0 = first ? 0u : (keys[i - 1] & (kSubCells - 1)) + 1;
        uint32_t* row = sub + (size_t)g * kSubCells;
        for (uint32_t q = q0; q <= xq; ++q) row[q] = (uint32_t)i;
        if (last)
            for (uint32_t q = xq + 1; q < (uint32_t)kSubCells; ++q) row[q] = (uint32_t)i + 1;
    }
}

// One particle of a work item (sorted slot i of scene s).  Kept out of the
// persistent loop so that the loop has no divergent exits: every lane reaches
// the loop's barriers together (a divergent continue there lets the compiler
// retire lanes one by one, and a parked lane 0 would stop fetching items).
// ROWS: density records each row's window and the rejection bits of its last
// 32 candidates at rows + 3 (r * rows_stride + i) (Win::rows of the plain
// kernels); the force modes read them instead of looking the rows up, and
// take the last chunk's bits from the record.
template <int MODE, bool ROWS>
__device__ __forceinline__ void pairs_one(int i, int s, const SceneDev& S, bool deleted,
                                          const float4* __restrict__ pos_s, const uint32_t* __restrict__ start,
                                          const uint32_t* __restrict__ end, const uint32_t* __restrict__ keys,
                                          float4* __restrict__ ve, const float4* __restrict__ vel_s,
                                          const uint32_t* __restrict__ idx, float4* __restrict__ pos_u,
                                          float4* __restrict__ vel_u, float4* __restrict__ veleval_u,
                                          float4* __restrict__ force_out, float4* __restrict__ xsph_out,
                                          const uint32_t* __restrict__ sub_all, float4* __restrict__ stage,
                                          uint32_t* __restrict__ rows, uint32_t rows_stride) {
    const rtps_sph_params& sp = S.sp;
    const uint32_t* sub = sub_all ? sub_all + (size_t)S.cell_base * kSubCells : nullptr;
    const float scale = sp.simulation_scale;
    const float h = sp.smoothing_distance, h2 = h * h;
    float4 pi = pos_s[i];
    const float4 p_unscaled = pi;
    pi.x *= scale;
    pi.y *= scale;
    pi.z *= scale;
    const Grid& g = S.gs;
    const int3 c = grid_cell(pi.x, pi.y, pi.z, g);
    const WindowCtx wc = window_ctx(pi.x, pi.y, pi.z, g, c, h, keys, S.res_x, S.res_y, S.res_z);
    const uint32_t* st = start + S.cell_base;
    const uint32_t* en = end + S.cell_base;
    const uint32_t prefix = (uint32_t)s << kSceneShift;
    // Each wave copies the union of its lanes' windows of a row (when at most
    // kBStageCap candidates) into its LDS slice `stage` with full-line loads
    // and the lanes read their candidates from there; wider unions read
    // global memory.  Same candidates in the same order either way.
    const uint64_t live = __ballot(true);
    const uint32_t nlive = (uint32_t)__popcll(live);
    const uint32_t rank = (uint32_t)__popcll(live & ((1ull << (threadIdx.x & 63)) - 1ull));
    // the row's wave-wide candidate range [A, B) (A >= B: no lane has candidates)
    auto wave_range = [&](uint32_t a, uint32_t b, uint32_t& A, uint32_t& B) {
        A = __ockl_wfred_min_u32(a < b ? a : 0xFFFFFFFFu);
        B = __ockl_wfred_max_u32(a < b ? b : 0u);
    };
    // the row's record of particle i (ROWS)
    auto rec = [&](int dz, int dy) { return rows + 3 * ((size_t)(3 * (dz + 1) + dy + 1) * rows_stride + (size_t)i); };
    if (MODE == kPairsDensity) {
        float sum = 0.0f;
        uint32_t rej = 0u;   // ROWS: the current row's rejection bits
        auto term = [&](float4 pj) {   // poly6_add_r, keeping h^2 - r^2 for the record
            const float u = h2_minus_r2(pi.x - pj.x * scale, pi.y - pj.y * scale, pi.z - pj.z * scale, h2);
            const float t = fmaxf(u, 0.0f);
            sum = fmaf(t * t, t, sum);
            if (ROWS) rej = rej_bit(rej, u);
        };
        // candidate j at P[j - off]
        auto walk = [&](const float4* __restrict__ P, uint32_t a, uint32_t b, uint32_t off) {
            const float4* __restrict__ pg = P + (a - off);
            const uint32_t n = a < b ? b - a : 0u;
            for (uint32_t gr = n / kGroup; gr != 0u; --gr, pg += kGroup) {   // loads of a group issued together
                float4 p[kGroup];
#pragma unroll
                for (int k = 0; k < kGroup; ++k) p[k] = pg[k];
#pragma unroll
                for (int k = 0; k < kGroup; ++k) term(p[k]);
            }
            for (uint32_t t = n % kGroup; t != 0u; --t, ++pg) term(*pg);
        };
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy) {
                uint32_t a, b, A, B;
                row_window(wc, c, dy, dz, h2, g, st, en, keys, a, b, prefix, sub);
                wave_range(a, b, A, B);
                if (B > A) {
                    if (B - A <= (uint32_t)(2 * kBStageCap)) {   // density uses the whole slice
                        for (uint32_t t = rank; t < B - A; t += nlive) stage[t] = pos_s[A + t];
                        __builtin_amdgcn_wave_barrier();   // the copies before any lane reads them
                        walk(stage, a, b, A);
                        __builtin_amdgcn_wave_barrier();   // every read before the next row's copies
                    } else {
                        walk(pos_s, a, b, 0u);
                    }
                }
                if (ROWS) {
                    uint32_t* __restrict__ o = rec(dz, dy);
                    o[0] = a;
                    o[1] = b;
                    o[2] = rej;
                    rej = 0u;
                }
            }
        ve[i].w = sp.wpoly6_coef * sp.mass * sum;
    } else {
        const float eps = sp.EPSILON, K = sp.K, rho0 = sp.rest_density;
        const float4 vi = ve[i];
        const float di = vi.w;
        const float Pi = K * (di - rho0);
        ForceAcc acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        float4* __restrict__ stage_v = stage + kBStageCap;
        // candidate j at P[j - off] / V[j - off]; inv_w: P[.].w holds 1 / d_j
        // (the staged copies), else one reciprocal per pair
        auto walk = [&](const float4* __restrict__ P, const float4* __restrict__ V, uint32_t a, uint32_t b,
                        uint32_t off, bool inv_w) {
            uint32_t j = a;
            for (; j + kGroup <= b; j += kGroup) {   // positions, then accepted velocities, loaded together
                const float4* __restrict__ pg = P + (j - off);
                const float4* __restrict__ vg = V + (j - off);
                float4 p[kGroup], v[kGroup];
#pragma unroll
                for (int k = 0; k < kGroup; ++k) p[k] = pg[k];
                float rx[kGroup], ry[kGroup], rz[kGroup], r2[kGroup];
                bool ok[kGroup];
#pragma unroll
                for (int k = 0; k < kGroup; ++k) {
                    rx[k] = pi.x - p[k].x * scale;
                    ry[k] = pi.y - p[k].y * scale;
                    rz[k] = pi.z - p[k].z * scale;
                    r2[k] = rx[k] * rx[k] + ry[k] * ry[k] + rz[k] * rz[k];
                    ok[k] = within_h(h2_minus_r2(rx[k], ry[k], rz[k], h2)) && j + k != (uint32_t)i;
                }
#pragma unroll
                for (int k = 0; k < kGroup; ++k)
                    if (ok[k]) v[k] = vg[k];
#pragma unroll
                for (int k = 0; k < kGroup; ++k)
                    if (ok[k])
                        force_pair_r(acc, rx[k], ry[k], rz[k], r2[k], h, h2, eps, K, rho0, Pi, di, vi, v[k], v[k].w,
                                     inv_w ? p[k].w : fast_rcp(v[k].w));
            }
            for (; j < b; ++j) {
                float4 pj = P[j - off];
                float rx = pi.x - pj.x * scale, ry = pi.y - pj.y * scale, rz = pi.z - pj.z * scale;
                float r2 = rx * rx + ry * ry + rz * rz;
                if (within_h(h2_minus_r2(rx, ry, rz, h2)) && j != (uint32_t)i) {
                    const float4 vj = V[j - off];
                    force_pair_r(acc, rx, ry, rz, r2, h, h2, eps, K, rho0, Pi, di, vi, vj, vj.w,
                                 inv_w ? pj.w : fast_rcp(vj.w));
                }
            }
        };
        // the staged copies, a lane's window in chunks (a ragged first one of
        // 1-32 candidates, then 32s, so the last chunk is the window's last
        // min(b - a, 32)): the tests of a chunk shift one rejection bit per
        // candidate into a word (rej_bit: the chunk's first candidate ends in
        // its highest bit), or (ROWS) the last chunk's word comes from the
        // record; the particle's own bit is cleared, then the pair terms run
        // once per accepted candidate, in ascending order (the same order and
        // arithmetic as walk, so the same sums)
        auto walk_bits = [&](uint32_t a, uint32_t b, uint32_t off, uint32_t rej_last) {
            uint32_t n = ((b - a - 1u) & 31u) + 1u;
            for (uint32_t j0 = a; j0 < b; j0 += n, n = 32u) {
                const float4* __restrict__ pc = stage + (j0 - off);
                uint32_t rej = rej_last;
                if (!ROWS || j0 + n != b) {
                    rej = 0u;
                    uint32_t k = 0u;
                    auto test = [&](float4 pj) {
                        const float rx = pi.x - pj.x * scale, ry = pi.y - pj.y * scale, rz = pi.z - pj.z * scale;
                        rej = rej_bit(rej, h2_minus_r2(rx, ry, rz, h2));
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
                // bit t = candidate t (reversed once): lowest set bit first
                uint32_t m = __builtin_bitreverse32(~rej) >> (32u - n);
                const uint32_t self = (uint32_t)i - j0;   // the particle itself passes (r = 0): drop it
                if (self < n) m &= ~(1u << self);
                while (m != 0u) {
                    const uint32_t t = (uint32_t)__builtin_ctz(m);
                    m &= m - 1u;
                    const float4 pj = pc[t];
                    const float rx = pi.x - pj.x * scale, ry = pi.y - pj.y * scale, rz = pi.z - pj.z * scale;
                    const float r2 = rx * rx + ry * ry + rz * rz;
                    const float4 vj = stage_v[j0 + t - off];
                    force_pair_r(acc, rx, ry, rz, r2, h, h2, eps, K, rho0, Pi, di, vi, vj, vj.w, pj.w);
                }
            }
        };
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy) {
                uint32_t a, b, A, B, rej_last = 0u;
                if constexpr (ROWS) {
                    const uint32_t* __restrict__ o = rec(dz, dy);
                    a = o[0];
                    b = o[1];
                    rej_last = o[2];
                } else {
                    row_window(wc, c, dy, dz, h2, g, st, en, keys, a, b, prefix, sub);
                }
                wave_range(a, b, A, B);
                if (B <= A) continue;
                if (B - A <= (uint32_t)kBStageCap) {
                    // the copy keeps 1 / d_j in the position's unused w
                    for (uint32_t t = rank; t < B - A; t += nlive) {
                        const float4 q = pos_s[A + t], e = ve[A + t];
                        stage[t] = make_float4(q.x, q.y, q.z, fast_rcp(e.w));
                        stage_v[t] = e;
                    }
                    __builtin_amdgcn_wave_barrier();   // the copies before any lane reads them
                    walk_bits(a, b, A, rej_last);
                    __builtin_amdgcn_wave_barrier();   // every read before the next row's copies
                } else {
                    walk(pos_s, ve, a, b, 0u, false);
                }
            }
        float4 F, X;
        force_epilogue(acc, di, sp, F, X);
        if (MODE == kPairsForceStore) {
            force_out[i] = F;
            xsph_out[i] = X;
            return;
        }
        const float4 v = vel_s[i];
        if (S.gsrc.n) gravity_sources(F, p_unscaled, S.gsrc, sp.simulation_scale);   // SPH.cpp:289-296
        if (S.collision) collision_wall(F, p_unscaled, v, sp, S.bmin, S.bmax);
        float4 po, vo, eo;
        leapfrog(p_unscaled, v, F, X, sp, S.dt, po, vo, eo);
        const uint32_t o = deleted ? (uint32_t)i : idx[i];
        pos_u[o] = po;
        vel_u[o] = vo;
        veleval_u[o] = eo;
    }
}

// MODE: kPairsDensity, kPairsForceStep (force + walls + leapfrog, scattered to
// the unsorted arrays) or kPairsForceStore (force and XSPH stored in sorted
// order for the adjoint's recorded step).  BLOCK threads per block = slots per
// work item; compiled for WAVES waves per SIMD (the register budget that keeps
// the grid's WAVES * 256 / BLOCK blocks per CU resident).
template <int MODE, int BLOCK, int WAVES, bool ROWS>
__global__ __launch_bounds__(BLOCK, WAVES) void k_pairs_b(const BatchItem* __restrict__ items, int n_items,
                                                     int* __restrict__ queue, const SceneDev* __restrict__ scenes,
                                                     const int* __restrict__ d_num, const float4* __restrict__ pos_s,
                                                     const uint32_t* __restrict__ start,
                                                     const uint32_t* __restrict__ end,
                                                     const uint32_t* __restrict__ keys, float4* __restrict__ ve,
                                                     const float4* __restrict__ vel_s,
                                                     const uint32_t* __restrict__ idx, float4* __restrict__ pos_u,
                                                     float4* __restrict__ vel_u, float4* __restrict__ veleval_u,
                                                     float4* __restrict__ force_out, float4* __restrict__ xsph_out,
                                                     const uint32_t* __restrict__ sub, uint32_t* __restrict__ rows,
                                                     uint32_t rows_stride) {
    __shared__ float4 stage_b[BLOCK / 64][2 * kBStageCap];   // one 6 KB slice per wave
    int q = (int)(blockIdx.x % kTileQueues);   // thread 0's current queue
    // the loop has one block-uniform exit (it < 0): every block leaves once all queues are drained
    for (int it = first_tile(n_items);;) {
        if (it < 0) it = claim_tile(queue, n_items, q);
        if (it < 0) break;
        const BatchItem I = items[it];
        const int s = __builtin_amdgcn_readfirstlane(I.scene);
        const SceneDev& S = scenes[s];
        int num;
        const bool deleted = scene_num(S, d_num[s], start, num);
        const int i = I.first + (int)threadIdx.x;
        if ((int)threadIdx.x < I.count && i - S.off < num)
            pairs_one<MODE, ROWS>(i, s, S, deleted, pos_s, start, end, keys, ve, vel_s, idx, pos_u, vel_u, veleval_u,
                                  force_out, xsph_out, sub, stage_b[threadIdx.x >> 6], rows, rows_stride);
        it = -1;
    }
    finish_tiles(queue);
}

__global__ void k_commit_b(int S, const SceneDev* __restrict__ scenes, int* __restrict__ d_num,
                           const uint32_t* __restrict__ start) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    int num;
    if (scene_num(scenes[s], d_num[s], start, num)) d_num[s] = num;
}

hipError_t launch_hash_b(int T, const uint16_t* sid, const SceneDev* scenes, const int* d_num, const float4* pos_u,
                         uint32_t* keys, uint32_t* idx, uint32_t* start, int n_cells, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    const int n = T > n_cells ? T : n_cells;
    if (n <= 0) return hipSuccess;
    k_hash_b<<<bblocks(n), kBBlock, 0, st>>>(T, sid, scenes, d_num, pos_u, keys, idx, start, n_cells);
    return hipGetLastError();
}

hipError_t launch_cellindices_b(int T, const uint32_t* keys, const SceneDev* scenes, uint32_t* start, uint32_t* end,
                                hipStream_t st, uint32_t* sub) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    k_cellindices_b<<<bblocks(T), kBBlock, 0, st>>>(T, keys, scenes, start, end, sub);
    return hipGetLastError();
}

int pairs_blocks_per_cu(const PairsCfg& c) { return c.waves * 256 / c.block; }

bool pairs_cfg_valid(int mode, const PairsCfg& c) {
    if (c.block != 64 && c.block != 128 && c.block != 256) return false;
    return mode == kPairsDensity ? (c.waves >= 4 && c.waves <= 6) : (c.waves >= 3 && c.waves <= 5);
}

PairsCfg pairs_default_cfg(int mode) { return mode == kPairsDensity ? PairsCfg{256, 5} : PairsCfg{256, 4}; }

hipError_t launch_pairs_b(int mode, const PairsCfg& cfg, int grid, const BatchItem* items, int n_items, int* queue,
                          const SceneDev* scenes, const int* d_num, const float4* pos_s, const uint32_t* start,
                          const uint32_t* end, const uint32_t* keys, float4* ve, const float4* vel_s,
                          const uint32_t* idx, float4* pos_u, float4* vel_u, float4* veleval_u, hipStream_t st,
                          float4* force_out, float4* xsph_out, const uint32_t* sub, uint32_t* rows,
                          uint32_t rows_stride) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (mode == kPairsForceStore && (!force_out || !xsph_out)) return hipErrorInvalidValue;
    if (rows && rows_stride == 0) return hipErrorInvalidValue;
    if (!pairs_cfg_valid(mode, cfg)) return hipErrorInvalidValue;
    // Instantiated shapes: density at 4, 5 and 6 waves, the force modes at 3,
    // 4 and 5 (pairs_cfg_valid has admitted only these).
    auto go = [&](auto m, auto b, auto w) {
        constexpr int M = decltype(m)::value, BL = decltype(b)::value, W = decltype(w)::value;
        if (rows)
            k_pairs_b<M, BL, W, true><<<grid, BL, 0, st>>>(items, n_items, queue, scenes, d_num, pos_s, start, end,
                                                           keys, ve, vel_s, idx, pos_u, vel_u, veleval_u, force_out,
                                                           xsph_out, sub, rows, rows_stride);
        else
            k_pairs_b<M, BL, W, false><<<grid, BL, 0, st>>>(items, n_items, queue, scenes, d_num, pos_s, start, end,
                                                            keys, ve, vel_s, idx, pos_u, vel_u, veleval_u, force_out,
                                                            xsph_out, sub, nullptr, 0u);
    };
    auto by_waves = [&](auto m, auto b) {
        if constexpr (decltype(m)::value == kPairsDensity) {
            if (cfg.waves == 4) go(m, b, std::integral_constant<int, 4>{});
            else if (cfg.waves == 5) go(m, b, std::integral_constant<int, 5>{});
            else go(m, b, std::integral_constant<int, 6>{});
        } else {
            if (cfg.waves == 3) go(m, b, std::integral_constant<int, 3>{});
            else if (cfg.waves == 4) go(m, b, std::integral_constant<int, 4>{});
            else go(m, b, std::integral_constant<int, 5>{});
        }
    };
    auto by_block = [&](auto m) {
        if (cfg.block == 256) by_waves(m, std::integral_constant<int, 256>{});
        else if (cfg.block == 128) by_waves(m, std::integral_constant<int, 128>{});
        else by_waves(m, std::integral_constant<int, 64>{});
    };
    switch (mode) {
        case kPairsDensity: by_block(std::integral_constant<int, kPairsDensity>{}); break;
        case kPairsForceStep: by_block(std::integral_constant<int, kPairsForceStep>{}); break;
        case kPairsForceStore: by_block(std::integral_constant<int, kPairsForceStore>{}); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_commit_b(int S, const SceneDev* scenes, int* d_num, const uint32_t* start, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    k_commit_b<<<(S + 63) / 64, 64, 0, st>>>(S, scenes, d_num, start);
    return hipGetLastError();
}

}  // namespace rtps
