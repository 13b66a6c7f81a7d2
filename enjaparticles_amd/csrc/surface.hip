// Colour field and iso-surface extraction (SPH::postProcess, SPH.cpp:338-386).
//
// Colour field (colorfield.cl:86-125, ColorField.cpp): node (s, t, r) of a
// res^3 lattice spanning the scaled boundary box takes
//     mass * sum_j Wspiky(max(|x - x_j|, eps), h) / rho_j,   |x - x_j| <= h,
// Wspiky(r, h) = 15 / (pi h^6) (h - r)^3 (cl_sph_kernels.h:71-78), over the
// 27 cells around the node, with the neighbour rows of the density kernel.
// Nodes are stored x fastest, then y, then z (a plain 3-D array instead of
// the reference's 2-D image of slices).
//
// Iso-surface (MarchingCubes.cpp with classifycubes.cl / constructhp.cl /
// traversehp.cl): the reference classifies cubes into a 2-D histogram
// pyramid and walks the pyramid once per output triangle.  Here the cubes
// are classified into per-cube triangle counts, an exclusive scan gives each
// cube its first triangle, and one thread per cube writes its triangles:
// vertices interpolated linearly along the crossing edges, normals from the
// interpolated negative central differences of the field (pointing away from
// the fluid), as traversehp.cl:470-515 computes them.  The case table, the
// corner bits and the edge ends are the reference's (mc_table.h); a cube's
// triangles come out in the table's order, cubes in lattice order (x fastest).
// Vertices are in world coordinates of the boundary box, or in the
// reference's frame (lattice coordinates / (res - 1) * 10, traversehp.cl:512).

#include <hip/hip_runtime.h>

#include <mutex>

#include "mc_table.h"
#include "rtps_internal.h"
#include "sph_device.h"

namespace rtps {

namespace {

constexpr int kBlockSurf = 256;
constexpr int kScanItems = 8;                            // per thread
constexpr int kScanTile = kBlockSurf * kScanItems;       // 2048 counts per tile

inline int blocks_of(long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

// Corner offsets of the case bits and the edges' two ends (mc_table.h).
__constant__ int8_t c_mc_corner[8][3];
__constant__ int8_t c_mc_edge[12][6];

namespace {

template <bool SCALED>
__global__ __launch_bounds__(kBlockSurf) void k_color_field(int res, const float4* __restrict__ pos_s,
                                                            const float4* __restrict__ ve_s,
                                                            const uint32_t* __restrict__ start,
                                                            const uint32_t* __restrict__ end, rtps_sph_params sp,
                                                            Grid g, Win win, rtps_float4 bmin, rtps_float4 bmax,
                                                            float* __restrict__ field) {
    const long n = (long)res * res * res;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int s = (int)(idx % res), t = (int)((idx / res) % res), r = (int)(idx / ((long)res * res));
    // texPos = (s, t, r) / (res - 1) * (bnd_max - bnd_min) + bnd_min (colorfield.cl:104-108)
    const float tmp = 1.0f / (float)(res - 1);
    const float px = (float)s * tmp * (bmax.x - bmin.x) + bmin.x;
    const float py = (float)t * tmp * (bmax.y - bmin.y) + bmin.y;
    const float pz = (float)r * tmp * (bmax.z - bmin.z) + bmin.z;
    const float scale = sp.simulation_scale;
    const float h = sp.smoothing_distance, h2 = h * h, eps = sp.EPSILON;
    const float h6 = h * h * h * h * h * h;
    const float alpha = 15.0f / 3.14159265f / h6;
    const int3 c = grid_cell(px, py, pz, g);
    const WindowCtx wc = window_ctx(px, py, pz, g, c, h, win.keys, win.rx, win.ry, win.rz);
    float sum = 0.0f;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            uint32_t a, b;
            row_window(wc, c, dy, dz, h2, g, start, end, win.keys, a, b, 0u, win.sub);
            // colorfield.cl:63-75 per candidate, kGroup candidates per
            // iteration with their position loads issued together
            auto term = [&](float4 pj, uint32_t j) {
                if (SCALED) { pj.x *= scale; pj.y *= scale; pj.z *= scale; }
                const float rx = px - pj.x, ry = py - pj.y, rz = pz - pj.z;
                const float r2 = rx * rx + ry * ry + rz * rz;
                if (r2 <= h2) {
                    const float rlen = fmaxf(__builtin_sqrtf(r2), eps);
                    const float hr = h - rlen;
                    sum += alpha * hr * hr * hr / ve_s[j].w;
                }
            };
            uint32_t j = a;
            for (; j + kGroup <= b; j += kGroup) {
                const float4* __restrict__ pg = pos_s + j;
                float4 q[kGroup];
#pragma unroll
                for (int k = 0; k < kGroup; ++k) q[k] = pg[k];
#pragma unroll
                for (int k = 0; k < kGroup; ++k) term(q[k], j + k);
            }
            for (; j < b; ++j) term(pos_s[j], j);
        }
    field[idx] = sum * sp.mass;
}

__device__ __forceinline__ float node(const float* __restrict__ f, int res, int x, int y, int z) {
    return f[((long)z * res + y) * res + x];
}

// Case of every cube (bit k set when the node at corner offset
// kMcCorner[k] exceeds the iso level, classifycubes.cl:44-55) and its
// triangle count (nrOfTriangles, classifycubes.cl:37).
__global__ __launch_bounds__(kBlockSurf) void k_mc_classify(int res, float iso, const float* __restrict__ f,
                                                            const uint8_t* __restrict__ ntri,
                                                            uint8_t* __restrict__ cases, uint32_t* __restrict__ counts) {
    const int m = res - 1;
    const long ncubes = (long)m * m * m;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncubes) return;
    const int x = (int)(idx % m), y = (int)((idx / m) % m), z = (int)(idx / ((long)m * m));
    uint32_t cs = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        cs |= (uint32_t)(node(f, res, x + c_mc_corner[k][0], y + c_mc_corner[k][1], z + c_mc_corner[k][2]) > iso) << k;
    cases[idx] = (uint8_t)cs;
    counts[idx] = ntri[cs];
}

// Exclusive scan of n counts in tiles of kScanTile: within-tile offsets, and
// each tile's total in sums[tile].
__global__ __launch_bounds__(kBlockSurf) void k_scan_tiles(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                           uint32_t* __restrict__ sums, long n) {
    __shared__ uint32_t part[kBlockSurf];
    const long base = (long)blockIdx.x * kScanTile + (long)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        tot += v[k];
    }
    part[threadIdx.x] = tot;
    __syncthreads();
    // inclusive Hillis-Steele scan of the per-thread totals
    for (int d = 1; d < kBlockSurf; d <<= 1) {
        const uint32_t add = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - tot;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == kBlockSurf - 1) sums[blockIdx.x] = part[threadIdx.x];
}

// Exclusive scan of the tile totals in place, one block walking them in
// chunks; the grand total goes to *total.
__global__ __launch_bounds__(kBlockSurf) void k_scan_sums(uint32_t* __restrict__ sums, int ntiles,
                                                          uint32_t* __restrict__ total) {
    __shared__ uint32_t part[kBlockSurf];
    uint32_t carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += kBlockSurf) {
        const int i = c0 + (int)threadIdx.x;
        const uint32_t v = i < ntiles ? sums[i] : 0u;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int d = 1; d < kBlockSurf; d <<= 1) {
            const uint32_t add = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0u;
            __syncthreads();
            part[threadIdx.x] += add;
            __syncthreads();
        }
        if (i < ntiles) sums[i] = carry + part[threadIdx.x] - v;
        carry += part[kBlockSurf - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__device__ __forceinline__ float3 neg_gradient(const float* __restrict__ f, int res, int x, int y, int z) {
    // traversehp.cl:470-502: -(f(p + e) - f(p - e)) per axis, indices clamped to the lattice
    const int xp = min(x + 1, res - 1), xm = max(x - 1, 0);
    const int yp = min(y + 1, res - 1), ym = max(y - 1, 0);
    const int zp = min(z + 1, res - 1), zm = max(z - 1, 0);
    return make_float3(node(f, res, xm, y, z) - node(f, res, xp, y, z), node(f, res, x, ym, z) - node(f, res, x, yp, z),
                       node(f, res, x, y, zm) - node(f, res, x, y, zp));
}

// Triangles of every cube at its scanned offset (at most `capacity` in all):
// 9 floats per triangle in `verts` and `normals`; frame 0 = world
// coordinates of [wmin, wmax], 1 = the reference's lattice / (res - 1) * 10.
__global__ __launch_bounds__(kBlockSurf) void k_mc_emit(int res, float iso, const float* __restrict__ f,
                                                        const uint8_t* __restrict__ cases,
                                                        const uint32_t* __restrict__ counts,
                                                        const uint32_t* __restrict__ offs,
                                                        const uint32_t* __restrict__ tile_offs,
                                                        const uint64_t* __restrict__ words, rtps_float4 wmin,
                                                        rtps_float4 wmax, int frame, long capacity,
                                                        float* __restrict__ verts, float* __restrict__ normals) {
    const int m = res - 1;
    const long ncubes = (long)m * m * m;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= ncubes) return;
    const uint32_t nt = counts[idx];
    if (nt == 0) return;
    const int x = (int)(idx % m), y = (int)((idx / m) % m), z = (int)(idx / ((long)m * m));
    const long first = (long)offs[idx] + tile_offs[idx / kScanTile];
    const uint64_t w = words[cases[idx]];
    const float inv = 1.0f / (float)(res - 1);
    for (uint32_t q = 0; q < nt; ++q) {
        const long o = first + q;
        if (o >= capacity) return;
        for (int v = 0; v < 3; ++v) {
            const int e = (int)((w >> (4 * (3 * q + v))) & 0xFu);
            const int8_t* ee = c_mc_edge[e];
            const int ax = x + ee[0], ay = y + ee[1], az = z + ee[2];
            const int bx = x + ee[3], by = y + ee[4], bz = z + ee[5];
            const float v0 = node(f, res, ax, ay, az), v1 = node(f, res, bx, by, bz);
            const float d = (iso - v0) / (v1 - v0);   // traversehp.cl:506-511
            const float gx = (float)ax + d * (float)(bx - ax);
            const float gy = (float)ay + d * (float)(by - ay);
            const float gz = (float)az + d * (float)(bz - az);
            const float3 n0 = neg_gradient(f, res, ax, ay, az), n1 = neg_gradient(f, res, bx, by, bz);
            float nx = n0.x + d * (n1.x - n0.x), ny = n0.y + d * (n1.y - n0.y), nz = n0.z + d * (n1.z - n0.z);
            const float len2 = nx * nx + ny * ny + nz * nz;
            const float il = len2 > 0.0f ? __builtin_amdgcn_rsqf(len2) : 0.0f;
            float* pv = verts + (o * 3 + v) * 3;
            float* pn = normals + (o * 3 + v) * 3;
            if (frame == 1) {   // mix(p0, p1, d) / (res - 1) * 10 (traversehp.cl:512)
                const float rm = (float)(res - 1);
                pv[0] = gx / rm * 10.0f;
                pv[1] = gy / rm * 10.0f;
                pv[2] = gz / rm * 10.0f;
            } else {
                pv[0] = wmin.x + gx * inv * (wmax.x - wmin.x);
                pv[1] = wmin.y + gy * inv * (wmax.y - wmin.y);
                pv[2] = wmin.z + gz * inv * (wmax.z - wmin.z);
            }
            pn[0] = nx * il;
            pn[1] = ny * il;
            pn[2] = nz * il;
        }
    }
}

}  // namespace

size_t surface_scan_tiles(long ncubes) { return (size_t)((ncubes + kScanTile - 1) / kScanTile); }

hipError_t launch_color_field(int res, const float4* pos_s, const float4* ve_s, const uint32_t* start,
                              const uint32_t* end, const rtps_sph_params& sp, const Grid& g, const Win& win,
                              const rtps_float4& bmin_scaled, const rtps_float4& bmax_scaled, float* field,
                              hipStream_t st) {
    (void)hipGetLastError();
    const long n = (long)res * res * res;
    if (res < 2) return hipErrorInvalidValue;
    if (sp.simulation_scale != 1.0f)
        k_color_field<true><<<blocks_of(n, kBlockSurf), kBlockSurf, 0, st>>>(res, pos_s, ve_s, start, end, sp, g, win,
                                                                            bmin_scaled, bmax_scaled, field);
    else
        k_color_field<false><<<blocks_of(n, kBlockSurf), kBlockSurf, 0, st>>>(res, pos_s, ve_s, start, end, sp, g, win,
                                                                             bmin_scaled, bmax_scaled, field);
    return hipGetLastError();
}

hipError_t launch_mc_count(int res, float iso, const float* field, const uint8_t* ntri, uint8_t* cases,
                           uint32_t* counts, uint32_t* offs, uint32_t* tile_sums, uint32_t* total, hipStream_t st) {
    (void)hipGetLastError();
    if (res < 2) return hipErrorInvalidValue;
    const long m = res - 1, ncubes = m * m * m;
    const int ntiles = (int)surface_scan_tiles(ncubes);
    k_mc_classify<<<blocks_of(ncubes, kBlockSurf), kBlockSurf, 0, st>>>(res, iso, field, ntri, cases, counts);
    k_scan_tiles<<<ntiles, kBlockSurf, 0, st>>>(counts, offs, tile_sums, ncubes);
    k_scan_sums<<<1, kBlockSurf, 0, st>>>(tile_sums, ntiles, total);
    return hipGetLastError();
}

hipError_t mc_upload_geometry() {
    static std::mutex mu;
    static bool done[64] = {};
    int dev = 0;
    hipError_t err = hipGetDevice(&dev);
    if (err != hipSuccess) return err;
    if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(mu);
    if (done[dev]) return hipSuccess;   // each device holds its own constant memory
    err = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_corner), kMcCorner, sizeof(kMcCorner));
    if (err == hipSuccess) err = hipMemcpyToSymbol(HIP_SYMBOL(c_mc_edge), kMcEdgeEnds, sizeof(kMcEdgeEnds));
    if (err == hipSuccess) done[dev] = true;
    return err;
}

hipError_t launch_mc_emit(int res, float iso, const float* field, const uint8_t* cases, const uint32_t* counts,
                          const uint32_t* offs, const uint32_t* tile_offs, const uint64_t* words,
                          const rtps_float4& wmin, const rtps_float4& wmax, int frame, long capacity, float* verts,
                          float* normals, hipStream_t st) {
    (void)hipGetLastError();
    if (res < 2) return hipErrorInvalidValue;
    const long m = res - 1, ncubes = m * m * m;
    k_mc_emit<<<blocks_of(ncubes, kBlockSurf), kBlockSurf, 0, st>>>(res, iso, field, cases, counts, offs, tile_offs,
                                                                   words, wmin, wmax, frame, capacity, verts,
                                                                   normals);
    return hipGetLastError();
}

}  // namespace rtps
