// Triangle collision: CollisionTriangle::execute (Collision_triangle.cpp)
// with collision_tri.cl:120-231 and 418-470.  For every particle and every
// loaded triangle, the segment pos +- dist * v/|v| (dist = boundary_distance)
// is intersected with the triangle (Ericson, Real-Time Collision Detection,
// pp. 191-192); on a hit at tp, penetration = dist - |pos - tp| and, when it
// exceeds 1e-6, the repulsion and friction of cl_sph_collision.h:31-73 along
// the triangle's unit normal are added to the particle's force.
//
// Per triangle the host prepares, in the simulation scale, its vertices
// a, b, c (= v * scale), n = (b - a) x (c - a), the unit normal n / |n| and a
// bounding sphere.  A particle farther than radius + dist from the sphere's
// centre cannot get a hit closer than dist (tp lies in the triangle), so it
// skips the test; that changes no result.  A block keeps only the triangles
// whose widened sphere meets its particles' bounding box, in LDS, in order.

#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "rtps_internal.h"
#include "sph_device.h"

namespace rtps {

namespace {

constexpr int kBlockTri = 256;
constexpr int kWavesTri = kBlockTri / 64;
constexpr int kTriChunk = 256;   // triangles tested per chunk (one per thread); 20 KiB of TriPrep in LDS

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Blocks of consecutive sorted particles are compact in space: per chunk of
// triangles, each thread tests one triangle's reach sphere against the
// block's bounding box, and the candidates are compacted into LDS in triangle
// order (ballot ranks, then per-wave offsets), so every particle visits the
// same triangles in the same order as without the cull.
__global__ __launch_bounds__(kBlockTri) void k_collision_tri(int num_host, const int* __restrict__ d_num,
                                                             const uint32_t* __restrict__ start, int nb_cells,
                                                             const float4* __restrict__ pos_s,
                                                             const float4* __restrict__ vel_s,
                                                             float4* __restrict__ force,
                                                             const TriPrep* __restrict__ tris,
                                                             const float4* __restrict__ reach, int ntri,
                                                             rtps_sph_params sp) {
    __shared__ TriPrep lt[kTriChunk];
    __shared__ float box[kWavesTri][6];
    __shared__ int wave_n[kWavesTri];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    int num = d_num ? *d_num : num_host;
    if (d_num) effective_num(num, start, nb_cells, num);
    const bool live = i < num;
    const float s = sp.simulation_scale;
    const float dist = sp.boundary_distance;
    const float k = sp.boundary_stiffness, damp = sp.boundary_dampening;
    float4 p = make_float4(0.f, 0.f, 0.f, 0.f), v = p;
    if (live) {
        p = pos_s[i];
        v = vel_s[i];
    }
    const float px = p.x * s, py = p.y * s, pz = p.z * s;
    // bounding box of the block's live particles
    {
        const float inf = __builtin_inff();
        const float lo[3] = {live ? px : inf, live ? py : inf, live ? pz : inf};
        const float hi[3] = {live ? px : -inf, live ? py : -inf, live ? pz : -inf};
        for (int c = 0; c < 3; ++c) {
            const float a = wave_min(lo[c]), b = wave_max(hi[c]);
            if (lane == 0) {
                box[wave][c] = a;
                box[wave][3 + c] = b;
            }
        }
    }
    __syncthreads();
    float bl[3], bh[3];
    for (int c = 0; c < 3; ++c) {
        bl[c] = box[0][c];
        bh[c] = box[0][3 + c];
        for (int w = 1; w < kWavesTri; ++w) {
            bl[c] = fminf(bl[c], box[w][c]);
            bh[c] = fmaxf(bh[c], box[w][3 + c]);
        }
    }
    // v3normalize (collision_tri.cl:46-50): NaN for v = 0, and then no hit
    const float mag = __builtin_sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    const float vnx = v.x / mag, vny = v.y / mag, vnz = v.z / mag;
    // segment ends p = pos - vn dist, q = pos + vn dist (collision_tri.cl:135-137)
    const float sx = px - vnx * dist, sy = py - vny * dist, sz = pz - vnz * dist;
    const float qx = px + vnx * dist, qy = py + vny * dist, qz = pz + vnz * dist;
    const float qpx = sx - qx, qpy = sy - qy, qpz = sz - qz;
    float fx = 0.f, fy = 0.f, fz = 0.f;
    for (int c0 = 0; c0 < ntri; c0 += kTriChunk) {
        const int t = c0 + (int)threadIdx.x;
        bool cand = false;
        if (t < ntri) {
            const float4 r = reach[t];   // centre, reach squared
            const float dx = fmaxf(fmaxf(bl[0] - r.x, r.x - bh[0]), 0.f);
            const float dy = fmaxf(fmaxf(bl[1] - r.y, r.y - bh[1]), 0.f);
            const float dz = fmaxf(fmaxf(bl[2] - r.z, r.z - bh[2]), 0.f);
            cand = dx * dx + dy * dy + dz * dz <= r.w;
        }
        const unsigned long long m = __ballot(cand);
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        __syncthreads();   // the previous chunk's triangles are no longer read
        if (lane == 0) wave_n[wave] = __popcll(m);
        __syncthreads();
        int base = 0, total = 0;
        for (int w = 0; w < kWavesTri; ++w) {
            if (w < wave) base += wave_n[w];
            total += wave_n[w];
        }
        if (cand) lt[base + rank] = tris[t];
        __syncthreads();
        if (!live) continue;
        for (int q = 0; q < total; ++q) {
            const TriPrep& T = lt[q];
            const float cx = px - T.cx, cy = py - T.cy, cz = pz - T.cz;
            if (cx * cx + cy * cy + cz * cz > T.reach2) continue;   // no hit within dist
            // Ericson segment-triangle test (collision_tri.cl:120-170)
            const float d = qpx * T.nx + qpy * T.ny + qpz * T.nz;
            if (d <= 0.0f) continue;
            const float apx = sx - T.ax, apy = sy - T.ay, apz = sz - T.az;
            float tt = apx * T.nx + apy * T.ny + apz * T.nz;
            if (tt < 0.0f || tt > d) continue;
            const float ex = qpy * apz - qpz * apy, ey = qpz * apx - qpx * apz, ez = qpx * apy - qpy * apx;
            const float abx = T.bx - T.ax, aby = T.by - T.ay, abz = T.bz - T.az;
            const float acx = T.qx - T.ax, acy = T.qy - T.ay, acz = T.qz - T.az;
            float vv = acx * ex + acy * ey + acz * ez;
            if (vv < 0.0f || vv > d) continue;
            float ww = -(abx * ex + aby * ey + abz * ez);
            if (ww < 0.0f || vv + ww > d) continue;
            const float ood = 1.0f / d;
            vv *= ood;
            ww *= ood;
            const float uu = 1.0f - vv - ww;
            const float tx = uu * T.ax + vv * T.bx + ww * T.qx;   // tp = u a + v b + w c
            const float ty = uu * T.ay + vv * T.by + ww * T.qy;
            const float tz = uu * T.az + vv * T.bz + ww * T.qz;
            const float ddx = px - tx, ddy = py - ty, ddz = pz - tz;
            const float pen = dist - __builtin_sqrtf(ddx * ddx + ddy * ddy + ddz * ddz);
            if (!(pen > 1e-6f)) continue;   // collision_tri.cl:196
            // calculateRepulsionForce + calculateFrictionForce (zero coefficients)
            const float ndv = T.ux * v.x + T.uy * v.y + T.uz * v.z;
            const float m2 = k * pen - damp * ndv;
            fx += m2 * T.ux;
            fy += m2 * T.uy;
            fz += m2 * T.uz;
            const float vtx = v.x - v.x * ndv, vty = v.y - v.y * ndv, vtz = v.z - v.z * ndv;
            if ((vtx + vty + vtz) / 3.0f > 0.0f) {
                fx -= vtx;
                fy -= vty;
                fz -= vtz;
            }
        }
    }
    if (!live) return;
    float4 f = force[i];
    f.x += fx;
    f.y += fy;
    f.z += fz;
    force[i] = f;
}

}  // namespace

std::vector<TriPrep> prepare_triangles(const rtps_triangle* tris, int n, float scale) {
#pragma clang fp contract(off)
    std::vector<TriPrep> out((size_t)n);
    for (int q = 0; q < n; ++q) {
        const rtps_triangle& T = tris[q];
        TriPrep& P = out[(size_t)q];
        const float ax = T.verts[0].x * scale, ay = T.verts[0].y * scale, az = T.verts[0].z * scale;
        const float bx = T.verts[1].x * scale, by = T.verts[1].y * scale, bz = T.verts[1].z * scale;
        const float cx = T.verts[2].x * scale, cy = T.verts[2].y * scale, cz = T.verts[2].z * scale;
        P.ax = ax; P.ay = ay; P.az = az;
        P.bx = bx; P.by = by; P.bz = bz;
        P.qx = cx; P.qy = cy; P.qz = cz;
        const float abx = bx - ax, aby = by - ay, abz = bz - az;
        const float acx = cx - ax, acy = cy - ay, acz = cz - az;
        P.nx = aby * acz - abz * acy;   // cross_product(ab, ac), collision_tri.cl:139
        P.ny = abz * acx - abx * acz;
        P.nz = abx * acy - aby * acx;
        const float nm = std::sqrt(P.nx * P.nx + P.ny * P.ny + P.nz * P.nz);
        P.ux = P.nx / nm; P.uy = P.ny / nm; P.uz = P.nz / nm;
        // bounding sphere: the centroid and the farthest vertex, widened by
        // the boundary distance later and by a margin for rounding
        const double mx = ((double)ax + bx + cx) / 3.0, my = ((double)ay + by + cy) / 3.0,
                     mz = ((double)az + bz + cz) / 3.0;
        double r2 = 0.0;
        const double vx[3] = {ax, bx, cx}, vy[3] = {ay, by, cy}, vz[3] = {az, bz, cz};
        for (int k = 0; k < 3; ++k)
            r2 = std::max(r2, (vx[k] - mx) * (vx[k] - mx) + (vy[k] - my) * (vy[k] - my) + (vz[k] - mz) * (vz[k] - mz));
        P.cx = (float)mx; P.cy = (float)my; P.cz = (float)mz;
        P.radius = (float)std::sqrt(r2);
        P.reach2 = 0.0f;   // set by set_triangle_reach once the boundary distance is known
    }
    return out;
}

void set_triangle_reach(std::vector<TriPrep>& t, float dist) {
    for (TriPrep& P : t) {
        const double r = ((double)P.radius + dist) * (1.0 + 1e-5) + 1e-6;
        P.reach2 = (float)(r * r);
    }
}

std::vector<float4> triangle_reach(const std::vector<TriPrep>& t) {
    std::vector<float4> out(t.size());
    for (size_t q = 0; q < t.size(); ++q) out[q] = make_float4(t[q].cx, t[q].cy, t[q].cz, t[q].reach2);
    return out;
}

hipError_t launch_collision_tri(int num, const int* d_num, const uint32_t* start, int nb_cells, const float4* pos_s,
                                const float4* vel_s, float4* force, const TriPrep* tris, const float4* reach, int ntri,
                                const rtps_sph_params& sp, hipStream_t st) {
    (void)hipGetLastError();
    if (num <= 0 || ntri <= 0) return hipSuccess;
    k_collision_tri<<<(num + kBlockTri - 1) / kBlockTri, kBlockTri, 0, st>>>(num, d_num, start, nb_cells, pos_s, vel_s,
                                                                           force, tris, reach, ntri, sp);
    return hipGetLastError();
}

}  // namespace rtps
