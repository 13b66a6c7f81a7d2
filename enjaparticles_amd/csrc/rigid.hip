// Rigid-body coupling of SPH::interact (rtpslib/system/SPH.cpp:610-663): the
// fluid against the particles of a rigid-body system, found through that
// system's own cell tables on the fluid's grid.
//
// A set of rigid-body particles (positions, and for a moving body velocities
// and masses) is hashed into the fluid's cell grid, sorted and given cell
// start/end tables with the same kernels as the fluid (System::hash_and_sort,
// System.cpp:145-170, on the rigid system).  k_rigid_force then adds, for each
// sorted fluid particle i and every set particle j in the 27 cells around i's
// cell with |r| <= 2h (cl_sph_neighbors.h:99-140 iteration, 2h cut-off):
//   norm = (x_j - x_i) / max(|r|, eps), stiff = s0 m
//   F += -stiff (2h - |r|) norm + s1 sqrt(stiff mn) ((vrel . norm) norm)
//   V += w (h - |r|)          (Wvisc_lapl, negative between h and 2h)
// static set  (staticrigidbody_force.cl:45-146): vrel = -e_i, mn = m / 2,
//             w = vrel, force_i += F / m - m mu c V / rho_i^2
// moving set  (rigidbody_force.cl:118-193): vrel = v_j - e_i,
//             mn = m m_j / (m + m_j), w = tangential part of vrel,
//             force_i += F / m + m mu c V / rho_i^2
// with e_i the fluid particle's veleval and c = wvisc_dd_coef.

#include <hip/hip_runtime.h>

#include "rtps_internal.h"
#include "sph_device.h"

namespace rtps {

namespace {
constexpr int kRBlock = 256;
inline int rblocks(long n) { return (int)((n + kRBlock - 1) / kRBlock); }
}  // namespace

__global__ __launch_bounds__(kRBlock) void k_rigid_gather(int n, const uint32_t* __restrict__ idx,
                                                          const float4* __restrict__ pu,
                                                          const float4* __restrict__ vu,
                                                          const float* __restrict__ mu, float4* __restrict__ ps,
                                                          float4* __restrict__ vs, float* __restrict__ ms) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = idx[i];
    ps[i] = pu[s];
    if (vs) vs[i] = vu[s];
    if (ms) ms[i] = mu[s];
}

template <bool MOVING>
__global__ __launch_bounds__(kRBlock) void k_rigid_force(int num_host, const int* __restrict__ d_num,
                                                         const uint32_t* __restrict__ fstart, int nb_cells,
                                                         const float4* __restrict__ pos_s,
                                                         const float4* __restrict__ ve_s, float4* __restrict__ force,
                                                         const float4* __restrict__ bpos,
                                                         const float4* __restrict__ bvel,
                                                         const float* __restrict__ bmass,
                                                         const uint32_t* __restrict__ bstart,
                                                         const uint32_t* __restrict__ bend, rtps_sph_params sp,
                                                         Grid gs, float s0, float s1) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int num = d_num ? *d_num : num_host;
    if (d_num) effective_num(num, fstart, nb_cells, num);
    if (i >= num) return;
    const float scale = sp.simulation_scale, h = sp.smoothing_distance, eps = sp.EPSILON, m = sp.mass;
    const float h2x = 2.0f * h;
    const float4 p = pos_s[i];
    const float px = p.x * scale, py = p.y * scale, pz = p.z * scale;
    const float4 e = ve_s[i];   // veleval.xyz | density
    const float stiff = s0 * m;
    float fx = 0.f, fy = 0.f, fz = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
    const int3 c = grid_cell(px, py, pz, gs);
    for (int z = c.z - 1; z <= c.z + 1; ++z)
        for (int y = c.y - 1; y <= c.y + 1; ++y)
            for (int x = c.x - 1; x <= c.x + 1; ++x) {
                const uint32_t ch = (uint32_t)grid_hash(x, y, z, gs);
                if (ch >= (uint32_t)nb_cells) continue;
                const uint32_t a = bstart[ch];
                if (a == kEmpty) continue;
                const uint32_t b = bend[ch];
                for (uint32_t j = a; j < b; ++j) {
                    const float4 q = bpos[j];
                    const float rx = q.x * scale - px, ry = q.y * scale - py, rz = q.z * scale - pz;
                    float rl = __builtin_sqrtf(rx * rx + ry * ry + rz * rz);
                    if (!(rl <= h2x)) continue;
                    rl = fmaxf(rl, eps);
                    const float inv = 1.0f / rl;
                    const float nx = rx * inv, ny = ry * inv, nz = rz * inv;
                    float ux = -e.x, uy = -e.y, uz = -e.z, mn;
                    if (MOVING) {
                        const float4 v = bvel[j];
                        ux += v.x;
                        uy += v.y;
                        uz += v.z;
                        const float mj = bmass[j];
                        mn = (m * mj) / (m + mj);
                    } else {
                        mn = (m * m) / (m + m);
                    }
                    const float un = ux * nx + uy * ny + uz * nz;
                    const float spr = -stiff * (h2x - rl);
                    const float dmp = s1 * __builtin_sqrtf(stiff * mn);
                    fx += spr * nx + dmp * (un * nx);
                    fy += spr * ny + dmp * (un * ny);
                    fz += spr * nz + dmp * (un * nz);
                    const float lap = h - rl;
                    if (MOVING) {
                        ux -= un * nx;
                        uy -= un * ny;
                        uz -= un * nz;
                    }
                    vx += ux * lap;
                    vy += uy * lap;
                    vz += uz * lap;
                }
            }
    const float cv = m * (1.0f / (e.w * e.w)) * (sp.viscosity * sp.wvisc_dd_coef);
    const float sg = MOVING ? cv : -cv;
    float4 f = force[i];
    f.x += fx / m + sg * vx;
    f.y += fy / m + sg * vy;
    f.z += fz / m + sg * vz;
    force[i] = f;
}

hipError_t rigid_prepare(int n, const float4* pos_u, const float4* vel_u, const float* mass_u, RigidDev& r,
                         const Grid& g, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    const int bits = sort_key_bits(g.nb_cells);
    const uint32_t pad = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    hipError_t e = launch_hash(n, nullptr, n, pad, pos_u, r.keys, r.idx, g, r.start, g.nb_cells + 1, 0, st);
    if (e != hipSuccess) return e;
    uint32_t *ks = r.keys, *is = r.idx;
    e = sort_pairs_pingpong(n, bits, r.keys, r.idx, r.sort_tmp, st, &ks, &is);
    if (e != hipSuccess) return e;
    e = launch_cellindices(n, nullptr, ks, g.nb_cells, r.start, r.end, 0, st);
    if (e != hipSuccess) return e;
    if (n > 0)
        k_rigid_gather<<<rblocks(n), kRBlock, 0, st>>>(n, is, pos_u, vel_u, mass_u, r.pos, vel_u ? r.vel : nullptr,
                                                       mass_u ? r.mass : nullptr);
    return hipGetLastError();
}

hipError_t launch_rigid_force(int num, const int* d_num, const uint32_t* fstart, int nb_cells, const float4* pos_s,
                              const float4* ve_s, float4* force, const RigidDev& r, const rtps_sph_params& sp,
                              const Grid& gs, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0 || r.n <= 0) return hipSuccess;
    if (r.moving)
        k_rigid_force<true><<<rblocks(num), kRBlock, 0, st>>>(num, d_num, fstart, nb_cells, pos_s, ve_s, force, r.pos,
                                                              r.vel, r.mass, r.start, r.end, sp, gs, r.s0, r.s1);
    else
        k_rigid_force<false><<<rblocks(num), kRBlock, 0, st>>>(num, d_num, fstart, nb_cells, pos_s, ve_s, force,
                                                               r.pos, nullptr, nullptr, r.start, r.end, sp, gs, r.s0,
                                                               r.s1);
    return hipGetLastError();
}

}  // namespace rtps
