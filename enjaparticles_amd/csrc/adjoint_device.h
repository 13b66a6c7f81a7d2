// Per-particle bodies of the adjoint kernels, shared by the single-system
// launches (adjoint.hip) and the batched ones (batch_adjoint.hip).  No
// reference counterpart; the forward step they differentiate is
// density.cl, force.cl, collision_wall.cl and leapfrog.cl (simulation_scale
// 1, no point sources, neighbour pairs held fixed):
//   rho_i = c_rho sum_j (h^2 - q^2)^3
//   a_i   = sum_j [ c_ij phi(q) r_ij + beta_ij psi(q) (e_j - e_i) ]
//           c_ij = m cp K (rho_i + rho_j - 2 rho0) / (rho_i rho_j), phi = (h-q)^2/q
//           beta_ij = m mu cv / (rho_i rho_j),                 psi = h - q
//   X_i   = cx sum_j (h^2 - q^2)^3 (e_j - e_i) / (rho_i + rho_j)
//   f_i   = clamp(a_i + walls(x_i, v_i) + g)
//   v' = v + dt f + xi X,  e' = (v + v')/2,  x' = x + dt v'
// c_ij, beta_ij and the XSPH weight are symmetric in (i, j), so every pair
// term of the backward pass is a gather over the same neighbour rows:
//   B1  per particle: leapfrog, velocity clamp and walls backwards -> adjoints
//       of a (ga) and X (gX), lambda_v of step n, the direct part of lambda_x
//   B2  pair loop: d/dx, d/de and d/drho of the pressure, viscosity and XSPH
//       terms, plus this particle's share of dL/dmu and dL/dK
//   B3  pair loop: density backwards, d rho / dx with the grho of both ends,
//       then lambda_x, lambda_e of step n scattered to the original order.
#pragma once

#include <hip/hip_runtime.h>

#include "sph_device.h"

namespace rtps {

__device__ __forceinline__ float3 f3(float4 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// Neighbour rows of one launch: cell tables (already offset to the scene's
// cells), sorted keys for the x windows (nullptr = plain rows), grid
// resolution and the key bits shared by a whole row (the scene id of
// segmented keys).
struct AdjRows {
    const uint32_t* start;
    const uint32_t* end;
    const uint32_t* keys;
    float rx, ry, rz;
    uint32_t prefix;
};

// Candidates per LDS copy of a row in B2 (4 records each) and B3 (a position
// and a float each): 6 KB and 5 KB per wave.
constexpr int kAdjStageCap = 96;
constexpr int kAdjB3StageCap = 256;
// LDS slice of the calling wave (float4 units): B2 needs 4 x kAdjStageCap,
// B3 kAdjB3StageCap + kAdjB3StageCap / 4.
constexpr int kAdjStageSlice = 4 * kAdjStageCap > kAdjB3StageCap + kAdjB3StageCap / 4
                                   ? 4 * kAdjStageCap : kAdjB3StageCap + kAdjB3StageCap / 4;

// The wave's live lanes and this lane's rank among them; the union [A, B) of
// the live lanes' row windows (A >= B: no lane has candidates).
struct WaveLanes {
    uint32_t n, rank;
};
__device__ __forceinline__ WaveLanes wave_lanes() {
    const uint64_t live = __ballot(true);
    return WaveLanes{(uint32_t)__popcll(live),
                     (uint32_t)__popcll(live & ((1ull << (threadIdx.x & 63)) - 1ull))};
}
__device__ __forceinline__ void wave_range(uint32_t a, uint32_t b, uint32_t& A, uint32_t& B) {
    A = __ockl_wfred_min_u32(a < b ? a : 0xFFFFFFFFu);
    B = __ockl_wfred_max_u32(a < b ? b : 0u);
}

// B1 for sorted particle i.  lam_x/v/e: adjoints of step n+1 (original
// order).  Writes ga/gX (sorted), lam_v of step n (original order) and the
// direct part of lam_x (sorted).
__device__ __forceinline__ void adj_b1_body(int i, const uint32_t* __restrict__ idx, const float4* __restrict__ pos_s,
                                            const float4* __restrict__ vel_s, const float4* __restrict__ force_s,
                                            const rtps_sph_params& sp, const rtps_float4& bmin,
                                            const rtps_float4& bmax, int collision, float dt,
                                            const float4* __restrict__ lam_x, const float4* __restrict__ lam_v,
                                            const float4* __restrict__ lam_e, float4* __restrict__ ga_s,
                                            float4* __restrict__ gX_s, float4* __restrict__ lam_v_out,
                                            float4* __restrict__ lxp_s) {
    const uint32_t o = idx[i];
    const float3 lx = f3(lam_x[o]), lv = f3(lam_v[o]), le = f3(lam_e[o]);
    // adjoint of v' (x' = x + dt v', e' = (v + v')/2)
    const float3 lvb = make_float3(lv.x + dt * lx.x + 0.5f * le.x, lv.y + dt * lx.y + 0.5f * le.y,
                                   lv.z + dt * lx.z + 0.5f * le.z);
    const float4 p = pos_s[i], v4 = vel_s[i];
    const float3 v = f3(v4);
    // f = force + walls + g, then the velocity clamp
    float4 F = force_s[i];
    if (collision) collision_wall(F, p, v4, sp, bmin, bmax);
    const float3 f = make_float3(F.x + sp.gravity.x, F.y + sp.gravity.y, F.z + sp.gravity.z);
    float3 gfc = make_float3(dt * lvb.x, dt * lvb.y, dt * lvb.z);   // adjoint of the clamped force
    float3 gf = gfc;
    const float speed = sqrtf(dot3(f, f));
    if (speed > sp.velocity_limit) {
        const float k = sp.velocity_limit / speed;
        const float3 u = make_float3(f.x / speed, f.y / speed, f.z / speed);
        const float ug = dot3(u, gfc);
        gf = make_float3(k * (gfc.x - u.x * ug), k * (gfc.y - u.y * ug), k * (gfc.z - u.z * ug));
    }
    const float xi = sp.xsph_factor;
    ga_s[i] = make_float4(gf.x, gf.y, gf.z, 0.f);
    gX_s[i] = make_float4(xi * lvb.x, xi * lvb.y, xi * lvb.z, 0.f);
    float3 gv = make_float3(lvb.x + 0.5f * le.x, lvb.y + 0.5f * le.y, lvb.z + 0.5f * le.z);
    float3 gx = lx;
    if (collision) {
        // active walls: d w / d x = -k s n n^T, d w / d v = -d n n^T, friction -v_t (v_t = v (1 - n.v))
        const float s = sp.simulation_scale, bd = sp.rest_distance * 0.5f;
        const float px = p.x * s, py = p.y * s, pz = p.z * s;
        const float diffs[6] = {bd - (pz - bmin.z), bd - (bmax.z - pz), bd - (py - bmin.y),
                                bd - (bmax.y - py), bd - (px - bmin.x), bd - (bmax.x - px)};
        const float3 ns[6] = {make_float3(0, 0, 1), make_float3(0, 0, -1), make_float3(0, 1, 0),
                              make_float3(0, -1, 0), make_float3(1, 0, 0), make_float3(-1, 0, 0)};
#pragma unroll
        for (int w = 0; w < 6; ++w) {
            if (!(diffs[w] > sp.EPSILON)) continue;
            const float3 n = ns[w];
            const float ng = dot3(n, gf);
            const float kx = -sp.boundary_stiffness * s * ng, kv = -sp.boundary_dampening * ng;
            gx.x += kx * n.x; gx.y += kx * n.y; gx.z += kx * n.z;
            gv.x += kv * n.x; gv.y += kv * n.y; gv.z += kv * n.z;
            const float ndv = dot3(n, v);
            const float3 vt = make_float3(v.x - v.x * ndv, v.y - v.y * ndv, v.z - v.z * ndv);
            if ((vt.x + vt.y + vt.z) / 3.0f > 0.0f) {
                // friction f = -v_t:  (d f / d v)^T g = -(1 - n.v) g + n (v . g)
                const float vg = dot3(v, gf);
                gv.x += -(1.f - ndv) * gf.x + n.x * vg;
                gv.y += -(1.f - ndv) * gf.y + n.y * vg;
                gv.z += -(1.f - ndv) * gf.z + n.z * vg;
            }
        }
    }
    lam_v_out[o] = make_float4(gv.x, gv.y, gv.z, 0.f);
    lxp_s[i] = make_float4(gx.x, gx.y, gx.z, 0.f);
}

// B2 for sorted particle i: pair terms.  ve_s = veleval.xyz | density
// (sorted).  Writes gx2 (d/dx of the pair terms), ge (d/de), grho (sorted);
// adds particle i's terms of dL/dmu and dL/dK to gmu, gK.  Hardware rsq / rcp
// (~1 ulp) replace the square root and the divisions.
__device__ __forceinline__ void adj_b2_body(int i, const float4* __restrict__ pos_s, const float4* __restrict__ ve_s,
                                            const float4* __restrict__ ga_s, const float4* __restrict__ gX_s,
                                            const AdjRows& rows, const rtps_sph_params& sp, const Grid& g,
                                            float4* __restrict__ gx2_s, float4* __restrict__ ge_s,
                                            float* __restrict__ grho_s, float& gmu, float& gK,
                                            float4* __restrict__ stage) {
    const float h = sp.smoothing_distance, h2 = h * h, eps = sp.EPSILON, eps2 = eps * eps;
    const float m = sp.mass, rho0 = sp.rest_density, K = sp.K, mu = sp.viscosity;
    const float cp = -0.5f * sp.wspiky_d_coef, cv = sp.wvisc_dd_coef;
    const float cx = 2.0f * sp.mass * sp.wpoly6_coef;
    const float4 pi = pos_s[i], vei = ve_s[i];
    const float3 ei = f3(vei), gai = f3(ga_s[i]), gXi = f3(gX_s[i]);
    const float ri = vei.w;
    const float inv_ri = __builtin_amdgcn_rcpf(ri);
    float3 gx = make_float3(0, 0, 0), ge = make_float3(0, 0, 0);
    float grho = 0.f;
    // one accepted pair (within_h, j != i)
    auto pair = [&](const float3 r, const float q2, const float4 vej, const float4 gaj4, const float4 gXj4) {
        const float3 ej = f3(vej), gaj = f3(gaj4), gXj = f3(gXj4);
        const float rj = vej.w;
        const float q2c = fmaxf(q2, eps2);
        const float rinv = __builtin_amdgcn_rsqf(q2c);   // 1 / max(|r|, eps)
        const float q = q2c * rinv;                        // max(|r|, eps)
        const float dq = q2 > eps2 ? 1.f : 0.f;            // d q / d |r|
        const float inv_rr = __builtin_amdgcn_rcpf(ri * rj);
        const float3 dga = make_float3(gai.x - gaj.x, gai.y - gaj.y, gai.z - gaj.z);
        const float3 de = make_float3(ej.x - ei.x, ej.y - ei.y, ej.z - ei.z);
        const float3 rh = make_float3(r.x * rinv, r.y * rinv, r.z * rinv);   // used only where dq = 1
        // pressure: T_ij = c phi r
        const float hq = h - q;
        const float phi = hq * hq * rinv;
        const float dphi = -hq * (h + q) * rinv * rinv * dq;   // d phi / d |r|
        const float cK = m * cp * (ri + rj - 2.f * rho0) * inv_rr;   // c / K
        const float cc = K * cK;
        const float rdga = dot3(r, dga);
        const float sig = phi * rdga;
        gx.x += cc * (phi * dga.x + dphi * rdga * rh.x);
        gx.y += cc * (phi * dga.y + dphi * rdga * rh.y);
        gx.z += cc * (phi * dga.z + dphi * rdga * rh.z);
        grho += sig * m * cp * K * (2.f * rho0 - rj) * inv_rr * inv_ri;
        gK += cK * phi * dot3(r, gai);
        // viscosity: V_ij = beta psi (e_j - e_i), psi = h - q
        const float bmu = m * cv * inv_rr;           // beta / mu
        const float beta = mu * bmu;
        const float psi = hq;
        const float3 nga = make_float3(-dga.x, -dga.y, -dga.z);   // ga_j - ga_i
        ge.x += beta * psi * nga.x;
        ge.y += beta * psi * nga.y;
        ge.z += beta * psi * nga.z;
        const float deg = dot3(de, nga);
        gx.x += beta * deg * dq * rh.x;
        gx.y += beta * deg * dq * rh.y;
        gx.z += beta * deg * dq * rh.z;
        grho += -(beta * inv_ri) * psi * dot3(de, dga);
        gmu += bmu * psi * dot3(de, gai);
        // XSPH: Y_ij = cx w6 (e_j - e_i) / D
        const float t = h2 - q2;
        const float w6 = t * t * t;
        const float inv_D = __builtin_amdgcn_rcpf(ri + rj);
        const float3 dgX = make_float3(gXi.x - gXj.x, gXi.y - gXj.y, gXi.z - gXj.z);
        const float yk = cx * w6 * inv_D;
        ge.x -= yk * dgX.x;
        ge.y -= yk * dgX.y;
        ge.z -= yk * dgX.z;
        const float dedgX = dot3(de, dgX);
        const float yr = cx * inv_D * (-6.f) * t * t * dedgX;
        gx.x += yr * r.x;
        gx.y += yr * r.y;
        gx.z += yr * r.z;
        grho += -yk * inv_D * dedgX;
    };
    const int3 c = grid_cell(pi.x, pi.y, pi.z, g);
    const WindowCtx wc = window_ctx(pi.x, pi.y, pi.z, g, c, h, rows.keys, rows.rx, rows.ry, rows.rz);
    // candidate j of a row: P[j - off], V / GA / GX[j - off]
    auto walk = [&](const float4* __restrict__ P, const float4* __restrict__ V, const float4* __restrict__ GA,
                    const float4* __restrict__ GX, uint32_t a, uint32_t b, uint32_t off) {
        uint32_t j = a;
        for (; j + kGroup <= b; j += kGroup) {   // positions, then the accepted records, loaded together
            float4 p[kGroup];
#pragma unroll
            for (int k = 0; k < kGroup; ++k) p[k] = P[j + k - off];
            float3 r[kGroup];
            float q2[kGroup];
            bool ok[kGroup];
#pragma unroll
            for (int k = 0; k < kGroup; ++k) {
                r[k] = make_float3(pi.x - p[k].x, pi.y - p[k].y, pi.z - p[k].z);
                q2[k] = dot3(r[k], r[k]);
                ok[k] = within_h(h2_minus_r2(r[k].x, r[k].y, r[k].z, h2)) && j + k != (uint32_t)i;
            }
            float4 v[kGroup], ga[kGroup], gX[kGroup];
#pragma unroll
            for (int k = 0; k < kGroup; ++k)
                if (ok[k]) {
                    v[k] = V[j + k - off];
                    ga[k] = GA[j + k - off];
                    gX[k] = GX[j + k - off];
                }
#pragma unroll
            for (int k = 0; k < kGroup; ++k)
                if (ok[k]) pair(r[k], q2[k], v[k], ga[k], gX[k]);
        }
        for (; j < b; ++j) {
            const float4 pj = P[j - off];
            const float3 r = make_float3(pi.x - pj.x, pi.y - pj.y, pi.z - pj.z);
            const float q2 = dot3(r, r);
            if (within_h(h2_minus_r2(r.x, r.y, r.z, h2)) && j != (uint32_t)i)
                pair(r, q2, V[j - off], GA[j - off], GX[j - off]);
        }
    };
    // rows of at most kAdjStageCap wave-wide candidates are copied once per
    // wave into `stage` (full-line loads) and read from there
    const WaveLanes wl = wave_lanes();
    float4* __restrict__ sp_ = stage;
    float4* __restrict__ sv = stage + kAdjStageCap;
    float4* __restrict__ sga = stage + 2 * kAdjStageCap;
    float4* __restrict__ sgx = stage + 3 * kAdjStageCap;
    // the staged copies, a lane's window in chunks of up to 32 candidates:
    // the tests of a chunk shift one bit per candidate into a mask (the
    // chunk's first candidate ends in its highest bit), the particle's own
    // bit is cleared, then the pair terms run once per set bit, highest
    // first: the same pairs in the same ascending order as walk (the same
    // sums), with the wave's pair iterations its lanes' largest accepted
    // count instead of their longest window
    auto walk_bits = [&](uint32_t a, uint32_t b, uint32_t off) {
        for (uint32_t j0 = a; j0 < b; j0 += 32u) {
            const uint32_t n = min(b - j0, 32u);
            const float4* __restrict__ pc = sp_ + (j0 - off);
            uint32_t mk = 0u, k = 0u;
            auto test = [&](const float4 p) {
                const float3 r = make_float3(pi.x - p.x, pi.y - p.y, pi.z - p.z);
                mk = mk + mk + (within_h(h2_minus_r2(r.x, r.y, r.z, h2)) ? 1u : 0u);
            };
            for (; k + kGroup <= n; k += kGroup) {
                float4 q[kGroup];
#pragma unroll
                for (int t = 0; t < kGroup; ++t) q[t] = pc[k + t];
#pragma unroll
                for (int t = 0; t < kGroup; ++t) test(q[t]);
            }
            for (; k < n; ++k) test(pc[k]);
            const uint32_t self = (uint32_t)i - j0;   // the particle itself passes (r = 0): drop it
            if (self < n) mk &= ~(1u << (n - 1u - self));
            while (mk != 0u) {
                const uint32_t hb = 31u - (uint32_t)__builtin_clz(mk);
                mk ^= 1u << hb;
                const uint32_t t = n - 1u - hb;
                const float4 p = pc[t];
                const float3 r = make_float3(pi.x - p.x, pi.y - p.y, pi.z - p.z);
                const uint32_t o = j0 + t - off;
                pair(r, dot3(r, r), sv[o], sga[o], sgx[o]);
            }
        }
    };
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            uint32_t a, b, A, B;
            row_window(wc, c, dy, dz, h2, g, rows.start, rows.end, rows.keys, a, b, rows.prefix);
            wave_range(a, b, A, B);
            if (B <= A) continue;
            if (B - A <= (uint32_t)kAdjStageCap) {
                for (uint32_t t = wl.rank; t < B - A; t += wl.n) {
                    sp_[t] = pos_s[A + t];
                    sv[t] = ve_s[A + t];
                    sga[t] = ga_s[A + t];
                    sgx[t] = gX_s[A + t];
                }
                __builtin_amdgcn_wave_barrier();   // the copies before any lane reads them
                walk_bits(a, b, A);
                __builtin_amdgcn_wave_barrier();   // every read before the next row's copies
            } else {
                walk(pos_s, ve_s, ga_s, gX_s, a, b, 0u);
            }
        }
    gx2_s[i] = make_float4(gx.x, gx.y, gx.z, 0.f);
    ge_s[i] = make_float4(ge.x, ge.y, ge.z, 0.f);
    grho_s[i] = grho;
}

// B3 for sorted particle i: density backwards and the scatter of lambda_x,
// lambda_e of step n.
__device__ __forceinline__ void adj_b3_body(int i, const uint32_t* __restrict__ idx, const float4* __restrict__ pos_s,
                                            const float* __restrict__ grho_s, const float4* __restrict__ lxp_s,
                                            const float4* __restrict__ gx2_s, const float4* __restrict__ ge_s,
                                            const AdjRows& rows, const rtps_sph_params& sp, const Grid& g,
                                            float4* __restrict__ lam_x_out, float4* __restrict__ lam_e_out,
                                            float4* __restrict__ stage) {
    const float h = sp.smoothing_distance, h2 = h * h;
    const float crho = sp.wpoly6_coef * sp.mass;
    const float4 pi = pos_s[i];
    const float gri = grho_s[i];
    float3 gx = make_float3(0, 0, 0);
    const int3 c = grid_cell(pi.x, pi.y, pi.z, g);
    const WindowCtx wc = window_ctx(pi.x, pi.y, pi.z, g, c, h, rows.keys, rows.rx, rows.ry, rows.rz);
    auto term = [&](float4 pj, float grj, bool ok) {
        const float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
        const float q2 = rx * rx + ry * ry + rz * rz;
        if (!(ok && q2 <= h2)) return;
        const float t = h2 - q2;
        const float k = crho * (-6.f) * t * t * (gri + grj);
        gx.x += k * rx;
        gx.y += k * ry;
        gx.z += k * rz;
    };
    // candidate j of a row: P[j - off], GR[j - off]
    auto walk = [&](const float4* __restrict__ P, const float* __restrict__ GR, uint32_t a, uint32_t b,
                    uint32_t off) {
        uint32_t j = a;
        for (; j + kGroup <= b; j += kGroup) {   // loads of a group issued together
            float4 p[kGroup];
            float gr[kGroup];
#pragma unroll
            for (int k = 0; k < kGroup; ++k) {
                p[k] = P[j + k - off];
                gr[k] = GR[j + k - off];
            }
#pragma unroll
            for (int k = 0; k < kGroup; ++k) term(p[k], gr[k], j + k != (uint32_t)i);
        }
        for (; j < b; ++j) term(P[j - off], GR[j - off], j != (uint32_t)i);
    };
    // rows of at most kAdjB3StageCap wave-wide candidates are copied once per
    // wave into `stage` (full-line loads) and read from there
    const WaveLanes wl = wave_lanes();
    float4* __restrict__ sp_ = stage;
    float* __restrict__ sgr = (float*)(stage + kAdjB3StageCap);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy) {
            uint32_t a, b, A, B;
            row_window(wc, c, dy, dz, h2, g, rows.start, rows.end, rows.keys, a, b, rows.prefix);
            wave_range(a, b, A, B);
            if (B <= A) continue;
            if (B - A <= (uint32_t)kAdjB3StageCap) {
                for (uint32_t t = wl.rank; t < B - A; t += wl.n) {
                    sp_[t] = pos_s[A + t];
                    sgr[t] = grho_s[A + t];
                }
                __builtin_amdgcn_wave_barrier();   // the copies before any lane reads them
                walk(sp_, sgr, a, b, A);
                __builtin_amdgcn_wave_barrier();   // every read before the next row's copies
            } else {
                walk(pos_s, grho_s, a, b, 0u);
            }
        }
    const float4 lp = lxp_s[i], g2 = gx2_s[i];
    const uint32_t o = idx[i];
    lam_x_out[o] = make_float4(lp.x + g2.x + gx.x, lp.y + g2.y + gx.y, lp.z + g2.z + gx.z, 0.f);
    lam_e_out[o] = ge_s[i];
}

// Forward integration of a recorded step for sorted particle i: walls (on a
// copy of the stored SPH force, which B1 needs unchanged) and leapfrog,
// scattered to the next state.
__device__ __forceinline__ void adj_integrate_body(int i, const float4* __restrict__ pos_s,
                                                   const float4* __restrict__ vel_s,
                                                   const float4* __restrict__ force_s,
                                                   const float4* __restrict__ xsph_s,
                                                   const uint32_t* __restrict__ idx, const rtps_sph_params& sp,
                                                   const rtps_float4& bmin, const rtps_float4& bmax, int collision,
                                                   float dt, float4* __restrict__ pos_u, float4* __restrict__ vel_u,
                                                   float4* __restrict__ veleval_u) {
    const float4 p = pos_s[i], v = vel_s[i];
    float4 F = force_s[i];
    if (collision) collision_wall(F, p, v, sp, bmin, bmax);
    float4 po, vo, eo;
    leapfrog(p, v, F, xsph_s[i], sp, dt, po, vo, eo);
    const uint32_t o = idx[i];
    pos_u[o] = po;
    vel_u[o] = vo;
    veleval_u[o] = eo;
}

// Deterministic wave64 + block sum of (gmu, gK) into partial[slot]; every
// thread of the block must call it.
template <int BLOCK>
__device__ __forceinline__ void adj_block_partial(float gmu, float gK, float2* __restrict__ partial, int slot) {
    __shared__ float2 red[BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) {
        gmu += __shfl_down(gmu, off, 64);
        gK += __shfl_down(gK, off, 64);
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_float2(gmu, gK);
    __syncthreads();
    if (threadIdx.x == 0) {
        float2 t = make_float2(0.f, 0.f);
        for (int w = 0; w < BLOCK / 64; ++w) {
            t.x += red[w].x;
            t.y += red[w].y;
        }
        partial[slot] = t;
    }
}

// Sum of partial[first, first + n) in FP64, fixed order, by one block of 256
// threads; adds the totals to acc[0], acc[1].
__device__ __forceinline__ void adj_reduce_body(const float2* __restrict__ partial, int first, int n,
                                                double* __restrict__ acc) {
    __shared__ double s0[256], s1[256];
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        a += (double)partial[first + k].x;
        b += (double)partial[first + k].y;
    }
    s0[threadIdx.x] = a;
    s1[threadIdx.x] = b;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        acc[0] += s0[0];
        acc[1] += s1[0];
    }
}

}  // namespace rtps
