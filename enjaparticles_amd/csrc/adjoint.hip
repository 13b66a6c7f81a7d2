// Single-system launches of the hand-written adjoint of the SPH step (no
// reference counterpart).  The per-particle bodies and the derivation live in
// adjoint_device.h; here each kernel maps thread i to sorted particle i.

#include <hip/hip_runtime.h>

#include "adjoint.h"
#include "adjoint_device.h"
#include "rtps_internal.h"
#include "sph_device.h"

namespace rtps {

constexpr int kABlock = 256;

static inline int ablocks(long n) { return (int)((n + kABlock - 1) / kABlock); }

static inline AdjRows rows_of(const uint32_t* start, const uint32_t* end, const Win& win) {
    return AdjRows{start, end, win.keys, win.rx, win.ry, win.rz, 0u};
}

__global__ __launch_bounds__(kABlock) void k_adj_b1(int num, const uint32_t* __restrict__ idx,
                                                    const float4* __restrict__ pos_s,
                                                    const float4* __restrict__ vel_s,
                                                    const float4* __restrict__ force_s, rtps_sph_params sp,
                                                    rtps_float4 bmin, rtps_float4 bmax, int collision, float dt,
                                                    const float4* __restrict__ lam_x, const float4* __restrict__ lam_v,
                                                    const float4* __restrict__ lam_e, float4* __restrict__ ga_s,
                                                    float4* __restrict__ gX_s, float4* __restrict__ lam_v_out,
                                                    float4* __restrict__ lxp_s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    adj_b1_body(i, idx, pos_s, vel_s, force_s, sp, bmin, bmax, collision, dt, lam_x, lam_v, lam_e, ga_s, gX_s,
                lam_v_out, lxp_s);
}

// B2 plus the per-block partial sums of dL/dmu, dL/dK (partial[blockIdx.x]).
__global__ __launch_bounds__(kABlock) void k_adj_b2(int num, const float4* __restrict__ pos_s,
                                                    const float4* __restrict__ ve_s,
                                                    const float4* __restrict__ ga_s,
                                                    const float4* __restrict__ gX_s, AdjRows rows,
                                                    rtps_sph_params sp, Grid g, float4* __restrict__ gx2_s,
                                                    float4* __restrict__ ge_s, float* __restrict__ grho_s,
                                                    float2* __restrict__ partial) {
    __shared__ float4 stage[kABlock / 64][kAdjStageSlice];   // one slice per wave
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float gmu = 0.f, gK = 0.f;
    if (i < num)
        adj_b2_body(i, pos_s, ve_s, ga_s, gX_s, rows, sp, g, gx2_s, ge_s, grho_s, gmu, gK, stage[threadIdx.x >> 6]);
    adj_block_partial<kABlock>(gmu, gK, partial, blockIdx.x);
}

__global__ __launch_bounds__(kABlock) void k_adj_b3(int num, const uint32_t* __restrict__ idx,
                                                    const float4* __restrict__ pos_s,
                                                    const float* __restrict__ grho_s,
                                                    const float4* __restrict__ lxp_s,
                                                    const float4* __restrict__ gx2_s,
                                                    const float4* __restrict__ ge_s, AdjRows rows,
                                                    rtps_sph_params sp, Grid g, float4* __restrict__ lam_x_out,
                                                    float4* __restrict__ lam_e_out) {
    __shared__ float4 stage[kABlock / 64][kAdjStageSlice];   // one slice per wave
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= num) return;
    adj_b3_body(i, idx, pos_s, grho_s, lxp_s, gx2_s, ge_s, rows, sp, g, lam_x_out, lam_e_out, stage[threadIdx.x >> 6]);
}

// Walls + leapfrog of a recorded step.  Flags particles that left the grid
// (the adjoint does not follow the deletion path).
__global__ __launch_bounds__(kABlock) void k_adj_integrate(int num, const float4* __restrict__ pos_s,
                                                           const float4* __restrict__ vel_s,
                                                           const float4* __restrict__ force_s,
                                                           const float4* __restrict__ xsph_s,
                                                           const uint32_t* __restrict__ idx,
                                                           const uint32_t* __restrict__ start, int nb_cells,
                                                           rtps_sph_params sp, rtps_float4 bmin, rtps_float4 bmax,
                                                           int collision, float dt, float4* __restrict__ pos_u,
                                                           float4* __restrict__ vel_u,
                                                           float4* __restrict__ veleval_u, int* __restrict__ left) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && start[nb_cells] != kEmpty) *left = 1;
    if (i >= num) return;
    adj_integrate_body(i, pos_s, vel_s, force_s, xsph_s, idx, sp, bmin, bmax, collision, dt, pos_u, vel_u,
                       veleval_u);
}

// Adds the per-block partial sums (fixed order) to the running dL/dmu, dL/dK.
__global__ void k_adj_reduce(const float2* __restrict__ partial, int n, double* __restrict__ acc) {
    adj_reduce_body(partial, 0, n, acc);
}

int adjoint_blocks(int num) { return ablocks(num); }

hipError_t launch_adj_integrate(int num, const float4* pos_s, const float4* vel_s, const float4* force_s,
                                const float4* xsph_s, const uint32_t* idx, const uint32_t* start, int nb_cells,
                                const rtps_sph_params& sp, const rtps_grid_params& gps, int collision, float dt,
                                float4* pos_u, float4* vel_u, float4* veleval_u, int* left_grid, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_adj_integrate<<<ablocks(num), kABlock, 0, st>>>(num, pos_s, vel_s, force_s, xsph_s, idx, start, nb_cells, sp,
                                                      gps.bnd_min, gps.bnd_max, collision, dt, pos_u, vel_u,
                                                      veleval_u, left_grid);
    return hipGetLastError();
}

hipError_t launch_adj_b1(int num, const uint32_t* idx, const float4* pos_s, const float4* vel_s,
                         const float4* force_s, const rtps_sph_params& sp, const rtps_grid_params& gps, int collision,
                         float dt, const float4* lam_x, const float4* lam_v, const float4* lam_e, float4* ga_s,
                         float4* gX_s, float4* lam_v_out, float4* lxp_s, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_adj_b1<<<ablocks(num), kABlock, 0, st>>>(num, idx, pos_s, vel_s, force_s, sp, gps.bnd_min, gps.bnd_max,
                                               collision, dt, lam_x, lam_v, lam_e, ga_s, gX_s, lam_v_out, lxp_s);
    return hipGetLastError();
}

hipError_t launch_adj_b2(int num, const float4* pos_s, const float4* ve_s, const float4* ga_s, const float4* gX_s,
                         const uint32_t* start, const uint32_t* end, const rtps_sph_params& sp, const Grid& g,
                         const Win& win, float4* gx2_s, float4* ge_s, float* grho_s, float2* partial,
                         hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_adj_b2<<<ablocks(num), kABlock, 0, st>>>(num, pos_s, ve_s, ga_s, gX_s, rows_of(start, end, win), sp, g, gx2_s,
                                               ge_s, grho_s, partial);
    return hipGetLastError();
}

hipError_t launch_adj_b3(int num, const uint32_t* idx, const float4* pos_s, const float* grho_s, const float4* lxp_s,
                         const float4* gx2_s, const float4* ge_s, const uint32_t* start, const uint32_t* end,
                         const rtps_sph_params& sp, const Grid& g, const Win& win, float4* lam_x_out,
                         float4* lam_e_out, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num <= 0) return hipSuccess;
    k_adj_b3<<<ablocks(num), kABlock, 0, st>>>(num, idx, pos_s, grho_s, lxp_s, gx2_s, ge_s, rows_of(start, end, win),
                                               sp, g, lam_x_out, lam_e_out);
    return hipGetLastError();
}

hipError_t launch_adj_reduce(const float2* partial, int n, double* acc, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    k_adj_reduce<<<1, 256, 0, st>>>(partial, n, acc);
    return hipGetLastError();
}

}  // namespace rtps
