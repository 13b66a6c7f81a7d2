// Launchers of the adjoint kernels (adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sph_device.h"

namespace rtps {

int adjoint_blocks(int num);
hipError_t launch_adj_b1(int num, const uint32_t* idx, const float4* pos_s, const float4* vel_s,
                         const float4* force_s, const rtps_sph_params& sp, const rtps_grid_params& gps, int collision,
                         float dt, const float4* lam_x, const float4* lam_v, const float4* lam_e, float4* ga_s,
                         float4* gX_s, float4* lam_v_out, float4* lxp_s, hipStream_t st);
hipError_t launch_adj_b2(int num, const float4* pos_s, const float4* ve_s, const float4* ga_s, const float4* gX_s,
                         const uint32_t* start, const uint32_t* end, const rtps_sph_params& sp, const Grid& g,
                         const Win& win, float4* gx2_s, float4* ge_s, float* grho_s, float2* partial,
                         hipStream_t st);
hipError_t launch_adj_b3(int num, const uint32_t* idx, const float4* pos_s, const float* grho_s, const float4* lxp_s,
                         const float4* gx2_s, const float4* ge_s, const uint32_t* start, const uint32_t* end,
                         const rtps_sph_params& sp, const Grid& g, const Win& win, float4* lam_x_out,
                         float4* lam_e_out, hipStream_t st);
hipError_t launch_adj_reduce(const float2* partial, int n, double* acc, hipStream_t st);
hipError_t launch_adj_integrate(int num, const float4* pos_s, const float4* vel_s, const float4* force_s,
                                const float4* xsph_s, const uint32_t* idx, const uint32_t* start, int nb_cells,
                                const rtps_sph_params& sp, const rtps_grid_params& gps, int collision, float dt,
                                float4* pos_u, float4* vel_u, float4* veleval_u, int* left_grid, hipStream_t st);

}  // namespace rtps
