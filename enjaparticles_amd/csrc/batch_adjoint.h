// Launchers of the batched adjoint kernels (batch_adjoint.hip).  Every
// launch runs one block per work item of the batch layout.
#pragma once

#include <hip/hip_runtime.h>

#include "batch_types.h"

namespace rtps {

hipError_t launch_badj_integrate(int n_items, const BatchItem* items, const SceneDev* scenes, const float4* pos_s,
                                 const float4* vel_s, const float4* force_s, const float4* xsph_s,
                                 const uint32_t* idx, const uint32_t* start, float4* pos_u, float4* vel_u,
                                 float4* veleval_u, int* left, hipStream_t st);
hipError_t launch_badj_b1(int n_items, const BatchItem* items, const SceneDev* scenes, const uint32_t* idx,
                          const float4* pos_s, const float4* vel_s, const float4* force_s, const float4* lam_x,
                          const float4* lam_v, const float4* lam_e, float4* ga_s, float4* gX_s, float4* lam_v_out,
                          float4* lxp_s, hipStream_t st);
hipError_t launch_badj_b2(int n_items, const BatchItem* items, const SceneDev* scenes, const float4* pos_s,
                          const float4* ve_s, const float4* ga_s, const float4* gX_s, const uint32_t* start,
                          const uint32_t* end, const uint32_t* keys, float4* gx2_s, float4* ge_s, float* grho_s,
                          float2* partial, hipStream_t st);
hipError_t launch_badj_b3(int n_items, const BatchItem* items, const SceneDev* scenes, const uint32_t* idx,
                          const float4* pos_s, const float* grho_s, const float4* lxp_s, const float4* gx2_s,
                          const float4* ge_s, const uint32_t* start, const uint32_t* end, const uint32_t* keys,
                          float4* lam_x_out, float4* lam_e_out, hipStream_t st);
hipError_t launch_badj_reduce(int num_scenes, const int* item_first, const float2* partial, double* acc,
                              hipStream_t st);

}  // namespace rtps
