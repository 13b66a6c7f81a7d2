// Batched adjoint launches (no reference counterpart): the adjoint bodies of
// adjoint_device.h run over every scene of a batch at once.  One block per
// work item (a 256-slot tile of one scene), so the scene's parameters are
// uniform per block, and the dL/dmu, dL/dK partial sum of block b belongs to
// the scene of item b; k_badj_reduce then adds each scene's items in a fixed
// order (FP64), one block per scene.  Items tile each scene from its first
// slot, exactly as the single-system launches tile a system from sorted
// index 0, so every scene gets the single system's block partition.

#include <hip/hip_runtime.h>

#include "adjoint_device.h"
#include "batch_adjoint.h"
#include "rtps_internal.h"
#include "sph_device.h"

namespace rtps {

constexpr int kBABlock = 256;

struct ItemCtx {
    int s, i;
    bool valid;
};

__device__ __forceinline__ ItemCtx item_ctx(const BatchItem* __restrict__ items) {
    const BatchItem it = items[blockIdx.x];
    ItemCtx c;
    c.s = __builtin_amdgcn_readfirstlane(it.scene);
    c.i = it.first + (int)threadIdx.x;
    c.valid = (int)threadIdx.x < it.count;
    return c;
}

__device__ __forceinline__ AdjRows scene_rows(const SceneDev& S, int s, const uint32_t* start, const uint32_t* end,
                                              const uint32_t* keys) {
    return AdjRows{start + S.cell_base, end + S.cell_base, keys, S.res_x, S.res_y, S.res_z,
                   (uint32_t)s << kSceneShift};
}

__global__ __launch_bounds__(kBABlock) void k_badj_integrate(const BatchItem* __restrict__ items,
                                                             const SceneDev* __restrict__ scenes,
                                                             const float4* __restrict__ pos_s,
                                                             const float4* __restrict__ vel_s,
                                                             const float4* __restrict__ force_s,
                                                             const float4* __restrict__ xsph_s,
                                                             const uint32_t* __restrict__ idx,
                                                             const uint32_t* __restrict__ start,
                                                             float4* __restrict__ pos_u, float4* __restrict__ vel_u,
                                                             float4* __restrict__ veleval_u, int* __restrict__ left) {
    const ItemCtx c = item_ctx(items);
    const SceneDev& S = scenes[c.s];
    if (threadIdx.x == 0 && c.i == S.off && start[S.cell_base + S.g.nb_cells] != kEmpty) *left = 1;
    if (!c.valid) return;
    adj_integrate_body(c.i, pos_s, vel_s, force_s, xsph_s, idx, S.sp, S.bmin, S.bmax, S.collision, S.dt, pos_u,
                       vel_u, veleval_u);
}

__global__ __launch_bounds__(kBABlock) void k_badj_b1(const BatchItem* __restrict__ items,
                                                      const SceneDev* __restrict__ scenes,
                                                      const uint32_t* __restrict__ idx,
                                                      const float4* __restrict__ pos_s,
                                                      const float4* __restrict__ vel_s,
                                                      const float4* __restrict__ force_s,
                                                      const float4* __restrict__ lam_x,
                                                      const float4* __restrict__ lam_v,
                                                      const float4* __restrict__ lam_e, float4* __restrict__ ga_s,
                                                      float4* __restrict__ gX_s, float4* __restrict__ lam_v_out,
                                                      float4* __restrict__ lxp_s) {
    const ItemCtx c = item_ctx(items);
    if (!c.valid) return;
    const SceneDev& S = scenes[c.s];
    adj_b1_body(c.i, idx, pos_s, vel_s, force_s, S.sp, S.bmin, S.bmax, S.collision, S.dt, lam_x, lam_v, lam_e, ga_s,
                gX_s, lam_v_out, lxp_s);
}

__global__ __launch_bounds__(kBABlock) void k_badj_b2(const BatchItem* __restrict__ items,
                                                      const SceneDev* __restrict__ scenes,
                                                      const float4* __restrict__ pos_s,
                                                      const float4* __restrict__ ve_s,
                                                      const float4* __restrict__ ga_s,
                                                      const float4* __restrict__ gX_s,
                                                      const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ end,
                                                      const uint32_t* __restrict__ keys, float4* __restrict__ gx2_s,
                                                      float4* __restrict__ ge_s, float* __restrict__ grho_s,
                                                      float2* __restrict__ partial) {
    __shared__ float4 stage[kBABlock / 64][kAdjStageSlice];   // one slice per wave
    const ItemCtx c = item_ctx(items);
    const SceneDev& S = scenes[c.s];
    float gmu = 0.f, gK = 0.f;
    if (c.valid)
        adj_b2_body(c.i, pos_s, ve_s, ga_s, gX_s, scene_rows(S, c.s, start, end, keys), S.sp, S.gs, gx2_s, ge_s,
                    grho_s, gmu, gK, stage[threadIdx.x >> 6]);
    adj_block_partial<kBABlock>(gmu, gK, partial, blockIdx.x);
}

__global__ __launch_bounds__(kBABlock) void k_badj_b3(const BatchItem* __restrict__ items,
                                                      const SceneDev* __restrict__ scenes,
                                                      const uint32_t* __restrict__ idx,
                                                      const float4* __restrict__ pos_s,
                                                      const float* __restrict__ grho_s,
                                                      const float4* __restrict__ lxp_s,
                                                      const float4* __restrict__ gx2_s,
                                                      const float4* __restrict__ ge_s,
                                                      const uint32_t* __restrict__ start,
                                                      const uint32_t* __restrict__ end,
                                                      const uint32_t* __restrict__ keys,
                                                      float4* __restrict__ lam_x_out,
                                                      float4* __restrict__ lam_e_out) {
    __shared__ float4 stage[kBABlock / 64][kAdjStageSlice];   // one slice per wave
    const ItemCtx c = item_ctx(items);
    if (!c.valid) return;
    const SceneDev& S = scenes[c.s];
    adj_b3_body(c.i, idx, pos_s, grho_s, lxp_s, gx2_s, ge_s, scene_rows(S, c.s, start, end, keys), S.sp, S.gs,
                lam_x_out, lam_e_out, stage[threadIdx.x >> 6]);
}

// Block s adds partial[item_first[s], item_first[s + 1]) to acc[2 s], acc[2 s + 1].
__global__ void k_badj_reduce(const int* __restrict__ item_first, const float2* __restrict__ partial,
                              double* __restrict__ acc) {
    const int s = blockIdx.x;
    const int f = item_first[s];
    adj_reduce_body(partial, f, item_first[s + 1] - f, acc + 2 * s);
}

hipError_t launch_badj_integrate(int n_items, const BatchItem* items, const SceneDev* scenes, const float4* pos_s,
                                 const float4* vel_s, const float4* force_s, const float4* xsph_s,
                                 const uint32_t* idx, const uint32_t* start, float4* pos_u, float4* vel_u,
                                 float4* veleval_u, int* left, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (n_items <= 0) return hipSuccess;
    k_badj_integrate<<<n_items, kBABlock, 0, st>>>(items, scenes, pos_s, vel_s, force_s, xsph_s, idx, start, pos_u,
                                                   vel_u, veleval_u, left);
    return hipGetLastError();
}

hipError_t launch_badj_b1(int n_items, const BatchItem* items, const SceneDev* scenes, const uint32_t* idx,
                          const float4* pos_s, const float4* vel_s, const float4* force_s, const float4* lam_x,
                          const float4* lam_v, const float4* lam_e, float4* ga_s, float4* gX_s, float4* lam_v_out,
                          float4* lxp_s, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (n_items <= 0) return hipSuccess;
    k_badj_b1<<<n_items, kBABlock, 0, st>>>(items, scenes, idx, pos_s, vel_s, force_s, lam_x, lam_v, lam_e, ga_s,
                                            gX_s, lam_v_out, lxp_s);
    return hipGetLastError();
}

hipError_t launch_badj_b2(int n_items, const BatchItem* items, const SceneDev* scenes, const float4* pos_s,
                          const float4* ve_s, const float4* ga_s, const float4* gX_s, const uint32_t* start,
                          const uint32_t* end, const uint32_t* keys, float4* gx2_s, float4* ge_s, float* grho_s,
                          float2* partial, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (n_items <= 0) return hipSuccess;
    k_badj_b2<<<n_items, kBABlock, 0, st>>>(items, scenes, pos_s, ve_s, ga_s, gX_s, start, end, keys, gx2_s, ge_s,
                                            grho_s, partial);
    return hipGetLastError();
}

hipError_t launch_badj_b3(int n_items, const BatchItem* items, const SceneDev* scenes, const uint32_t* idx,
                          const float4* pos_s, const float* grho_s, const float4* lxp_s, const float4* gx2_s,
                          const float4* ge_s, const uint32_t* start, const uint32_t* end, const uint32_t* keys,
                          float4* lam_x_out, float4* lam_e_out, hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (n_items <= 0) return hipSuccess;
    k_badj_b3<<<n_items, kBABlock, 0, st>>>(items, scenes, idx, pos_s, grho_s, lxp_s, gx2_s, ge_s, start, end, keys,
                                            lam_x_out, lam_e_out);
    return hipGetLastError();
}

hipError_t launch_badj_reduce(int num_scenes, const int* item_first, const float2* partial, double* acc,
                              hipStream_t st) {
    (void)hipGetLastError();   // drop a stale error left by unrelated earlier HIP calls
    if (num_scenes <= 0) return hipSuccess;
    k_badj_reduce<<<num_scenes, 256, 0, st>>>(item_first, partial, acc);
    return hipGetLastError();
}

}  // namespace rtps
