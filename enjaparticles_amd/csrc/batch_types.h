// Device-side description of one scene of a batch and of one work item of
// the persistent neighbour kernels.
#pragma once

#include <stdint.h>

#include "sph_device.h"

namespace rtps {

constexpr uint32_t kSceneShift = 24;   // segmented key: (scene << 24) | local key
constexpr int kMaxScenes = 256;
// The persistent pair kernels take their work items from the per-XCD tile
// queues of sph_device.h (kTileQueueWords ints, zero between launches; each
// launch's last block clears them), one set shared by the kernels of a step.

struct SceneDev {
    rtps_sph_params sp;          // SPHParams of the scene
    Grid g;                      // unscaled grid (hash)
    Grid gs;                     // scaled grid (neighbour search)
    rtps_float4 bmin, bmax;      // scaled domain bounds (walls)
    float res_x, res_y, res_z;   // grid resolution
    float dt;
    int off;                     // first slot of the scene in the particle arrays
    int cap;                     // slots of the scene (its max_num_particles)
    int cell_base;               // first entry of the scene's cell table
    int collision;               // wall collision on
    GravitySources gsrc;         // point sources of gravity.cl (n = 0: none)
};

struct BatchItem {
    int scene, first, count, pad_;
};

hipError_t launch_hash_b(int T, const uint16_t* sid, const SceneDev* scenes, const int* d_num, const float4* pos_u,
                         uint32_t* keys, uint32_t* idx, uint32_t* start, int n_cells, hipStream_t st);
// sub (optional): sub-cell start table, 16 entries per cell of every scene
hipError_t launch_cellindices_b(int T, const uint32_t* keys, const SceneDev* scenes, uint32_t* start, uint32_t* end,
                                hipStream_t st, uint32_t* sub = nullptr);
constexpr int kPairsDensity = 0, kPairsForceStep = 1, kPairsForceStore = 2;
// Launch configuration of a persistent pair kernel: `block` threads per block
// (= slots per work item: 64, 128 or 256) and the register budget of `waves`
// waves per SIMD (3, 4, 5 or 6: at most 168, 128, 96 or 80 VGPRs); the grid
// is CUs x waves x 256 / block blocks.  The density kernel is built for 4, 5
// and 6 waves; the force kernels for 3, 4 and 5 (5 fits their registers only
// on recorded rows: looking the rows up, their live values exceed the 96
// registers of 5 waves and spill to scratch memory).  The batched system's
// autotuner picks one per kernel; the defaults are 256 threads at 5 waves
// (density) and at 4 waves (force).
struct PairsCfg {
    int block = 256;
    int waves = 5;
};
int pairs_blocks_per_cu(const PairsCfg& c);
bool pairs_cfg_valid(int mode, const PairsCfg& c);
PairsCfg pairs_default_cfg(int mode);
hipError_t launch_pairs_b(int mode, const PairsCfg& cfg, int grid, const BatchItem* items, int n_items, int* queue,
                          const SceneDev* scenes, const int* d_num, const float4* pos_s, const uint32_t* start,
                          const uint32_t* end, const uint32_t* keys, float4* ve, const float4* vel_s,
                          const uint32_t* idx, float4* pos_u, float4* vel_u, float4* veleval_u, hipStream_t st,
                          float4* force_out = nullptr, float4* xsph_out = nullptr, const uint32_t* sub = nullptr,
                          uint32_t* rows = nullptr, uint32_t rows_stride = 0);
hipError_t launch_commit_b(int S, const SceneDev* scenes, int* d_num, const uint32_t* start, hipStream_t st);

}  // namespace rtps
