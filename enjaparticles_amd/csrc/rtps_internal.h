// Internal declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "rtps_c.h"
#include "rtps/RTPSSettings.h"

namespace rtps {

// Flags of the library's timing events (stage timers, autotuners): they only
// time work on their stream and are never used to make device memory visible
// to the host, so their records skip the system-scope fence (a cache
// write-back and invalidate per record, which also leaves the next kernel's
// caches cold).
constexpr unsigned kTimingEvent = hipEventDisableSystemFence;

struct Grid;
struct GravitySources;
struct Win;

// ---- error reporting (capi.cpp) ----
int set_error(int code, const std::string& msg);
int hip_error(hipError_t e, const char* where);

// ---- parameter derivation (params.cpp) ----
// SPH::calculate (SPHSettings.cpp:30-144) on `s` (writes the derived keys),
// then SPH::updateParams (:149-201) and System::setupDomain (System.cpp:294-329).
void derive_params(RTPSSettings& s, rtps_sph_params& sp, rtps_grid_params& gp, rtps_grid_params& gps);
// SPH::updateParams only (re-read of the settings after a change).
void read_sph_params(const RTPSSettings& s, rtps_sph_params& sp);
int sort_key_bits(int nb_cells);
void gravity_sources_of(const RTPSSettings& st, GravitySources& g);
int integrator_of(const RTPSSettings& st);

// ---- kernels (sph_kernels.hip) ----
extern int g_density_stage;   // 1 = density rows staged through LDS (default), 0 = read from global memory
extern int g_force_variant;   // force loop: 0 = direct, 4 = rows staged in LDS, 5 = staged with accepted bits
hipError_t launch_hash(int num, const int* d_num, int n_pad, uint32_t pad_key, const float4* pos_u, uint32_t* keys,
                       uint32_t* idx, const Grid& g, uint32_t* start_clear, int n_clear, int sub_bits,
                       hipStream_t st);
// sub (optional, sub_bits = 4 only): the sub-cell start table of the
// windowed neighbour rows, 16 entries per cell
hipError_t launch_cellindices(int num, const int* d_num, const uint32_t* keys, int nb_cells, uint32_t* start,
                              uint32_t* end, int sub_bits, hipStream_t st, uint32_t* sub = nullptr);
hipError_t launch_read_nc(const uint32_t* start, int nb_cells, int* nc, hipStream_t st);
// cu/cs (optional): particle colours, permuted like the positions
hipError_t launch_permute(int num, const int* d_num, const uint32_t* idx, const float4* pu, const float4* vu,
                          const float4* eu, float4* ps, float4* vs, float4* es, hipStream_t st,
                          const float4* cu = nullptr, float4* cs = nullptr);
// launch_cellindices and launch_permute in one kernel (the system's step)
hipError_t launch_cells_permute(int num, const int* d_num, const uint32_t* keys, int nb_cells, uint32_t* start,
                                uint32_t* end, int sub_bits, uint32_t* sub, const uint32_t* idx, const float4* pu,
                                const float4* vu, const float4* eu, float4* ps, float4* vs, float4* es,
                                hipStream_t st, const float4* cu = nullptr, float4* cs = nullptr);
hipError_t launch_density(int num, const int* d_num, const float4* pos_s, const uint32_t* start, const uint32_t* end,
                          const rtps_sph_params& sp, const Grid& g, const Win& win, float* dens, float4* packed,
                          hipStream_t st);
hipError_t launch_force(int num, int* d_num, const float4* pos_s, const float* dens, const float4* ve,
                        bool packed, const uint32_t* start, const uint32_t* end, const rtps_sph_params& sp,
                        const Grid& g, const Win& win, float4* force, float4* xsph, hipStream_t st);
hipError_t launch_force_fused(int num, int* d_num, const float4* pos_s, const float4* ve_packed,
                              const uint32_t* start, const uint32_t* end, const rtps_sph_params& sp, const Grid& g,
                              const Win& win, float4* force, float4* xsph, const float4* vel_s, const uint32_t* idx,
                              const GravitySources& gsrc, int collision, const rtps_grid_params& gps, float dt,
                              float4* pos_u, float4* vel_u, float4* veleval_u, hipStream_t st,
                              const float4* color_s = nullptr, float4* color_u = nullptr);
hipError_t launch_post_force(int num, const int* d_num, const uint32_t* start, int nb_cells, const float4* pos_s,
                             const float4* vel_s, float4* force, const rtps_sph_params& sp,
                             const GravitySources& gsrc, int collision, const rtps_grid_params& gps, hipStream_t st);
hipError_t launch_leapfrog(int num, const int* d_num, const uint32_t* start, int nb_cells, const float4* pos_s,
                           const float4* vel_s, const float4* force, const float4* xsph, const uint32_t* idx,
                           const rtps_sph_params& sp, float dt, float4* pos_u, float4* vel_u, float4* veleval_u,
                           hipStream_t st, const float4* color_s = nullptr, float4* color_u = nullptr);
// euler.cl (integrator "euler"); ve_s: the sorted veleval|density records,
// copied to veleval_u on a deletion step
hipError_t launch_euler(int num, const int* d_num, const uint32_t* start, int nb_cells, const float4* pos_s,
                        const float4* vel_s, const float4* force, const float4* ve_s, const uint32_t* idx,
                        const rtps_sph_params& sp, float dt, float4* pos_u, float4* vel_u, float4* veleval_u,
                        hipStream_t st, const float4* color_s = nullptr, float4* color_u = nullptr);
hipError_t launch_count_pairs(int num, const int* d_num, const float4* pos_s, const uint32_t* start,
                              const uint32_t* end, const rtps_sph_params& sp, const Grid& g, const Win& win,
                              unsigned long long* out, hipStream_t st);
hipError_t launch_commit_num(int* d_num, const uint32_t* start, int nb_cells, hipStream_t st);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, long n, hipStream_t st);

// Default particle colour of System::pushParticles (System.h:160).
constexpr rtps_float4 kDefaultColor{1.0f, 0.0f, 0.0f, 1.0f};

// ---- triangle collision (collision_tri.hip) ----
// One triangle in the simulation scale: vertices a, b, c (q* = c), the
// unnormalised normal n = (b - a) x (c - a), the unit normal u, and the
// bounding sphere (centre c*, radius) with reach2 = (radius + dist)^2 widened.
struct TriPrep {
    float ax, ay, az, bx, by, bz, qx, qy, qz;
    float nx, ny, nz, ux, uy, uz;
    float cx, cy, cz, radius, reach2;
};
std::vector<TriPrep> prepare_triangles(const rtps_triangle* tris, int n, float scale);
void set_triangle_reach(std::vector<TriPrep>& t, float dist);
// (centre, reach squared) per triangle, for the block cull
std::vector<float4> triangle_reach(const std::vector<TriPrep>& t);
hipError_t launch_collision_tri(int num, const int* d_num, const uint32_t* start, int nb_cells, const float4* pos_s,
                                const float4* vel_s, float4* force, const TriPrep* tris, const float4* reach, int ntri,
                                const rtps_sph_params& sp, hipStream_t st);

// ---- colour field and iso-surface (surface.hip) ----
size_t surface_scan_tiles(long ncubes);
hipError_t launch_color_field(int res, const float4* pos_s, const float4* ve_s, const uint32_t* start,
                              const uint32_t* end, const rtps_sph_params& sp, const Grid& g, const Win& win,
                              const rtps_float4& bmin_scaled, const rtps_float4& bmax_scaled, float* field,
                              hipStream_t st);
hipError_t launch_mc_count(int res, float iso, const float* field, const uint8_t* ntri, uint8_t* cases,
                           uint32_t* counts, uint32_t* offs, uint32_t* tile_sums, uint32_t* total, hipStream_t st);
// the corner and edge tables into the emit kernel's constant memory (once per process)
hipError_t mc_upload_geometry();
hipError_t launch_mc_emit(int res, float iso, const float* field, const uint8_t* cases, const uint32_t* counts,
                          const uint32_t* offs, const uint32_t* tile_offs, const uint64_t* words,
                          const rtps_float4& wmin, const rtps_float4& wmax, int frame, long capacity, float* verts,
                          float* normals, hipStream_t st);

// ---- rigid-body coupling (rigid.hip) ----
// One set of rigid-body particles in the fluid's grid: sorted positions (and,
// for a moving body, velocities and masses), its cell start/end tables, and
// the staging arrays of the host upload.  s0 / s1: the rigid system's
// "spring" and "dampening" settings.
struct RigidDev {
    int n = 0, cap = 0;
    bool moving = false;
    float s0 = 0.f, s1 = 0.f;
    float4 *pos = nullptr, *vel = nullptr, *in_pos = nullptr, *in_vel = nullptr;
    float *mass = nullptr, *in_mass = nullptr;
    uint32_t *keys = nullptr, *idx = nullptr, *start = nullptr, *end = nullptr;
    void* sort_tmp = nullptr;
};
// Hashes, sorts and tables n particles of pos_u (vel_u / mass_u may be null)
// into r on the unscaled grid g.
hipError_t rigid_prepare(int n, const float4* pos_u, const float4* vel_u, const float* mass_u, RigidDev& r,
                         const Grid& g, hipStream_t st);
// Adds the forces of set r to the sorted fluid forces (gs: the scaled grid).
hipError_t launch_rigid_force(int num, const int* d_num, const uint32_t* fstart, int nb_cells, const float4* pos_s,
                              const float4* ve_s, float4* force, const RigidDev& r, const rtps_sph_params& sp,
                              const Grid& gs, hipStream_t st);

// ---- hoses (hose.cpp) ----
std::vector<rtps_float4> hose_spray(rtps_hose* h);

// ---- radix sort (radix_sort.hip) ----
extern int g_sort_items;   // one-sweep tile: threads * 100 + keys per thread, 0 = by size (process-wide)
bool sort_config_valid(int block, int keys_per_thread);
int sort_passes(int key_bits);
size_t sort_temp_bytes(int n, int key_bits);
hipError_t sort_clear_temp(void* temp, int n, hipStream_t st);
hipError_t sort_pairs_pingpong(int n, int key_bits, uint32_t* keys, uint32_t* vals, void* temp, hipStream_t st,
                               uint32_t** out_keys, uint32_t** out_vals);
hipError_t sort_pairs_inplace(int n, int key_bits, uint32_t* keys, uint32_t* vals, void* temp, hipStream_t st);
hipError_t sort_pairs_passes(int n, const int* shifts, const int* widths, int npasses, uint32_t* keys,
                             uint32_t* vals, void* temp, hipStream_t st, uint32_t** out_keys, uint32_t** out_vals);
int sort_plan(int key_bits, int* shifts, int* widths);
// launch_hash (keys of pos_u, pad_key past num, start[0..n_clear) cleared)
// and sort_pairs_pingpong of its (keys, idx) with the hash fused into the
// sort's histogram kernel: one kernel fewer per step.
hipError_t hash_sort_pairs(int n, int key_bits, const float4* pos_u, int num_host, const int* d_num, uint32_t pad_key,
                           const Grid& g, int sub_bits, uint32_t* start_clear, int n_clear, uint32_t* keys,
                           uint32_t* vals, void* temp, hipStream_t st, uint32_t** out_keys, uint32_t** out_vals);
hipError_t sort_segments(int batch, int length, int dir, uint32_t* dkeys, uint32_t* dvals, const uint32_t* skeys,
                         const uint32_t* svals, void* temp, hipStream_t st);

// ---- bucket sort of the step's cell keys (bucket_sort.hip) ----
// True when hash_bucket_sort takes n keys of key_bits bits.
bool bucket_sort_fits(int n, int key_bits);
// Temp bytes of hash_bucket_sort for every count up to n (at least
// sort_temp_bytes(n, key_bits)).
size_t bucket_temp_bytes(int n, int key_bits);
// hash_sort_pairs's result (keys of pos_u, pad_key past num, start[0..n_clear)
// cleared, (keys, idx) sorted stably) always left in (keys, vals).
hipError_t hash_bucket_sort(int n, int key_bits, const float4* pos_u, int num_host, const int* d_num, uint32_t pad_key,
                            const Grid& g, int sub_bits, uint32_t* start_clear, int n_clear, uint32_t* keys,
                            uint32_t* vals, void* temp, hipStream_t st);

}  // namespace rtps
