/*
 * rtps_c.h -- C ABI of the MI355X-native SPH step (EnjaParticles / rtpslib drop-in).
 *
 * Plain C types only (no HIP, no torch): every entry point takes plain
 * pointers, sizes and an opaque stream (a hipStream_t passed as void*, NULL =
 * the default stream).  Device pointers are marked "dev", host pointers "host".
 * Every function returns an int status (RTPS_OK = 0 or a negative RTPS_ERR_*);
 * rtps_last_error() gives the message of the last failure on the calling
 * thread.  No C++ exception crosses this boundary.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference root).
 */
#ifndef RTPS_C_H
#define RTPS_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTPS_OK 0
#define RTPS_ERR_ARG (-1)           /* bad argument (null, negative size, size limit) */
#define RTPS_ERR_HIP (-2)           /* a HIP runtime call failed */
#define RTPS_ERR_CAPACITY (-3)      /* more particles than max_num_particles */
#define RTPS_ERR_UNKNOWN_SYSTEM (-4) /* settings "system" is not "sph" */
#define RTPS_ERR_STATE (-5)         /* call not valid in the handle's state */

/* float4 of rtpslib/structs.h:403-581 (16-byte aligned). */
typedef struct rtps_float4 {
    float x, y, z, w;
} __attribute__((aligned(16))) rtps_float4;

/* SPHParams, rtpslib/system/SPHSettings.h:38-136 (device twin
 * sph/cl_src/cl_sph_structs.h:30-75): 144 bytes, gravity at offset 128. */
typedef struct rtps_sph_params {
    float rest_density, mass, rest_distance, smoothing_distance, simulation_scale;
    float boundary_stiffness, boundary_dampening, boundary_distance, K;
    float viscosity, velocity_limit, xsph_factor;
    float friction_coef, restitution_coef, shear, attraction, spring;
    float EPSILON;
    float wpoly6_coef, wpoly6_d_coef, wpoly6_dd_coef, wspiky_coef, wspiky_d_coef;
    float wspiky_dd_coef, wvisc_coef, wvisc_d_coef, wvisc_dd_coef;
    int num, max_num;
    int pad_[3];
    rtps_float4 gravity;
} __attribute__((aligned(16))) rtps_sph_params;

/* GridParams, rtpslib/domain/Domain.h:144-172 (twin common/cl_src/cl_structs.h:29-43): 128 bytes. */
typedef struct rtps_grid_params {
    rtps_float4 grid_size, grid_min, grid_max, bnd_min, bnd_max, grid_res, grid_delta;
    int nb_cells;
    int pad_[3];
} __attribute__((aligned(16))) rtps_grid_params;

/* ------------------------------------------------------------------------ */
/* Errors and version                                                        */
const char* rtps_last_error(void);
int rtps_version(void);

/* ------------------------------------------------------------------------ */
/* a1. Parameter derivation: SPH::calculate + SPH::updateParams
 * (rtpslib/system/SPHSettings.cpp:30-201) with the 6-digit settings round
 * trip (rtpslib/RTPSSettings.h:65-85), and System::setupDomain /
 * Domain::calculateCells (rtpslib/system/System.cpp:294-329,
 * rtpslib/domain/Domain.cpp:38-81).  keys/vals: n host C strings. */
int rtps_derive_params(const char* const* keys, const char* const* vals, int n, rtps_sph_params* sphp,
                       rtps_grid_params* gp, rtps_grid_params* gp_scaled);

/* ------------------------------------------------------------------------ */
/* Stage operators (device pointers).  One per reference stage.             */

/* a2. Hash::execute (rtpslib/system/common/Hash.h:49-61, hash.cl:50-157):
 * keys[i] = linear cell of pos_u[i] (clamped to nb_cells), idx[i] = i for
 * i < num; entries num <= i < n_pad get the pad key (all ones in
 * rtps_sort_key_bits(nb_cells) bits) so they sort last. */
int rtps_hash(int num, int n_pad, const rtps_float4* pos_u /*dev*/, uint32_t* keys /*dev*/, uint32_t* idx /*dev*/,
              const rtps_grid_params* gp /*host*/, void* stream);

/* Number of low key bits the sort must order for a grid of nb_cells cells
 * (the sentinel nb_cells and the pad key included). */
int rtps_sort_key_bits(int nb_cells);

/* a3. Bitonic<T>::Sort (rtpslib/system/common/BitonicSort.h:26-28) as used by
 * System::bitonic_sort (System.cpp:407-480): ascending key/value sort of n
 * pairs, in place.  Implemented as a stable LSD radix sort over the low
 * key_bits bits.  temp: dev scratch of rtps_sort_temp_bytes(n, key_bits). */
size_t rtps_sort_temp_bytes(int n, int key_bits);
int rtps_sort_pairs(int n, int key_bits, uint32_t* keys /*dev*/, uint32_t* vals /*dev*/, void* temp /*dev*/,
                    size_t temp_bytes, void* stream);

/* rtps_sort_pairs without the temp argument and without any allocation: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc).  The sort's scratch is a fixed 11 MiB device array inside the
 * library (one per GPU, held until the process ends), so n is limited to
 * 2^20 keys; larger n returns RTPS_ERR_ARG.  The array keeps the last call's
 * contents and the sort clears what must start at zero; calls that overlap on
 * one GPU (two streams or threads at once) share the array and race. */
int rtps_sort_pairs_noalloc(int n, int key_bits, uint32_t* keys /*dev*/, uint32_t* vals /*dev*/, void* stream);

/* a3. Bitonic<T>::Sort(batch, arrayLength, dir, dstkey, dstval, srckey,
 * srcval) (rtpslib/system/common/BitonicSort.h:26-28, BitonicSort.cpp:62-183)
 * with its argument meaning: `batch` independent arrays of array_length
 * key/value pairs, read from src_*, written sorted to dst_* (dst may equal
 * src), ascending for dir != 0, descending for dir == 0, on all 32 key bits;
 * array_length < 2 does nothing, as in the reference.  Any array_length (the
 * reference needs powers of two); equal keys keep their source order (the
 * bitonic network leaves them in a network-defined order).  temp: dev scratch
 * of rtps_sort_temp_bytes(array_length, 32) bytes. */
int rtps_sort_segments(int batch, int array_length, int dir, uint32_t* dst_keys /*dev*/, uint32_t* dst_vals /*dev*/,
                       const uint32_t* src_keys /*dev*/, const uint32_t* src_vals /*dev*/, void* temp /*dev*/,
                       size_t temp_bytes, void* stream);

/* a4. CellIndices::execute (rtpslib/system/common/CellIndices.h:47-58,
 * cellindices.cl:12-107): start[0..nb_cells] := 0xFFFFFFFF, then start/end of
 * every occupied cell from the sorted keys.  end[] of empty cells is left
 * untouched (as in the reference).  nc: optional dev int receiving
 * (int)start[nb_cells] (the reference reads it back to the host). */
int rtps_cellindices(int num, const uint32_t* keys /*dev*/, int nb_cells, uint32_t* start /*dev*/,
                     uint32_t* end /*dev*/, int* nc /*dev, may be NULL*/, void* stream);

/* a5. Permute::execute (rtpslib/system/common/Permute.h:46-79, permute.cl:10-46):
 * X_s[i] = X_u[idx[i]] for position, velocity and veleval. */
int rtps_permute(int num, const uint32_t* idx, const rtps_float4* pos_u, const rtps_float4* vel_u,
                 const rtps_float4* veleval_u, rtps_float4* pos_s, rtps_float4* vel_s, rtps_float4* veleval_s,
                 void* stream);

/* a7. Density::execute (rtpslib/system/sph/Density.h:40-53, density.cl:86-126):
 * density[i] = wpoly6 * m * sum_{|r|<=h} (h^2-r^2)^3 over the 27 cells, self
 * included.  gps = the scaled grid (SPH.cpp:255-267). */
int rtps_density(int num, const rtps_float4* pos_s, const uint32_t* start, const uint32_t* end,
                 const rtps_sph_params* sphp /*host*/, const rtps_grid_params* gps /*host*/, float* density,
                 void* stream);

/* a8. Force::execute (rtpslib/system/sph/Force.h:34, force.cl:148-186):
 * pressure (spiky), viscosity (Laplacian) and XSPH sums on veleval;
 * force[i] is an acceleration, xsph[i] the XSPH velocity correction. */
int rtps_force(int num, const rtps_float4* pos_s, const float* density, const rtps_float4* veleval_s,
               const uint32_t* start, const uint32_t* end, const rtps_sph_params* sphp, const rtps_grid_params* gps,
               rtps_float4* force, rtps_float4* xsph, void* stream);

/* a9. Gravity::execute (rtpslib/system/common/Gravity.cpp:42-79, gravity.cl:9-36):
 * force[i] += sum_s alpha_s m_s dir/dist^2 for nsrc point sources (host arrays,
 * at most RTPS_MAX_GRAVITY_SOURCES). */
#define RTPS_MAX_GRAVITY_SOURCES 16
int rtps_gravity(int num, int nsrc, const rtps_float4* sources /*host*/, const float* masses /*host*/,
                 const float* alphas /*host*/, const rtps_float4* pos_s, rtps_float4* force, float scale,
                 void* stream);

/* a10. CollisionWall::execute (rtpslib/system/sph/Collision_wall.h:35,
 * collision_wall.cl:31-119): six-wall penalty + friction added to force. */
int rtps_collision_wall(int num, const rtps_float4* pos_s, const rtps_float4* vel_s, rtps_float4* force,
                        const rtps_sph_params* sphp, const rtps_grid_params* gps, void* stream);

/* Triangle record of CollisionTriangle (collision_tri.cl:63-67): three
 * vertices (w unused) and a normal slot (recomputed from the vertices). */
typedef struct {
    rtps_float4 verts[3];
    rtps_float4 normal;
} rtps_triangle;

/* a10'. CollisionTriangle::execute (Collision_triangle.cpp,
 * collision_tri.cl:120-231, 418-470): for every particle and triangle, the
 * segment pos +- boundary_distance * v/|v| against the triangle (Ericson);
 * a hit closer than boundary_distance adds the penalty and friction of
 * cl_sph_collision.h along the triangle's unit normal to force.  tris is a
 * host array of n triangles in world units. */
int rtps_collision_triangles(int num, const rtps_float4* pos_s, const rtps_float4* vel_s, rtps_float4* force,
                             const rtps_triangle* tris /*host*/, int n, const rtps_sph_params* sphp, void* stream);

/* a11. LeapFrog::execute (rtpslib/system/sph/LeapFrog.h:44-63, leapfrog.cl:30-121):
 * integrates sorted particle i and scatters it to pos_u/vel_u/veleval_u[idx[i]]. */
int rtps_leapfrog(int num, const rtps_float4* pos_s, const rtps_float4* vel_s, const rtps_float4* force,
                  const rtps_float4* xsph, const uint32_t* idx, const rtps_sph_params* sphp, float dt,
                  rtps_float4* pos_u, rtps_float4* vel_u, rtps_float4* veleval_u, void* stream);

/* ------------------------------------------------------------------------ */
/* a12/a13. The SPH system: RTPS::generateSystemInstance (rtpslib/RTPS.h:45,
 * RTPS.cpp:37-67) + SPH (rtpslib/system/SPH.h:56-155).  A handle owns its
 * device buffers and one HIP stream; calls on one handle are not thread-safe,
 * calls on different handles are. */
typedef struct rtps_sph rtps_sph;

/* Settings as key/value strings (RTPSSettings).  Fails with
 * RTPS_ERR_UNKNOWN_SYSTEM unless "system" is "sph" (RTPS.cpp:62-66 returns NULL). */
int rtps_sph_create(const char* const* keys, const char* const* vals, int n, rtps_sph** out);
int rtps_sph_destroy(rtps_sph* s);

/* SPH::pushParticles (SPH.cpp:560-602): append n particles (host arrays of
 * float4; vel may be NULL for zero).  RTPS_ERR_CAPACITY if num + n > max_num
 * (the reference prints and returns). */
int rtps_sph_push_particles(rtps_sph* s, const rtps_float4* pos, const rtps_float4* vel, int n);
/* The same with the colour argument of SPH::pushParticles (every pushed
 * particle gets `color`; rtps_sph_push_particles uses the default (1,0,0,1)).
 * From the first push of another colour on, the system keeps colours sorted
 * and renumbered with their particles every step, as permute.cl / leapfrog.cl
 * do (RTPS_OPT_COLOR). */
int rtps_sph_push_particles_color(rtps_sph* s, const rtps_float4* pos, const rtps_float4* vel, int n,
                                  rtps_float4 color);

/* SPH::update (SPH.cpp:117-125,149-300) and SPH::integrate (SPH.cpp:389-441). */
int rtps_sph_update(rtps_sph* s);
int rtps_sph_integrate(rtps_sph* s);
/* nsteps x (update + integrate), captured in a hipGraph when enabled. */
int rtps_sph_step(rtps_sph* s, int nsteps);
/* Wait for all work on the handle's stream. */
int rtps_sph_synchronize(rtps_sph* s);

/* System::getNum (System.h:82-89); reads the device count (it can shrink
 * through the out-of-grid deletion path, SPH.cpp:220-250). */
int rtps_sph_get_num(rtps_sph* s, int* num);

/* Fields for rtps_sph_read / rtps_sph_device_ptr. */
enum rtps_field {
    RTPS_FIELD_POS = 0,        /* float4, unsorted (getPositionBufferUnsorted) */
    RTPS_FIELD_VEL = 1,        /* float4, unsorted */
    RTPS_FIELD_VELEVAL = 2,    /* float4, unsorted */
    RTPS_FIELD_POS_S = 3,      /* float4, sorted (getPositionBuffer) */
    RTPS_FIELD_VEL_S = 4,      /* float4, sorted */
    RTPS_FIELD_VELEVAL_S = 5,  /* float4, sorted; .w holds the density after a step */
    RTPS_FIELD_DENSITY = 6,    /* float, sorted (getDensityBuffer) */
    RTPS_FIELD_FORCE = 7,      /* float4, sorted (when force output is enabled) */
    RTPS_FIELD_XSPH = 8,       /* float4, sorted (when force output is enabled) */
    RTPS_FIELD_KEYS = 9,       /* uint32, sorted cell keys (device pointer: cell << 4 | in-cell x sixteenth) */
    RTPS_FIELD_IDX = 10,       /* uint32, sorted -> original index */
    RTPS_FIELD_CELL_START = 11, /* uint32[nb_cells+1] (getCellStartBuffer) */
    RTPS_FIELD_CELL_END = 12,  /* uint32[nb_cells+1] (getCellEndBuffer) */
    RTPS_FIELD_COLOR = 13,     /* float4, unsorted (getColorBufferUnsorted) */
    RTPS_FIELD_COLOR_S = 14,   /* float4, sorted (getColorBuffer); current while colours are kept */
    RTPS_FIELD_MASS = 15,      /* float, unsorted (getMassBufferUnsorted): the system mass; read only */
    RTPS_FIELD_MASS_S = 16     /* float, sorted (getMassBuffer); read only */
};
/* Blocking copy of the first `count` elements of a field into host memory. */
int rtps_sph_read(rtps_sph* s, int field, void* host, int count);
/* Device pointer of a field (zero-copy access on the handle's stream). */
int rtps_sph_device_ptr(rtps_sph* s, int field, void** dev);
/* Overwrite the first `count` elements of an unsorted float4 field (POS, VEL,
 * VELEVAL, COLOR; writing COLOR turns RTPS_OPT_COLOR on). */
int rtps_sph_write(rtps_sph* s, int field, const void* host, int count);

/* Host export for renderers (in place of the reference's GL buffer sharing,
 * System.cpp:214-238): rtps_sph_read_async enqueues the copy of the first
 * `count` elements of a float4 field, IDX or a cell array on the handle's
 * stream and returns; the data is in `host` after the next
 * rtps_sph_synchronize (or rtps_sph_read).  Steps enqueued after it do not
 * disturb it.  `host` should be page-locked memory from rtps_host_alloc, so
 * the copy overlaps host work. */
int rtps_host_alloc(size_t bytes, void** host);
int rtps_host_free(void* host);
int rtps_sph_read_async(rtps_sph* s, int field, void* host, int count);

/* RTPSSettings::SetSetting / GetSettingAs<string> on the handle's settings;
 * a change is applied at the next update (SPH.cpp:154). */
int rtps_sph_set_setting(rtps_sph* s, const char* key, const char* val);
int rtps_sph_get_setting(rtps_sph* s, const char* key, char* buf, int buflen);
int rtps_sph_get_params(rtps_sph* s, rtps_sph_params* sphp, rtps_grid_params* gp, rtps_grid_params* gps);

/* Options (not in the reference):
 *   RTPS_OPT_COLLISION   1 = wall collision on (default; the shipped SPH.cpp:297 has it off)
 *   RTPS_OPT_GRAPH       1 = capture rtps_sph_step's inner loop in a hipGraph (default 1)
 *   RTPS_OPT_FORCE_OUT   1 = also store force/xsph per step (default 0)
 *   RTPS_OPT_TIMERS      1 = per-stage hipEvent timers, recorded inside the step
 *                        graph (or around eager launches); two event sets
 *                        alternate, so the host reads step k-2's times while
 *                        step k-1 runs and rtps_sph_step returns after the
 *                        last step finished; 2 = the same for the force
 *                        stage only (two events per step); 3 = the force
 *                        stage only on the first of every four steps of an
 *                        rtps_sph_step call (the others run without events:
 *                        each event record costs a few microseconds between
 *                        kernels)
 *   RTPS_OPT_FORCE_VARIANT  force-loop kernel, process-wide default for handles
 *                        that are not autotuned: 0 = each lane walks its
 *                        windows in global memory, one branch per candidate;
 *                        4 = 0 with each row's candidates copied once per wave
 *                        into LDS; 5 = 4 with each chunk of 32 candidates
 *                        tested first into a bit mask and the pair terms run
 *                        once per accepted candidate (default).  The three give
 *                        the same bits.
 *   RTPS_OPT_WINDOW      1 = scan only the x window of each neighbour row, its
 *                        bounds looked up in a table of sub-cell starts that
 *                        cellindices writes (default; 64 bytes per grid cell),
 *                        2 = the same window found by binary search of the
 *                        sorted keys, 0 = whole rows
 *   RTPS_OPT_AUTOTUNE    1 = the first rtps_sph_step tunes the neighbour launches
 *                        over the force loops 5, 4 and 0 (default; see
 *                        rtps_sph_autotune), 0 = fixed defaults
 *   RTPS_OPT_SORT_TILE   radix-sort tile, process-wide: threads * 100 + keys
 *                        per thread (1024 x 1, 2, 4, 8; 512 x 4, 8; 256 x 4,
 *                        8, 16, e.g. 102404), or 0 (default) = by size
 *   RTPS_OPT_DENSITY_STAGE  1 = the density kernel copies each row's candidates
 *                        once per wave into LDS (default), 0 = every lane reads
 *                        its candidates from global memory; process-wide, the
 *                        same densities either way
 *   RTPS_OPT_BLOCK_DENSITY, RTPS_OPT_BLOCK_FORCE  threads per block (64, 128 or
 *                        256) of the density / force launches of this handle;
 *                        setting either fixes the launch configuration (the
 *                        handle is then not autotuned; the force loop is the
 *                        process default of RTPS_OPT_FORCE_VARIANT)
 *   RTPS_OPT_COLOR       1 = permute colours every step and renumber them with
 *                        the survivors of a deletion (permute.cl:41,
 *                        leapfrog.cl:107); 0 = colours stay where pushed.
 *                        Default 0; the first push of a non-default colour
 *                        sets 1.
 *   RTPS_OPT_SURFACE_FRAME  vertices of rtps_sph_surface / the post-process mesh:
 *                        0 = world coordinates of the boundary box (default),
 *                        1 = the reference's frame, lattice coordinates /
 *                        (res - 1) * 10 (traversehp.cl:512)
 *   RTPS_OPT_TILE_REMAP  tile order of the neighbour kernels over the 8 XCDs:
 *                        0 = hardware round-robin (default), 1 = one
 *                        contiguous range per XCD, C in 2..4096 = runs of C
 *                        consecutive tiles per XCD, interleaved
 *   RTPS_OPT_TILE_QUEUES 1 = the density and force kernels run as resident
 *                        grids taking their tiles from eight queues, one per
 *                        XCD over a contiguous eighth of the sorted particles,
 *                        a drained queue's blocks moving on to the next (each
 *                        XCD's L2 then serves one stretch of the neighbour
 *                        rows); 0 = one tile per block in RTPS_OPT_TILE_REMAP
 *                        order.  The same results either way.
 *   RTPS_OPT_BUCKET_SORT 1 = the step sorts its cell keys by bucket (a
 *                        histogram of key buckets, the keys placed by bucket,
 *                        each bucket sorted by one wave) where the key space
 *                        has at most 2^16 buckets averaging at most 1024 keys
 *                        (default), 0 = the LSD radix sort of rtps_sort_pairs.
 *                        The same order either way (keys ascending, equal keys
 *                        in particle order).
 *   RTPS_OPT_SHARED_ROWS 1 = with force loop 5, the density kernel records each
 *                        particle's nine neighbour windows and the within-h
 *                        bits of their last 32 candidates, and the force
 *                        kernel reads them instead of looking the windows up
 *                        and testing those candidates again (default; 108
 *                        bytes of device memory per particle slot, allocated
 *                        at creation); 0 = the force kernel looks them up.
 *                        The same results either way. */
#define RTPS_OPT_COLLISION 0
#define RTPS_OPT_GRAPH 1
#define RTPS_OPT_FORCE_OUT 2
#define RTPS_OPT_TIMERS 3
#define RTPS_OPT_FORCE_VARIANT 4
#define RTPS_OPT_WINDOW 5
#define RTPS_OPT_TILE_REMAP 6
#define RTPS_OPT_AUTOTUNE 7
#define RTPS_OPT_SORT_TILE 8
#define RTPS_OPT_COLOR 9
#define RTPS_OPT_DENSITY_STAGE 10
#define RTPS_OPT_BLOCK_DENSITY 11
#define RTPS_OPT_BLOCK_FORCE 12
#define RTPS_OPT_SURFACE_FRAME 13
#define RTPS_OPT_TILE_QUEUES 14
#define RTPS_OPT_BUCKET_SORT 15
#define RTPS_OPT_SHARED_ROWS 16
int rtps_sph_set_option(rtps_sph* s, int option, int value);

/* Autotuner (no reference counterpart): runs one step (as rtps_sph_step(s, 1))
 * during which the threads per block of the density and force launches (64,
 * 128 or 256) and the force loop (5, 4 and 0) are timed on the live state
 * and the fastest density + force pair kept
 * (a candidate tried later must be 5% faster).  The two kernels are idempotent on one sorted state, so the
 * step's result does not depend on the choice.  Choices are cached per
 * process by device and size class (and in the file named by the
 * RTPS_TUNE_CACHE environment variable, when set); measure = 0 takes a cached
 * choice when there is one.  With RTPS_OPT_AUTOTUNE on (default) the first
 * rtps_sph_step does this by itself.  Reports the configuration in use. */
int rtps_sph_autotune(rtps_sph* s, int measure, int* block_density, int* block_force, int* force_variant);

/* Per-stage GPU times (SPH::printTimers, SPH.cpp:664-672): average ms per
 * call since the last reset, in the order hash, sort, cellindices, permute,
 * density, force (force + gravity + collision + leapfrog when fused).  n <= 8. */
#define RTPS_NUM_STAGES 6
int rtps_sph_stage_times(rtps_sph* s, float* avg_ms, int* counts, int n);
int rtps_sph_reset_timers(rtps_sph* s);

/* Pair totals on the last step's sorted state: K = candidates of the 27-cell
 * neighbour loop, M = accepted pairs (self included) -- the FLOP model's
 * inputs -- and K_tested (may be NULL) = candidates the windowed rows test. */
int rtps_sph_count_pairs(rtps_sph* s, long long* K, long long* M, long long* K_tested);

/* The handle's stream (hipStream_t as void*). */
int rtps_sph_stream(rtps_sph* s, void** stream);

/* ------------------------------------------------------------------------ */
/* Hoses: Hose (rtpslib/system/common/Hose.cpp:30-128) with the disc emitter
 * addDiscRandom (rtpslib/domain/IV.cpp:313-365), host code.  A hose sprays up
 * to total_n particles, a perturbed disc of the given radius perpendicular to
 * its velocity every em = (int)(1 + spacing / dt / |velocity| / 10) calls.
 * The reference draws from libc rand(); a hose here draws from its own
 * SplitMix64 stream started at `seed`.  rtps_hose_spray writes the sprayed
 * positions (w = 1) to out (capacity >= the particles left) and their number
 * to *n. */
typedef struct rtps_hose rtps_hose;
int rtps_hose_create(float dt, int total_n, rtps_float4 center, rtps_float4 velocity, float radius, float spacing,
                     unsigned long long seed, rtps_hose** out);
int rtps_hose_destroy(rtps_hose* h);
int rtps_hose_update(rtps_hose* h, rtps_float4 center, rtps_float4 velocity, float radius, float spacing);
int rtps_hose_refill(rtps_hose* h, int num);
int rtps_hose_spray(rtps_hose* h, rtps_float4* out, int capacity, int* n);
/* Emission state: particles left, calls since the last spray, the cadence em,
 * the disc axes u, w and the velocity (after the reference's zero guard). */
int rtps_hose_state(rtps_hose* h, int* n_left, int* em_count, int* em, rtps_float4* u, rtps_float4* w,
                    rtps_float4* velocity);

/* SPH::addHose / updateHose / refillHose (SPH.cpp:523-545): hoses of the
 * system, sprayed at the start of every update (SPH::sprayHoses, SPH.cpp:166,
 * 547-557) with the hose velocity for the new particles.  radius is in units
 * of the settings' "spacing" (SPH.cpp:526).  Sprays are pushed like
 * rtps_sph_push_particles_color with the hose colour, so rtps_sph_step takes
 * one host round trip per step while a hose exists. */
int rtps_sph_add_hose(rtps_sph* s, int total_n, rtps_float4 center, rtps_float4 velocity, float radius,
                      rtps_float4 color, int* index);
int rtps_sph_update_hose(rtps_sph* s, int index, rtps_float4 center, rtps_float4 velocity, float radius,
                         rtps_float4 color);
int rtps_sph_refill_hose(rtps_sph* s, int index, int refill);
/* SPH::loadTriangles (SPH.cpp, Collision_triangle.cpp:41-53): triangles
 * (world units) the particles collide with after the walls in every update
 * (SPH::collision, with RTPS_OPT_COLLISION on); n = 0 removes them.  While
 * triangles are loaded a step runs force, gravity and walls, triangles and
 * the integration as separate kernels. */
int rtps_sph_load_triangles(rtps_sph* s, const rtps_triangle* tris, int n);

/* Rigid-body coupling of SPH::interact (SPH.cpp:610-663): the particles of a
 * rigid-body system (System::addInteractionSystem, System.h:167) push the
 * fluid.  rtps_sph_set_rigid_particles gives a moving body's particles
 * (rigidbody_force.cl:118-193: positions, velocities and masses, host arrays
 * in world units, as ParticleRigidBody's getPositionBuffer /
 * getVelocityBuffer / getMassBuffer), rtps_sph_set_static_boundary the
 * static boundary particles (staticrigidbody_force.cl:45-146,
 * getStaticPositionBuffer); spring and dampening are the rigid system's
 * "spring" and "dampening" settings (its friction settings are not read by
 * either kernel).  Each set is hashed into the fluid's grid at the call and
 * used until replaced; n = 0 removes it.  rtps_sph_interact adds the moving
 * body's forces, then the static boundary's, to the forces of the preceding
 * rtps_sph_update (call it between update and integrate, as the application
 * loop does, TestApplication.cpp:815-835); rtps_sph_step does this by itself
 * while a set is loaded (force, gravity and walls, triangles, rigid bodies and
 * the integration then run as separate kernels). */
int rtps_sph_set_rigid_particles(rtps_sph* s, const rtps_float4* pos, const rtps_float4* vel, const float* mass,
                                 int n, float spring, float dampening);
int rtps_sph_set_static_boundary(rtps_sph* s, const rtps_float4* pos, int n, float spring, float dampening);
int rtps_sph_interact(rtps_sph* s);

/* System::testDelete (System.cpp:366-376): moves the first 40 particles to
 * (100, 100, 100), outside the grid, so the next update deletes them. */
int rtps_sph_test_delete(rtps_sph* s);

/* ------------------------------------------------------------------------ */
/* Colour field and iso-surface: SPH::postProcess (SPH.cpp:338-386) with
 * ColorField (colorfield.cl, ColorField.cpp) and MarchingCubes
 * (MarchingCubes.cpp, classifycubes.cl, constructhp.cl, traversehp.cl), on
 * the sorted state of the last step or update.
 *
 * rtps_sph_color_field: the res^3 lattice over the scaled boundary box, node
 * (s, t, r) = mass * sum_j Wspiky(max(|x - x_j|, eps), h) / rho_j over the
 * neighbours within h; copied to `field` (may be NULL) x fastest, then y, z.
 *
 * rtps_sph_surface: the iso-surface of that field at `iso`, *ntris
 * triangles, the first `capacity` of them copied as 9 floats each: vertices
 * in world coordinates of the boundary box (or, with RTPS_OPT_SURFACE_FRAME
 * 1, in the reference's [0, 10]^3 frame) and unit normals pointing away from
 * the fluid.  Cubes are triangulated with the reference's case table, corner
 * bits and edge ends (rtps_mc_case_table, rtps_mc_geometry); triangles come in
 * cube order (x fastest), each cube's in the table's order (the reference
 * emits them in histogram-pyramid order).
 *
 * rtps_sph_post_process: SPH::postProcess: with setting "use_color_field"
 * = 1, the surface for "color_field_res" (default 2) and "color_field_iso"
 * (default 0.0001, the reference's isolevel), kept in the handle;
 * rtps_sph_mesh copies it (SPH::getMCMesh). */
int rtps_sph_color_field(rtps_sph* s, int res, float* field);
int rtps_sph_surface(rtps_sph* s, int res, float iso, int capacity, float* vertices, float* normals, int* ntris);
int rtps_sph_post_process(rtps_sph* s);
int rtps_sph_mesh(rtps_sph* s, int capacity, float* vertices, float* normals, int* ntris);
/* The marching-cubes case table of traversehp.cl:74 (triTable) with
 * classifycubes.cl:37's counts: tri[256 * 16] edge ids (-1 padded, three per
 * triangle), ntri[256] triangles per case.  rtps_mc_geometry: corners[8 * 3]
 * = the lattice offset of case bit k (classifycubes.cl:44-55: bit k set when
 * that node exceeds the iso level) and edges[12 * 6] = the two ends of edge e
 * (traversehp.cl:35-72, offsets3; vertices are interpolated from the first
 * toward the second).  Host only; either pointer may be NULL. */
int rtps_mc_case_table(signed char* tri, unsigned char* ntri);
int rtps_mc_geometry(signed char* corners, signed char* edges);

/* ------------------------------------------------------------------------ */
/* a14. Batched scenes (no reference counterpart): num_scenes independent SPH
 * systems (<= 256), each with its own settings, stepped together with one
 * launch per stage (segmented keys (scene << 24) | cell, persistent
 * density/force kernels over a queue of (scene, tile) items) and one hipGraph
 * per step.  Settings of scene s are the n_settings[s] key/value pairs that
 * follow those of scene s-1 in keys/vals.  Every scene computes what an
 * rtps_sph system with the same settings and particles computes. */
typedef struct rtps_batch rtps_batch;
int rtps_batch_create(int num_scenes, const int* n_settings, const char* const* keys, const char* const* vals,
                      rtps_batch** out);
int rtps_batch_destroy(rtps_batch* b);
int rtps_batch_num_scenes(rtps_batch* b, int* num_scenes, int* total_slots);
int rtps_batch_push_particles(rtps_batch* b, int scene, const rtps_float4* pos, const rtps_float4* vel, int n);
int rtps_batch_step(rtps_batch* b, int nsteps);
int rtps_batch_synchronize(rtps_batch* b);
int rtps_batch_get_num(rtps_batch* b, int scene, int* num);
/* Fields as for rtps_sph_read, per scene: unsorted fields in the scene's own
 * index order; KEYS are the segmented keys (scene << 24) | cell; IDX are
 * scene-local original indices. */
int rtps_batch_read(rtps_batch* b, int scene, int field, void* host, int count);
/* RTPS_OPT_GRAPH, RTPS_OPT_TIMERS, RTPS_OPT_COLLISION (all scenes),
 * RTPS_OPT_AUTOTUNE (1 = the first rtps_batch_step tunes the pair kernels,
 * default; see rtps_batch_autotune) and RTPS_OPT_BLOCK_DENSITY /
 * RTPS_OPT_BLOCK_FORCE = block * 10 + waves (e.g. 2565: 256 threads per block
 * = slots per work item, compiled for 5 waves per SIMD; block 64, 128 or 256,
 * waves 4..6 for density and 3..5 for force, whose registers spill at 5 waves
 * unless the rows are recorded), which fix that kernel's launch;
 * RTPS_OPT_SHARED_ROWS (1 = the
 * density kernel records each particle's neighbour windows and within-h bits
 * for the force kernel, default; 0 = the force kernel looks them up; the same
 * results either way; 108 bytes of device memory per slot). */
int rtps_batch_set_option(rtps_batch* b, int option, int value);
/* Autotuner of the batched pair kernels (no reference counterpart): runs one
 * step (as rtps_batch_step(b, 1)) during which the density and the force
 * kernel are each timed on the live state over block sizes 64, 128 and 256
 * (slots per work item) and register budgets of 4, 5 and 6 waves per SIMD
 * (density) or 3 and 4 (force), and keep their fastest (a configuration
 * tried later must be 3% faster).
 * Cached per device and layout size class (and in RTPS_TUNE_CACHE when set);
 * measure = 0 takes a cached choice.  Every configuration computes the same
 * step.  Reports the configurations in use. */
int rtps_batch_autotune(rtps_batch* b, int measure, int* block_density, int* waves_density, int* block_force,
                        int* waves_force);
int rtps_batch_stage_times(rtps_batch* b, float* avg_ms, int* counts, int n);
int rtps_batch_reset_timers(rtps_batch* b);

/* ------------------------------------------------------------------------ */
/* a15. Adjoint (no reference counterpart): gradients of a loss on the final
 * state with respect to the initial velocities, the viscosity and the gas
 * constant, through `nsteps` steps of the SPH system (density, force/XSPH,
 * walls, leapfrog; neighbour pairs of each step held fixed; simulation_scale
 * 1, no point sources, no deletion).  The forward keeps the state every
 * checkpoint_every steps; the backward replays each segment from its
 * checkpoint and runs the hand-written adjoint kernels through it.
 * All arrays are device float4[num] in original particle order; the calls
 * run on `stream`. */
typedef struct rtps_adj rtps_adj;
int rtps_adj_create(const char* const* keys, const char* const* vals, int n, int num_particles,
                    int checkpoint_every, int max_steps, rtps_adj** out);
int rtps_adj_destroy(rtps_adj* a);
/* RTPS_ERR_STATE if particles left the grid (the adjoint has no deletion path). */
int rtps_adj_forward(rtps_adj* a, const rtps_float4* pos0, const rtps_float4* vel0, int nsteps, float viscosity,
                     float gas_constant, int collision, rtps_float4* pos_T, rtps_float4* vel_T /* may be NULL */,
                     void* stream);
int rtps_adj_backward(rtps_adj* a, const rtps_float4* dL_dpos_T, const rtps_float4* dL_dvel_T /* may be NULL */,
                      rtps_float4* dL_dvel0, rtps_float4* dL_dpos0 /* may be NULL */, double* dL_dviscosity,
                      double* dL_dgas_constant, void* stream);

/* a14 + a15. Batched adjoint (no reference counterpart): the adjoint of
 * rtps_adj_* for num_scenes independent scenes (<= 256) at once, one launch
 * per stage for the whole batch, with per-scene viscosity and gas constant
 * and their gradients.  Scene settings are passed as for rtps_batch_create;
 * scene s holds exactly its max_num_particles particles, in slots
 * [offsets[s], offsets[s + 1]) of the concatenated device float4 arrays.
 * viscosity, gas_constant, dL_dviscosity and dL_dgas_constant are host
 * arrays of num_scenes entries.  Each scene gets what rtps_adj_* gives for
 * that scene alone. */
typedef struct rtps_badj rtps_badj;
int rtps_badj_create(int num_scenes, const int* n_settings, const char* const* keys, const char* const* vals,
                     int checkpoint_every, int max_steps, rtps_badj** out);
int rtps_badj_destroy(rtps_badj* a);
/* offsets: num_scenes + 1 entries (may be NULL). */
int rtps_badj_layout(rtps_badj* a, int* num_scenes, int* total, int* offsets);
int rtps_badj_forward(rtps_badj* a, const rtps_float4* pos0, const rtps_float4* vel0, int nsteps,
                      const float* viscosity, const float* gas_constant, int collision, rtps_float4* pos_T,
                      rtps_float4* vel_T /* may be NULL */, void* stream);
int rtps_badj_backward(rtps_badj* a, const rtps_float4* dL_dpos_T, const rtps_float4* dL_dvel_T /* may be NULL */,
                       rtps_float4* dL_dvel0, rtps_float4* dL_dpos0 /* may be NULL */, double* dL_dviscosity,
                       double* dL_dgas_constant, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RTPS_C_H */
