// The SPH system handle: RTPS::generateSystemInstance + SPH (rtpslib/RTPS.cpp:37-67,
// rtpslib/system/SPH.cpp) on one HIP stream.
//
// Device state (all sized for max_num_particles, one HIP allocation each):
//   unsorted  pos_u, vel_u, veleval_u        float4   (original index order)
//   sorted    pos_s, vel_s, ve_s             float4   (ve_s = veleval.xyz | density)
//             force_s, xsph_s                float4   (unfused path / force output)
//   keys/idx (+ alternates in the sort temp) u32
//   start/end                                u32[nb_cells + 1]
//   d_num                                    the particle count, on the device
// The step never synchronises with the host: the per-step clear of start[]
// is fused into the hash kernel, nc stays on the device and the deletion
// path's new count is committed by a one-thread kernel, so rtps_sph_step can
// replay one captured hipGraph per step.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <vector>

#include "rtps_internal.h"
#include "mc_table.h"
#include "sph_device.h"

using namespace rtps;

enum Stage { ST_HASH = 0, ST_SORT, ST_CELL, ST_PERMUTE, ST_DENSITY, ST_FORCE, ST_N };

struct rtps_sph {
    RTPSSettings settings;
    rtps_sph_params sp{};
    rtps_grid_params gp{}, gps{};
    Grid g{}, gs{};
    float dt = 0.f;
    int integrator = 0;   // settings "integrator": 0 = leapfrog, 1 = euler (SPH.cpp:89-95, 392-409)
    int max_num = 0, num_host = 0, nb_cells = 0, key_bits = 0;
    uint32_t pad_key = 0;
    hipStream_t stream = nullptr;
    float4 *pos_u = nullptr, *vel_u = nullptr, *veleval_u = nullptr;
    float4 *pos_s = nullptr, *vel_s = nullptr, *ve_s = nullptr, *force_s = nullptr, *xsph_s = nullptr;
    float4 *color_u = nullptr, *color_s = nullptr;   // cl_color_u / cl_color_s (System.cpp:236-237)
    uint32_t *keys = nullptr, *idx = nullptr, *start = nullptr, *end = nullptr;
    uint32_t *keys_sorted = nullptr, *idx_sorted = nullptr;
    // sub-cell starts of the windowed rows (16 per cell; nullptr when the grid
    // is too large for it: the kernels then search the sorted keys)
    uint32_t* sub_start = nullptr;
    bool sub_valid = false;   // the last cellindices wrote the table for the current sorted state
    int* d_num = nullptr;
    int* tile_q = nullptr;   // per-XCD tile queue counters of the neighbour kernels (kTileQueueWords)
    void* sort_tmp = nullptr;
    // the neighbour rows the density kernel records for force loop 5 (Win::rows):
    // 9 rows x 3 words per particle slot, then 9 x 2 words per wave of 64 slots
    uint32_t* rows = nullptr;
    GravitySources gsrc{};
    int opt_collision = 1, opt_graph = 1, opt_force_out = 0, opt_timers = 0, opt_window = 1, opt_remap = 0,
        opt_queues = 0;
    int opt_bucket = 1;   // the step's keys: bucket sort where the key space allows it (1) or LSD radix sort (0)
    int opt_rows = 1;     // density records the neighbour rows for force loop 5 (1) or force looks them up (0)
    int opt_autotune = 1;
    int opt_color = 0;   // permute colours every step (set by the first non-default colour pushed)
    int opt_frame = 0;   // surface vertices: 0 = world coordinates, 1 = the reference's [0, 10]^3 frame
    // launch configuration of the neighbour kernels (0 / -1 = defaults), set by the autotuner
    int block_density = 0, block_force = 0, force_variant = -1;
    bool tuned = false;
    hipGraphExec_t gexec = nullptr;
    bool graph_valid = false;
    // Timers: two event sets alternate between steps (set k & 1 for step k),
    // each recorded by its own captured step graph when graphs are on.
    hipGraphExec_t gexec_t[2] = {};
    bool timed_graph_valid = false;
    int graph_events_ok = -1;   // captured external event records give elapsed times (-1 = not probed)
    hipEvent_t ev[2][ST_N + 1] = {};
    bool ev_pending[2] = {};
    double stage_ms[ST_N] = {};
    long stage_cnt[ST_N] = {};
    bool update_pending = false;
    bool sorted_ready = false;   // a step or update has filled the sorted arrays
    // colour field and iso-surface (SPH::postProcess)
    int cf_res = 0;
    float* cf_field = nullptr;
    uint8_t* mc_cases = nullptr;
    uint32_t *mc_counts = nullptr, *mc_offs = nullptr, *mc_tiles = nullptr, *mc_total = nullptr;
    uint64_t* mc_words = nullptr;   // the case table, one packed word per case
    uint8_t* mc_ntri = nullptr;
    float *mesh_v = nullptr, *mesh_n = nullptr;
    long mesh_cap = 0;
    int mesh_ntri = 0;
    std::vector<rtps_hose*> hoses;   // SPH::hoses (SPH.h), sprayed at the start of every update
    std::vector<rtps_triangle> triangles;   // SPH::loadTriangles, world units
    TriPrep* tri_dev = nullptr;   // tri_cap records, then tri_cap reach spheres (tri_reach)
    float4* tri_reach = nullptr;
    size_t tri_cap = 0;
    int ntri = 0;
    std::vector<rtps_float4> hose_color;
    // rigid-body particle sets of SPH::interact: [0] a moving body, [1] the static boundary
    RigidDev rigid[2];

    ~rtps_sph() {
        for (rtps_hose* h : hoses) rtps_hose_destroy(h);
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto& g : gexec_t)
            if (g) (void)hipGraphExecDestroy(g);
        void* bufs[] = {pos_u, vel_u, veleval_u, pos_s, vel_s, ve_s,  force_s,  xsph_s,   color_u,
                        color_s, keys, idx, start, end, d_num, sort_tmp, sub_start, tile_q, rows};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        void* sbufs[] = {cf_field, mc_cases, mc_counts, mc_offs, mc_tiles, mc_total, mc_words, mc_ntri, mesh_v, mesh_n,
                         tri_dev};
        for (void* b : sbufs)
            if (b) (void)hipFree(b);
        for (RigidDev& r : rigid) {
            void* rb[] = {r.pos, r.vel, r.in_pos, r.in_vel, r.mass, r.in_mass, r.keys, r.idx, r.start, r.end,
                          r.sort_tmp};
            for (void* b : rb)
                if (b) (void)hipFree(b);
        }
        for (auto& set : ev)
            for (auto& e : set)
                if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

// Sub-cell start table of the windowed rows, when it is in use.
// Written by cellindices while the window option asks for it; read by the
// neighbour kernels only when the current sorted state's table was written
// (entries of cells that are empty now may be stale or uninitialised).
uint32_t* sub_table_to_write(const rtps_sph* s) { return s->opt_window == 1 ? s->sub_start : nullptr; }
uint32_t* sub_table(const rtps_sph* s) { return s->sub_valid ? s->sub_start : nullptr; }

// Colour buffers handed to the kernels (none unless the system keeps colours).
float4* col_u(const rtps_sph* s) { return s->opt_color ? s->color_u : nullptr; }
float4* col_s(const rtps_sph* s) { return s->opt_color ? s->color_s : nullptr; }

// Launch inputs of the density (force = false) or force kernel.
Win win(const rtps_sph* s, bool force = false) {
    Win w{s->opt_window ? s->keys_sorted : nullptr, s->gps.grid_res.x, s->gps.grid_res.y, s->gps.grid_res.z,
          s->opt_remap, force ? s->block_force : s->block_density, force ? s->force_variant : -1};
    w.sub = sub_table(s);
    w.tq = s->opt_queues ? s->tile_q : nullptr;
    // both launches of a step take the same choice: density records the rows
    // exactly when the force launch will read them (loop 5)
    const int var = s->force_variant >= 0 ? s->force_variant : g_force_variant;
    if (s->opt_rows && s->rows && var == 5) {
        w.rows = s->rows;
        w.rows_stride = (uint32_t)s->max_num;
    }
    return w;
}

template <typename T>
hipError_t dalloc(T** p, size_t count) {
    return hipMalloc((void**)p, count * sizeof(T) + 16);
}

#define RTPS_TRY_HIP(expr)                                     \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return hip_error(e_, #expr);     \
    } while (0)

void load_gravity_sources(rtps_sph* s) { gravity_sources_of(s->settings, s->gsrc); }

// Triangles in the current simulation scale and boundary distance, on the device.
hipError_t upload_triangles(rtps_sph* s) {
    s->ntri = 0;
    if (s->triangles.empty()) return hipSuccess;
    std::vector<TriPrep> prep = prepare_triangles(s->triangles.data(), (int)s->triangles.size(),
                                                  s->sp.simulation_scale);
    set_triangle_reach(prep, s->sp.boundary_distance);
    const std::vector<float4> reach = triangle_reach(prep);
    if (prep.size() > s->tri_cap) {
        hipError_t e = hipStreamSynchronize(s->stream);
        if (e == hipSuccess && s->tri_dev) e = hipFree(s->tri_dev);
        s->tri_dev = nullptr;
        s->tri_reach = nullptr;
        s->tri_cap = 0;
        if (e == hipSuccess) e = hipMalloc(&s->tri_dev, prep.size() * (sizeof(TriPrep) + sizeof(float4)));
        if (e != hipSuccess) return e;
        s->tri_cap = prep.size();
        s->tri_reach = reinterpret_cast<float4*>(s->tri_dev + s->tri_cap);
    }
    hipError_t e = hipMemcpyAsync(s->tri_dev, prep.data(), prep.size() * sizeof(TriPrep), hipMemcpyHostToDevice,
                                  s->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->tri_reach, reach.data(), reach.size() * sizeof(float4), hipMemcpyHostToDevice,
                           s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) s->ntri = (int)prep.size();
    return e;
}

void apply_settings(rtps_sph* s) {
    read_sph_params(s->settings, s->sp);
    s->integrator = std::max(0, integrator_of(s->settings));
    s->sp.num = s->num_host;
    s->dt = s->settings.GetSettingAs<float>("time_step");
    load_gravity_sources(s);
    if (!s->triangles.empty()) (void)upload_triangles(s);   // a failure leaves ntri = 0: reported by the next load
    s->settings.updated();
    s->graph_valid = false;
    s->timed_graph_valid = false;
}

int maybe_update_params(rtps_sph* s) {
    if (s->settings.hasChanged()) apply_settings(s);   // SPH.cpp:154
    return RTPS_OK;
}

struct StageTimer {
    rtps_sph* s;
    bool on;
    int set = 0;
    bool external = false;   // inside a graph capture: record as external event nodes
    hipError_t err = hipSuccess;
    explicit StageTimer(rtps_sph* s_, int set_ = 0) : s(s_), on(s_->opt_timers != 0), set(set_) {}
    void mark(int i) {
        if (!on) return;
        if (s->opt_timers >= 2 && i != ST_FORCE && i != ST_N) return;   // force stage only
        hipError_t e = external ? hipEventRecordWithFlags(s->ev[set][i], s->stream, hipEventRecordExternal)
                                : hipEventRecord(s->ev[set][i], s->stream);
        if (e != hipSuccess && err == hipSuccess) err = e;
    }
};

// hash -> sort -> cellindices -> permute (System::hash_and_sort, SPH.cpp:170-213)
int launch_front(rtps_sph* s, StageTimer& t) {
    const int n = s->num_host;
    s->sorted_ready = true;
    // the hash runs inside the sort's first (histogram) kernel: the "sort"
    // stage holds both, the "hash" stage is empty
    t.mark(ST_HASH);
    t.mark(ST_SORT);
    if (s->opt_bucket && bucket_sort_fits(n, s->key_bits)) {
        RTPS_TRY_HIP(hash_bucket_sort(n, s->key_bits, s->pos_u, n, s->d_num, s->pad_key, s->g, kSubBits, s->start,
                                      s->nb_cells + 1, s->keys, s->idx, s->sort_tmp, s->stream));
        s->keys_sorted = s->keys;
        s->idx_sorted = s->idx;
    } else {
        RTPS_TRY_HIP(hash_sort_pairs(n, s->key_bits, s->pos_u, n, s->d_num, s->pad_key, s->g, kSubBits, s->start,
                                     s->nb_cells + 1, s->keys, s->idx, s->sort_tmp, s->stream, &s->keys_sorted,
                                     &s->idx_sorted));
    }
    // cellindices and permute in one kernel: the "cellindices" stage holds
    // both, the "permute" stage is empty
    t.mark(ST_CELL);
    RTPS_TRY_HIP(launch_cells_permute(n, s->d_num, s->keys_sorted, s->nb_cells, s->start, s->end, kSubBits,
                                      sub_table_to_write(s), s->idx_sorted, s->pos_u, s->vel_u, s->veleval_u,
                                      s->pos_s, s->vel_s, s->ve_s, s->stream, col_u(s), col_s(s)));
    s->sub_valid = sub_table_to_write(s) != nullptr;
    t.mark(ST_PERMUTE);
    t.mark(ST_DENSITY);
    RTPS_TRY_HIP(launch_density(n, s->d_num, s->pos_s, s->start, s->end, s->sp, s->gs, win(s), nullptr, s->ve_s,
                                s->stream));
    return RTPS_OK;
}

bool rigid_on(const rtps_sph* s) { return s->rigid[0].n > 0 || s->rigid[1].n > 0; }

// SPH::interact (SPH.cpp:610-663): the moving body's particles, then the
// static boundary's, add their forces to the sorted forces of update().
int launch_interact(rtps_sph* s) {
    for (const RigidDev& r : s->rigid)
        RTPS_TRY_HIP(launch_rigid_force(s->num_host, s->d_num, s->start, s->nb_cells, s->pos_s, s->ve_s, s->force_s, r,
                                        s->sp, s->gs, s->stream));
    return RTPS_OK;
}

// SPH::integrate (SPH.cpp:389-441) on the stored forces: euler.cl or leapfrog.cl.
int launch_integrate(rtps_sph* s) {
    if (s->integrator == 1)
        RTPS_TRY_HIP(launch_euler(s->num_host, s->d_num, s->start, s->nb_cells, s->pos_s, s->vel_s, s->force_s,
                                  s->ve_s, s->idx_sorted, s->sp, s->dt, s->pos_u, s->vel_u, s->veleval_u, s->stream,
                                  col_s(s), col_u(s)));
    else
        RTPS_TRY_HIP(launch_leapfrog(s->num_host, s->d_num, s->start, s->nb_cells, s->pos_s, s->vel_s, s->force_s,
                                     s->xsph_s, s->idx_sorted, s->sp, s->dt, s->pos_u, s->vel_u, s->veleval_u,
                                     s->stream, col_s(s), col_u(s)));
    return RTPS_OK;
}

// Force through integration + count commit, as the application loop runs a
// step (update, interact, integrate; TestApplication.cpp:815-835).  One
// fused kernel, or, with the euler integrator, triangles in the collisions or
// rigid-body sets loaded: force; gravity sources and walls; triangles
// (SPH::collision, SPH.cpp:303-336); rigid-body forces; the integration, as
// separate kernels.
int launch_back(rtps_sph* s) {
    const bool fo = s->opt_force_out != 0;
    const bool tri = s->ntri > 0 && s->opt_collision;
    if (s->integrator == 1 || tri || rigid_on(s)) {
        RTPS_TRY_HIP(launch_force(s->num_host, s->d_num, s->pos_s, nullptr, s->ve_s, true, s->start, s->end, s->sp,
                                  s->gs, win(s, true), s->force_s, s->xsph_s, s->stream));
        RTPS_TRY_HIP(launch_post_force(s->num_host, s->d_num, s->start, s->nb_cells, s->pos_s, s->vel_s, s->force_s,
                                       s->sp, s->gsrc, s->opt_collision, s->gps, s->stream));
        if (tri)
            RTPS_TRY_HIP(launch_collision_tri(s->num_host, s->d_num, s->start, s->nb_cells, s->pos_s, s->vel_s,
                                              s->force_s, s->tri_dev, s->tri_reach, s->ntri, s->sp, s->stream));
        int rc = launch_interact(s);
        if (rc) return rc;
        rc = launch_integrate(s);
        if (rc) return rc;
    } else {
        RTPS_TRY_HIP(launch_force_fused(s->num_host, s->d_num, s->pos_s, s->ve_s, s->start, s->end, s->sp, s->gs,
                                        win(s, true), fo ? s->force_s : nullptr, fo ? s->xsph_s : nullptr, s->vel_s,
                                        s->idx_sorted, s->gsrc, s->opt_collision, s->gps, s->dt, s->pos_u, s->vel_u,
                                        s->veleval_u, s->stream, col_s(s), col_u(s)));
        return RTPS_OK;   // the fused kernel commits the count of a deletion step itself
    }
    RTPS_TRY_HIP(launch_commit_num(s->d_num, s->start, s->nb_cells, s->stream));
    return RTPS_OK;
}

// One step: front + launch_back.
int launch_step(rtps_sph* s, StageTimer& t) {
    int rc = launch_front(s, t);
    if (rc) return rc;
    t.mark(ST_FORCE);
    rc = launch_back(s);
    if (rc) return rc;
    t.mark(ST_N);
    return RTPS_OK;
}

// Adds the elapsed times of event set `set` (a finished or finishing step).
int collect_timers(rtps_sph* s, int set) {
    if (!s->ev_pending[set]) return RTPS_OK;
    s->ev_pending[set] = false;
    RTPS_TRY_HIP(hipEventSynchronize(s->ev[set][ST_N]));
    for (int i = s->opt_timers >= 2 ? (int)ST_FORCE : 0; i < ST_N; ++i) {
        float ms = 0.f;
        RTPS_TRY_HIP(hipEventElapsedTime(&ms, s->ev[set][i], s->ev[set][i + 1]));
        s->stage_ms[i] += ms;
        s->stage_cnt[i] += 1;
    }
    return RTPS_OK;
}

// Captures one step into *out; with timers, event set `set` is recorded
// around the stages inside the graph.
int capture_step(rtps_sph* s, bool timed, int set, hipGraphExec_t* out) {
    if (*out) {
        (void)hipGraphExecDestroy(*out);
        *out = nullptr;
    }
    hipGraph_t graph = nullptr;
    RTPS_TRY_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    StageTimer t(s, set);
    t.on = timed;
    t.external = true;
    int rc = launch_step(s, t);
    hipError_t e = hipStreamEndCapture(s->stream, &graph);
    if (rc || t.err != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc ? rc : hip_error(t.err, "hipEventRecord (graph capture)");
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_error(e, "hipGraphInstantiate");
    return RTPS_OK;
}

int capture_graph(rtps_sph* s) {
    int rc = capture_step(s, false, 0, &s->gexec);
    if (rc) return rc;
    s->graph_valid = true;
    return RTPS_OK;
}

// Whether events recorded by a replayed graph (external event nodes) give
// elapsed times: a graph of two event records is replayed once.
bool probe_graph_events(rtps_sph* s) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t ex = nullptr;
    bool ok = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        bool rec = hipEventRecordWithFlags(s->ev[0][0], s->stream, hipEventRecordExternal) == hipSuccess &&
                   hipEventRecordWithFlags(s->ev[0][1], s->stream, hipEventRecordExternal) == hipSuccess;
        ok = hipStreamEndCapture(s->stream, &graph) == hipSuccess && rec;
    }
    ok = ok && hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0) == hipSuccess;
    ok = ok && hipGraphLaunch(ex, s->stream) == hipSuccess && hipStreamSynchronize(s->stream) == hipSuccess;
    float ms = -1.f;
    ok = ok && hipEventElapsedTime(&ms, s->ev[0][0], s->ev[0][1]) == hipSuccess && ms >= 0.f;
    if (ex) (void)hipGraphExecDestroy(ex);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return ok;
}

// Timed steps: step k records event set k & 1; the set is read back just
// before it is recorded again (two steps later), so the host never waits for
// the step the GPU is running and the timers add no gap between steps.
// RTPS_OPT_TIMERS 3 times the force stage on one step in this many.
constexpr int kTimerSample = 4;

int step_timed(rtps_sph* s, int nsteps) {
    if (s->opt_graph && s->graph_events_ok < 0) s->graph_events_ok = probe_graph_events(s) ? 1 : 0;
    const bool graph = s->opt_graph && s->graph_events_ok == 1;
    if (graph && !s->timed_graph_valid) {
        for (int k = 0; k < 2; ++k) {
            int rc = capture_step(s, true, k, &s->gexec_t[k]);
            if (rc) return rc;
        }
        s->timed_graph_valid = true;
    }
    // RTPS_OPT_TIMERS 3: only every kTimerSample-th step carries the events;
    // the steps between run the untimed graph (or launches)
    const int every = s->opt_timers == 3 ? kTimerSample : 1;
    if (graph && every > 1 && !s->graph_valid) {
        int rc = capture_graph(s);
        if (rc) return rc;
    }
    for (int k = 0; k < nsteps; ++k) {
        if (k % every != 0) {
            if (graph) {
                RTPS_TRY_HIP(hipGraphLaunch(s->gexec, s->stream));
            } else {
                StageTimer t(s);
                t.on = false;
                int rc = launch_step(s, t);
                if (rc) return rc;
            }
            continue;
        }
        const int set = (k / every) & 1;
        int rc = collect_timers(s, set);
        if (rc) return rc;
        if (graph) {
            RTPS_TRY_HIP(hipGraphLaunch(s->gexec_t[set], s->stream));
        } else {
            StageTimer t(s, set);
            rc = launch_step(s, t);
            if (rc) return rc;
            if (t.err != hipSuccess) return hip_error(t.err, "hipEventRecord");
        }
        s->ev_pending[set] = true;
    }
    for (int set = 0; set < 2; ++set) {
        int rc = collect_timers(s, set);
        if (rc) return rc;
    }
    return RTPS_OK;
}

float4* field_f4(rtps_sph* s, int field) {
    switch (field) {
        case RTPS_FIELD_POS: return s->pos_u;
        case RTPS_FIELD_VEL: return s->vel_u;
        case RTPS_FIELD_VELEVAL: return s->veleval_u;
        case RTPS_FIELD_POS_S: return s->pos_s;
        case RTPS_FIELD_VEL_S: return s->vel_s;
        case RTPS_FIELD_VELEVAL_S: return s->ve_s;
        case RTPS_FIELD_FORCE: return s->force_s;
        case RTPS_FIELD_XSPH: return s->xsph_s;
        case RTPS_FIELD_COLOR: return s->color_u;
        case RTPS_FIELD_COLOR_S: return s->color_s;
        default: return nullptr;
    }
}


// ---------------------------------------------------------------------------
// Autotuner: threads per block of the density and force launches and the
// force-loop variant, timed on the live sorted state.  Both kernels are
// idempotent there (density reads positions and writes veleval_s.w; the fused
// force kernel reads the sorted arrays and writes only the unsorted ones), so
// timing repeated launches and finishing with the chosen configuration is
// exactly one ordinary step.  Choices are cached per process, keyed by device
// and size class, and optionally in the text file named by RTPS_TUNE_CACHE
// ("key block_density block_force variant" per line).
struct Tune {
    int bd, bf, var;
};

std::mutex g_tune_mu;
std::map<std::string, Tune> g_tune_cache;
bool g_tune_file_loaded = false;

const char* tune_file() {
    const char* f = std::getenv("RTPS_TUNE_CACHE");
    return f && *f ? f : nullptr;
}

void tune_load_file_locked() {
    if (g_tune_file_loaded) return;
    g_tune_file_loaded = true;
    const char* f = tune_file();
    if (!f) return;
    std::ifstream in(f);
    std::string k;
    Tune t;
    auto block_ok = [](int b) { return b == 64 || b == 128 || b == 256; };
    while (in >> k >> t.bd >> t.bf >> t.var)
        if (block_ok(t.bd) && block_ok(t.bf) && (t.var == 0 || t.var == 4 || t.var == 5)) g_tune_cache[k] = t;
}

bool tune_get(const std::string& key, Tune& t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_file_locked();
    auto it = g_tune_cache.find(key);
    if (it == g_tune_cache.end()) return false;
    t = it->second;
    return true;
}

void tune_put(const std::string& key, const Tune& t) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load_file_locked();
    g_tune_cache[key] = t;
    if (const char* f = tune_file()) {
        std::ofstream out(f, std::ios::app);
        out << key << ' ' << t.bd << ' ' << t.bf << ' ' << t.var << '\n';
    }
}

std::string tune_key(const rtps_sph* s) {
    int dev = 0;
    hipDeviceProp_t prop{};
    std::string arch = "unknown";
    int cus = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        arch = prop.gcnArchName;
        cus = prop.multiProcessorCount;
        for (char& ch : arch)
            if (ch == ' ' || ch == ':') ch = '_';
    }
    unsigned long n2 = 1;
    while (n2 < (unsigned long)s->num_host) n2 <<= 1;
    return arch + "/cu" + std::to_string(cus) + "/n" + std::to_string(n2) + "/c" + std::to_string(s->nb_cells) +
           "/w" + std::to_string(s->opt_window) + "/r" + std::to_string(s->opt_remap) + "/q" + std::to_string(s->opt_queues) + "/s" +
           std::to_string(s->sp.simulation_scale != 1.0f ? 1 : 0) + "/i" + std::to_string(s->integrator);
}

void apply_tune(rtps_sph* s, const Tune& t) {
    s->block_density = t.bd;
    s->block_force = t.bf;
    s->force_variant = t.var;
    s->tuned = true;
    s->graph_valid = false;
    s->timed_graph_valid = false;
}

// One step of the system, tuning the launch configuration on the way
// (measure = false: take a cached choice when there is one).
int autotune_step(rtps_sph* s, bool measure) {
    const std::string key = tune_key(s);
    Tune t{};
    StageTimer off(s);
    off.on = false;
    if (!measure && tune_get(key, t)) {
        apply_tune(s, t);
        return launch_step(s, off);
    }
    int rc = launch_front(s, off);
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RTPS_TRY_HIP(hipEventCreateWithFlags(&e0, kTimingEvent));
    hipError_t e = hipEventCreateWithFlags(&e1, kTimingEvent);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_error(e, "hipEventCreate");
    }
    constexpr int kReps = 3;
    // average ms of kReps launches after one untimed launch; < 0 on error
    auto time_it = [&](auto&& launch) -> float {
        if (launch() != hipSuccess) return -1.f;
        if (hipEventRecord(e0, s->stream) != hipSuccess) return -1.f;
        for (int r = 0; r < kReps; ++r)
            if (launch() != hipSuccess) return -1.f;
        if (hipEventRecord(e1, s->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.f;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.f;
        return ms / kReps;
    };
    const int blocks[] = {64, 128, 256};
    // a later candidate (larger block, later loop) must be faster by this
    // factor: the first state is no guide to close calls (on the 1M dam the
    // block sizes time within 2% of each other at the first step, and 64-thread
    // blocks are 2-6% faster from a few hundred steps on)
    constexpr float kMargin = 0.95f;
    // RTPS_TUNE_VERBOSE=1 prints every timed configuration to stderr
    const char* verbose_env = std::getenv("RTPS_TUNE_VERBOSE");
    const bool verbose = verbose_env && verbose_env[0] == '1';
    const int vars[] = {5, 4, 0};   // force loops in the order they are tried
    float best_d = 1e30f;
    int bd = 256;
    bool failed = false;
    for (int bs : blocks) {
        Win w = win(s, false);
        w.block = bs;
        const float ms = time_it([&] {
            return launch_density(s->num_host, s->d_num, s->pos_s, s->start, s->end, s->sp, s->gs, w, nullptr,
                                  s->ve_s, s->stream);
        });
        if (verbose) std::fprintf(stderr, "rtps autotune: density block %d: %.4f ms\n", bs, ms);
        if (ms < 0.f) failed = true;
        else if (best_d >= 1e30f || ms < kMargin * best_d) { best_d = ms; bd = bs; }
    }
    const bool fo = s->opt_force_out != 0;
    float best_total = 1e30f;
    Tune best{256, 256, 0};
    for (int var : vars)
        for (int bs : blocks) {
            if (failed) break;
            Win w = win(s, true);
            w.block = bs;
            w.variant = var;
            const float ms = time_it([&] {
                // the euler step stores forces and leaves veleval alone: time the unfused force kernel
                if (s->integrator == 1)
                    return launch_force(s->num_host, s->d_num, s->pos_s, nullptr, s->ve_s, true, s->start, s->end,
                                        s->sp, s->gs, w, s->force_s, s->xsph_s, s->stream);
                return launch_force_fused(s->num_host, s->d_num, s->pos_s, s->ve_s, s->start, s->end, s->sp, s->gs, w,
                                          fo ? s->force_s : nullptr, fo ? s->xsph_s : nullptr, s->vel_s,
                                          s->idx_sorted, s->gsrc, s->opt_collision, s->gps, s->dt, s->pos_u,
                                          s->vel_u, s->veleval_u, s->stream);
            });
            if (verbose) std::fprintf(stderr, "rtps autotune: force loop %d block %d: %.4f ms\n", var, bs, ms);
            if (ms < 0.f) failed = true;
            else if (best_total >= 1e30f || ms + best_d < kMargin * best_total) {
                best_total = ms + best_d;
                best = Tune{bd, bs, var};
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (failed) return hip_error(hipGetLastError(), "autotune launch");
    apply_tune(s, best);
    tune_put(key, best);
    // finish the step with the chosen configuration
    RTPS_TRY_HIP(launch_density(s->num_host, s->d_num, s->pos_s, s->start, s->end, s->sp, s->gs, win(s), nullptr,
                                s->ve_s, s->stream));
    return launch_back(s);
}

}  // namespace

extern "C" {

int rtps_sph_create(const char* const* keys, const char* const* vals, int n, rtps_sph** out) {
    try {
        if (!out || n < 0 || (n > 0 && (!keys || !vals))) return set_error(RTPS_ERR_ARG, "rtps_sph_create: bad arguments");
        *out = nullptr;
        std::unique_ptr<rtps_sph> s(new rtps_sph());
        for (int i = 0; i < n; ++i) {
            if (!keys[i] || !vals[i]) return set_error(RTPS_ERR_ARG, "rtps_sph_create: null key or value");
            s->settings.SetSetting<std::string>(keys[i], vals[i]);
        }
        if (s->settings.GetSettingAs<std::string>("system") != "sph")
            return set_error(RTPS_ERR_UNKNOWN_SYSTEM,
                             "No system of type: " + s->settings.GetSettingAs<std::string>("system"));
        s->max_num = s->settings.GetSettingAs<int>("max_num_particles");
        if (s->max_num <= 0) return set_error(RTPS_ERR_ARG, "max_num_particles must be positive");
        if (integrator_of(s->settings) < 0)
            return set_error(RTPS_ERR_ARG, "unknown integrator: " + s->settings.GetSettingAs<std::string>("integrator") +
                                               " (leapfrog or euler)");
        derive_params(s->settings, s->sp, s->gp, s->gps);
        // keys are (cell << 4) | x sixteenth with cell <= nb_cells: 32 bits hold nb_cells <= 2^28 - 2
        if (s->gp.nb_cells <= 0 || s->gp.nb_cells > (1 << 28) - 2)
            return set_error(RTPS_ERR_ARG, "grid has no cells or more than 2^28 - 2 cells (" +
                                               std::to_string(s->gp.nb_cells) + ")");
        s->g = make_grid(s->gp);
        s->gs = make_grid(s->gps);
        s->nb_cells = s->gp.nb_cells;
        s->key_bits = sort_key_bits(s->nb_cells) + kSubBits;
        s->pad_key = (s->key_bits >= 32) ? 0xFFFFFFFFu : ((1u << s->key_bits) - 1u);
        apply_settings(s.get());

        RTPS_TRY_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        const size_t N = (size_t)s->max_num, C = (size_t)s->nb_cells + 1;
        RTPS_TRY_HIP(dalloc(&s->pos_u, N));
        RTPS_TRY_HIP(dalloc(&s->vel_u, N));
        RTPS_TRY_HIP(dalloc(&s->veleval_u, N));
        RTPS_TRY_HIP(dalloc(&s->pos_s, N));
        RTPS_TRY_HIP(dalloc(&s->vel_s, N));
        RTPS_TRY_HIP(dalloc(&s->ve_s, N));
        RTPS_TRY_HIP(dalloc(&s->force_s, N));
        RTPS_TRY_HIP(dalloc(&s->xsph_s, N));
        RTPS_TRY_HIP(dalloc(&s->color_u, N));
        RTPS_TRY_HIP(dalloc(&s->color_s, N));
        RTPS_TRY_HIP(dalloc(&s->keys, N));
        RTPS_TRY_HIP(dalloc(&s->idx, N));
        RTPS_TRY_HIP(dalloc(&s->start, C));
        RTPS_TRY_HIP(dalloc(&s->end, C));
        RTPS_TRY_HIP(dalloc(&s->d_num, 1));
        RTPS_TRY_HIP(dalloc(&s->tile_q, kTileQueueWords));
        RTPS_TRY_HIP(hipMemsetAsync(s->tile_q, 0, kTileQueueWords * sizeof(int), s->stream));
        if ((size_t)s->nb_cells * kSubCells * sizeof(uint32_t) <= ((size_t)2 << 30))   // at most 2 GiB
            RTPS_TRY_HIP(dalloc(&s->sub_start, (size_t)s->nb_cells * kSubCells));
        // the bucket sort's temp, a superset of the radix sort's, for every count up to N
        RTPS_TRY_HIP(hipMalloc(&s->sort_tmp, bucket_temp_bytes((int)N, s->key_bits)));
        RTPS_TRY_HIP(dalloc(&s->rows, 27 * N + 18 * ((N + 63) / 64)));
        float4* zf[] = {s->pos_u, s->vel_u, s->veleval_u, s->pos_s, s->vel_s, s->ve_s, s->force_s, s->xsph_s};
        for (float4* p : zf) RTPS_TRY_HIP(hipMemsetAsync(p, 0, N * sizeof(float4), s->stream));
        {   // every slot starts with the default colour of pushParticles (System.h:160)
            std::vector<rtps_float4> red(N, kDefaultColor);
            RTPS_TRY_HIP(hipMemcpyAsync(s->color_u, red.data(), N * sizeof(float4), hipMemcpyHostToDevice, s->stream));
            RTPS_TRY_HIP(hipMemcpyAsync(s->color_s, red.data(), N * sizeof(float4), hipMemcpyHostToDevice, s->stream));
            RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
        }
        RTPS_TRY_HIP(hipMemsetAsync(s->keys, 0, N * 4, s->stream));
        RTPS_TRY_HIP(hipMemsetAsync(s->idx, 0, N * 4, s->stream));
        RTPS_TRY_HIP(hipMemsetAsync(s->d_num, 0, sizeof(int), s->stream));
        // cell arrays start at 666 (SPH.cpp:512-517)
        RTPS_TRY_HIP(launch_fill_u32(s->start, 666u, (long)C, s->stream));
        RTPS_TRY_HIP(launch_fill_u32(s->end, 666u, (long)C, s->stream));
        RTPS_TRY_HIP(sort_clear_temp(s->sort_tmp, (int)N, s->stream));
        s->keys_sorted = s->keys;
        s->idx_sorted = s->idx;
        for (auto& set : s->ev)
            for (auto& e : set) RTPS_TRY_HIP(hipEventCreateWithFlags(&e, kTimingEvent));
        RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
        *out = s.release();
        return RTPS_OK;
    } catch (const std::exception& e) {
        return set_error(RTPS_ERR_ARG, std::string("rtps_sph_create: ") + e.what());
    }
}

int rtps_sph_destroy(rtps_sph* s) {
    delete s;
    return RTPS_OK;
}

int rtps_sph_get_num(rtps_sph* s, int* num) {
    if (!s || !num) return set_error(RTPS_ERR_ARG, "rtps_sph_get_num: null argument");
    int v = 0;
    RTPS_TRY_HIP(hipMemcpyAsync(&v, s->d_num, sizeof(int), hipMemcpyDeviceToHost, s->stream));
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    *num = v;
    return RTPS_OK;
}

int rtps_sph_push_particles_color(rtps_sph* s, const rtps_float4* pos, const rtps_float4* vel, int n,
                                  rtps_float4 color) {
    if (!s || (!pos && n > 0) || n < 0) return set_error(RTPS_ERR_ARG, "rtps_sph_push_particles: bad arguments");
    if (n == 0) return RTPS_OK;
    int num = 0;
    int rc = rtps_sph_get_num(s, &num);
    if (rc) return rc;
    if ((long)num + n > s->max_num)
        return set_error(RTPS_ERR_CAPACITY, "pushParticles: exceeded max nb (" + std::to_string(s->max_num) +
                                                ") of particles allowed");
    // SPH::pushParticles (SPH.cpp:560-602): positions, colours, velocities
    const bool other = color.x != kDefaultColor.x || color.y != kDefaultColor.y || color.z != kDefaultColor.z ||
                       color.w != kDefaultColor.w;
    if (other && !s->opt_color) {   // from now on colours follow their particles
        s->opt_color = 1;
        s->graph_valid = false;
        s->timed_graph_valid = false;
    }
    std::vector<rtps_float4> cols((size_t)n, color);
    RTPS_TRY_HIP(hipMemcpyAsync(s->pos_u + num, pos, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
    RTPS_TRY_HIP(hipMemcpyAsync(s->color_u + num, cols.data(), (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
    if (vel) RTPS_TRY_HIP(hipMemcpyAsync(s->vel_u + num, vel, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
    else RTPS_TRY_HIP(hipMemsetAsync(s->vel_u + num, 0, (size_t)n * 16, s->stream));
    num += n;
    RTPS_TRY_HIP(hipMemcpyAsync(s->d_num, &num, sizeof(int), hipMemcpyHostToDevice, s->stream));
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    s->num_host = num;
    s->settings.SetSetting("num_particles", num);   // SPH.cpp:588-589
    apply_settings(s);
    return RTPS_OK;
}

int rtps_sph_push_particles(rtps_sph* s, const rtps_float4* pos, const rtps_float4* vel, int n) {
    return rtps_sph_push_particles_color(s, pos, vel, n, kDefaultColor);
}

namespace {

// SPH::sprayHoses (SPH.cpp:547-557): each hose's spray is pushed with the
// hose velocity.  A push that would exceed max_num is dropped, as
// SPH::pushParticles does (SPH.cpp:563-567); the hose still counts it sprayed.
int spray_hoses(rtps_sph* s) {
    for (size_t k = 0; k < s->hoses.size(); ++k) {
        rtps_hose* h = s->hoses[k];
        std::vector<rtps_float4> p = hose_spray(h);
        if (p.empty()) continue;
        rtps_float4 v{};
        rtps_hose_state(h, nullptr, nullptr, nullptr, nullptr, nullptr, &v);
        int num = 0;
        int rc = rtps_sph_get_num(s, &num);
        if (rc) return rc;
        if ((long)num + (long)p.size() > s->max_num) continue;
        std::vector<rtps_float4> vel(p.size(), v);
        rc = rtps_sph_push_particles_color(s, p.data(), vel.data(), (int)p.size(), s->hose_color[k]);
        if (rc) return rc;
    }
    return RTPS_OK;
}

int hose_at(rtps_sph* s, int index, rtps_hose** h) {
    if (!s || index < 0 || index >= (int)s->hoses.size())
        return set_error(RTPS_ERR_ARG, "no hose with index " + std::to_string(index));
    *h = s->hoses[index];
    return RTPS_OK;
}

}  // namespace

int rtps_sph_add_hose(rtps_sph* s, int total_n, rtps_float4 center, rtps_float4 velocity, float radius,
                      rtps_float4 color, int* index) {
    if (!s || !index) return set_error(RTPS_ERR_ARG, "rtps_sph_add_hose: null argument");
    // SPH::addHose (SPH.cpp:523-532): radius in units of the particle spacing
    const float spacing = s->settings.GetSettingAs<float>("spacing");
    rtps_hose* h = nullptr;
    const unsigned long long seed = 0x5EEDull + 0x9E3779B97F4A7C15ull * (unsigned long long)s->hoses.size();
    int rc = rtps_hose_create(s->settings.GetSettingAs<float>("time_step"), total_n, center, velocity,
                              radius * spacing, spacing, seed, &h);
    if (rc) return rc;
    s->hoses.push_back(h);
    s->hose_color.push_back(color);
    *index = (int)s->hoses.size() - 1;
    return RTPS_OK;
}

int rtps_sph_update_hose(rtps_sph* s, int index, rtps_float4 center, rtps_float4 velocity, float radius,
                         rtps_float4 color) {
    rtps_hose* h = nullptr;
    int rc = hose_at(s, index, &h);
    if (rc) return rc;
    s->hose_color[index] = color;
    const float spacing = s->settings.GetSettingAs<float>("spacing");   // SPH::updateHose (SPH.cpp:533-541)
    return rtps_hose_update(h, center, velocity, radius * spacing, spacing);
}

int rtps_sph_refill_hose(rtps_sph* s, int index, int refill) {
    rtps_hose* h = nullptr;
    int rc = hose_at(s, index, &h);
    if (rc) return rc;
    return rtps_hose_refill(h, refill);   // SPH::refillHose (SPH.cpp:542-545)
}

int rtps_sph_load_triangles(rtps_sph* s, const rtps_triangle* tris, int n) {
    if (!s || n < 0 || (n > 0 && !tris)) return set_error(RTPS_ERR_ARG, "rtps_sph_load_triangles: bad arguments");
    s->triangles.assign(tris, tris + n);
    maybe_update_params(s);
    RTPS_TRY_HIP(upload_triangles(s));
    s->graph_valid = false;
    s->timed_graph_valid = false;
    return RTPS_OK;
}

int rtps_sph_test_delete(rtps_sph* s) {
    if (!s) return set_error(RTPS_ERR_ARG, "rtps_sph_test_delete: null handle");
    // System::testDelete (System.cpp:366-376): 40 particles moved outside the grid
    if (s->max_num < 40) return set_error(RTPS_ERR_CAPACITY, "testDelete needs max_num >= 40");
    std::vector<rtps_float4> far(40, rtps_float4{100.f, 100.f, 100.f, 1.f});
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    RTPS_TRY_HIP(hipMemcpy(s->pos_u, far.data(), far.size() * sizeof(rtps_float4), hipMemcpyHostToDevice));
    return RTPS_OK;
}

int rtps_sph_update(rtps_sph* s) {
    if (!s) return set_error(RTPS_ERR_ARG, "rtps_sph_update: null handle");
    maybe_update_params(s);
    if (!s->hoses.empty()) {
        int rc = spray_hoses(s);   // SPH.cpp:166
        if (rc) return rc;
    }
    StageTimer t(s);
    t.on = false;
    int rc = launch_front(s, t);
    if (rc) return rc;
    RTPS_TRY_HIP(launch_force(s->num_host, s->d_num, s->pos_s, nullptr, s->ve_s, true, s->start, s->end, s->sp, s->gs,
                              win(s, true), s->force_s, s->xsph_s, s->stream));
    RTPS_TRY_HIP(launch_post_force(s->num_host, s->d_num, s->start, s->nb_cells, s->pos_s, s->vel_s, s->force_s,
                                   s->sp, s->gsrc, s->opt_collision, s->gps, s->stream));
    if (s->ntri > 0 && s->opt_collision)
        RTPS_TRY_HIP(launch_collision_tri(s->num_host, s->d_num, s->start, s->nb_cells, s->pos_s, s->vel_s,
                                          s->force_s, s->tri_dev, s->tri_reach, s->ntri, s->sp, s->stream));
    s->update_pending = true;
    return RTPS_OK;
}

int rtps_sph_integrate(rtps_sph* s) {
    if (!s) return set_error(RTPS_ERR_ARG, "rtps_sph_integrate: null handle");
    if (!s->update_pending) return set_error(RTPS_ERR_STATE, "integrate() needs a preceding update()");
    int rc = launch_integrate(s);   // SPH::integrate (SPH.cpp:392-409)
    if (rc) return rc;
    RTPS_TRY_HIP(launch_commit_num(s->d_num, s->start, s->nb_cells, s->stream));
    s->update_pending = false;
    return RTPS_OK;
}

int rtps_sph_interact(rtps_sph* s) {
    if (!s) return set_error(RTPS_ERR_ARG, "rtps_sph_interact: null handle");
    if (!s->update_pending) return set_error(RTPS_ERR_STATE, "interact() needs a preceding update()");
    return launch_interact(s);
}

namespace {

// Loads set k (0 = moving body, 1 = static boundary) from host arrays.
int load_rigid(rtps_sph* s, int k, const rtps_float4* pos, const rtps_float4* vel, const float* mass, int n,
               float spring, float dampening) {
    RigidDev& r = s->rigid[k];
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    s->graph_valid = false;
    s->timed_graph_valid = false;
    if (n == 0) {
        r.n = 0;
        return RTPS_OK;
    }
    if (!r.start) {
        RTPS_TRY_HIP(dalloc(&r.start, (size_t)s->nb_cells + 1));
        RTPS_TRY_HIP(dalloc(&r.end, (size_t)s->nb_cells + 1));
        RTPS_TRY_HIP(launch_fill_u32(r.end, 666u, (long)s->nb_cells + 1, s->stream));   // as SPH.cpp:512-517
    }
    if (n > r.cap) {
        void* old[] = {r.pos, r.vel, r.in_pos, r.in_vel, r.mass, r.in_mass, r.keys, r.idx, r.sort_tmp};
        for (void* b : old)
            if (b) RTPS_TRY_HIP(hipFree(b));
        r = RigidDev{0, 0, false, 0.f, 0.f, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr, nullptr, r.start, r.end, nullptr};
        const size_t N = (size_t)n;
        RTPS_TRY_HIP(dalloc(&r.pos, N));
        RTPS_TRY_HIP(dalloc(&r.vel, N));
        RTPS_TRY_HIP(dalloc(&r.in_pos, N));
        RTPS_TRY_HIP(dalloc(&r.in_vel, N));
        RTPS_TRY_HIP(dalloc(&r.mass, N));
        RTPS_TRY_HIP(dalloc(&r.in_mass, N));
        RTPS_TRY_HIP(dalloc(&r.keys, N));
        RTPS_TRY_HIP(dalloc(&r.idx, N));
        RTPS_TRY_HIP(hipMalloc(&r.sort_tmp, sort_temp_bytes(n, 32)));
        RTPS_TRY_HIP(sort_clear_temp(r.sort_tmp, n, s->stream));
        r.cap = n;
    }
    RTPS_TRY_HIP(hipMemcpyAsync(r.in_pos, pos, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
    if (vel) RTPS_TRY_HIP(hipMemcpyAsync(r.in_vel, vel, (size_t)n * 16, hipMemcpyHostToDevice, s->stream));
    if (mass) RTPS_TRY_HIP(hipMemcpyAsync(r.in_mass, mass, (size_t)n * 4, hipMemcpyHostToDevice, s->stream));
    r.moving = k == 0;
    r.s0 = spring;
    r.s1 = dampening;
    RTPS_TRY_HIP(rigid_prepare(n, r.in_pos, r.moving ? r.in_vel : nullptr, r.moving ? r.in_mass : nullptr, r, s->g,
                               s->stream));
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    r.n = n;
    return RTPS_OK;
}

}  // namespace

int rtps_sph_set_static_boundary(rtps_sph* s, const rtps_float4* pos, int n, float spring, float dampening) {
    if (!s || n < 0 || (n > 0 && !pos))
        return set_error(RTPS_ERR_ARG, "rtps_sph_set_static_boundary: bad arguments");
    return load_rigid(s, 1, pos, nullptr, nullptr, n, spring, dampening);
}

int rtps_sph_set_rigid_particles(rtps_sph* s, const rtps_float4* pos, const rtps_float4* vel, const float* mass,
                                 int n, float spring, float dampening) {
    if (!s || n < 0 || (n > 0 && (!pos || !vel || !mass)))
        return set_error(RTPS_ERR_ARG, "rtps_sph_set_rigid_particles: bad arguments");
    for (int i = 0; i < n; ++i)
        if (!(mass[i] > 0.f)) return set_error(RTPS_ERR_ARG, "rtps_sph_set_rigid_particles: masses must be > 0");
    return load_rigid(s, 0, pos, vel, mass, n, spring, dampening);
}

namespace {

// eager: launch the kernels directly even when graphs are on (a step right
// after a push, whose graph would be captured for one replay only)
int run_steps(rtps_sph* s, int nsteps, bool eager = false) {
    if (s->num_host == 0) return RTPS_OK;
    if (s->opt_autotune && !s->tuned && nsteps > 0) {   // the first step also tunes the launches
        int rc = autotune_step(s, false);
        if (rc) return rc;
        if (--nsteps == 0) return RTPS_OK;
    }
    if (s->opt_timers) return step_timed(s, nsteps);
    if (s->opt_graph && !eager) {
        if (!s->graph_valid) {
            int rc = capture_graph(s);
            if (rc) return rc;
        }
        for (int k = 0; k < nsteps; ++k) RTPS_TRY_HIP(hipGraphLaunch(s->gexec, s->stream));
        return RTPS_OK;
    }
    for (int k = 0; k < nsteps; ++k) {
        StageTimer t(s);
        t.on = false;
        int rc = launch_step(s, t);
        if (rc) return rc;
    }
    return RTPS_OK;
}

}  // namespace

int rtps_sph_step(rtps_sph* s, int nsteps) {
    if (!s || nsteps < 0) return set_error(RTPS_ERR_ARG, "rtps_sph_step: bad arguments");
    if (s->update_pending) return set_error(RTPS_ERR_STATE, "rtps_sph_step after update() without integrate()");
    maybe_update_params(s);
    if (s->hoses.empty()) return run_steps(s, nsteps);
    for (int k = 0; k < nsteps; ++k) {   // hoses spray on the host before every update
        const int before = s->num_host;
        int rc = spray_hoses(s);
        if (rc) return rc;
        rc = run_steps(s, 1, s->num_host != before);
        if (rc) return rc;
    }
    return RTPS_OK;
}

int rtps_sph_synchronize(rtps_sph* s) {
    if (!s) return set_error(RTPS_ERR_ARG, "null handle");
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    return RTPS_OK;
}

int rtps_sph_read(rtps_sph* s, int field, void* host, int count) {
    if (!s || !host || count < 0) return set_error(RTPS_ERR_ARG, "rtps_sph_read: bad arguments");
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    if (field == RTPS_FIELD_CELL_START || field == RTPS_FIELD_CELL_END) {
        if (count > s->nb_cells + 1) return set_error(RTPS_ERR_ARG, "rtps_sph_read: count > nb_cells + 1");
        RTPS_TRY_HIP(hipMemcpy(host, field == RTPS_FIELD_CELL_START ? s->start : s->end, (size_t)count * 4,
                               hipMemcpyDeviceToHost));
        return RTPS_OK;
    }
    if (count > s->max_num) return set_error(RTPS_ERR_ARG, "rtps_sph_read: count > max_num");
    if (field == RTPS_FIELD_KEYS || field == RTPS_FIELD_IDX) {
        RTPS_TRY_HIP(hipMemcpy(host, field == RTPS_FIELD_KEYS ? s->keys_sorted : s->idx_sorted, (size_t)count * 4,
                               hipMemcpyDeviceToHost));
        if (field == RTPS_FIELD_KEYS) {   // internal keys carry kSubBits of in-cell x; return the cell keys
            uint32_t* k = (uint32_t*)host;
            for (int i = 0; i < count; ++i) k[i] >>= kSubBits;
        }
        return RTPS_OK;
    }
    if (field == RTPS_FIELD_MASS || field == RTPS_FIELD_MASS_S) {
        // every particle carries the system mass (SPH.cpp:572-573 fills mass_p with sphp.mass)
        float* m = (float*)host;
        for (int i = 0; i < count; ++i) m[i] = s->sp.mass;
        return RTPS_OK;
    }
    if (field == RTPS_FIELD_DENSITY) {
        std::vector<float> tmp((size_t)count * 4);
        RTPS_TRY_HIP(hipMemcpy(tmp.data(), s->ve_s, (size_t)count * 16, hipMemcpyDeviceToHost));
        float* d = (float*)host;
        for (int i = 0; i < count; ++i) d[i] = tmp[4 * (size_t)i + 3];
        return RTPS_OK;
    }
    float4* src = field_f4(s, field);
    if (!src) return set_error(RTPS_ERR_ARG, "rtps_sph_read: unknown field");
    RTPS_TRY_HIP(hipMemcpy(host, src, (size_t)count * 16, hipMemcpyDeviceToHost));
    return RTPS_OK;
}

namespace {

// Buffers of the colour field and the surface for a res^3 lattice.
int surface_buffers(rtps_sph* s, int res) {
    if (!s->mc_words) {   // case table, once per handle
        const McTable& t = mc_table();
        RTPS_TRY_HIP(mc_upload_geometry());
        RTPS_TRY_HIP(hipMalloc(&s->mc_words, sizeof(kMcCaseWords)));
        RTPS_TRY_HIP(hipMalloc(&s->mc_ntri, sizeof(t.ntri)));
        RTPS_TRY_HIP(hipMemcpy(s->mc_words, kMcCaseWords, sizeof(kMcCaseWords), hipMemcpyHostToDevice));
        RTPS_TRY_HIP(hipMemcpy(s->mc_ntri, t.ntri, sizeof(t.ntri), hipMemcpyHostToDevice));
        RTPS_TRY_HIP(hipMalloc(&s->mc_total, sizeof(uint32_t)));
    }
    if (s->cf_res == res) return RTPS_OK;
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    void* old[] = {s->cf_field, s->mc_cases, s->mc_counts, s->mc_offs, s->mc_tiles};
    for (void* b : old)
        if (b) RTPS_TRY_HIP(hipFree(b));
    s->cf_field = nullptr;
    s->mc_cases = nullptr;
    s->mc_counts = s->mc_offs = s->mc_tiles = nullptr;
    s->cf_res = 0;
    const size_t nodes = (size_t)res * res * res, m = (size_t)(res - 1), cubes = m * m * m;
    RTPS_TRY_HIP(hipMalloc(&s->cf_field, nodes * sizeof(float)));
    RTPS_TRY_HIP(hipMalloc(&s->mc_cases, cubes));
    RTPS_TRY_HIP(hipMalloc(&s->mc_counts, cubes * sizeof(uint32_t)));
    RTPS_TRY_HIP(hipMalloc(&s->mc_offs, cubes * sizeof(uint32_t)));
    RTPS_TRY_HIP(hipMalloc(&s->mc_tiles, surface_scan_tiles((long)cubes) * sizeof(uint32_t)));
    s->cf_res = res;
    return RTPS_OK;
}

// ColorField::execute (ColorField.cpp) on the last step's sorted state.
int color_field(rtps_sph* s, int res) {
    if (res < 2 || res > 1024) return set_error(RTPS_ERR_ARG, "colour field resolution must be 2..1024");
    if (!s->sorted_ready) return set_error(RTPS_ERR_STATE, "the colour field needs a preceding update or step");
    if (s->update_pending) return set_error(RTPS_ERR_STATE, "the colour field needs integrate() after update()");
    int rc = surface_buffers(s, res);
    if (rc) return rc;
    RTPS_TRY_HIP(launch_color_field(res, s->pos_s, s->ve_s, s->start, s->end, s->sp, s->gs, win(s), s->gps.bnd_min,
                                    s->gps.bnd_max, s->cf_field, s->stream));
    return RTPS_OK;
}

// MarchingCubes::execute: the iso-surface of the colour field into the
// handle's mesh buffers (grown to the triangle count).
int surface(rtps_sph* s, int res, float iso) {
    int rc = color_field(s, res);
    if (rc) return rc;
    RTPS_TRY_HIP(launch_mc_count(res, iso, s->cf_field, s->mc_ntri, s->mc_cases, s->mc_counts, s->mc_offs,
                                 s->mc_tiles, s->mc_total, s->stream));
    uint32_t total = 0;
    RTPS_TRY_HIP(hipMemcpyAsync(&total, s->mc_total, sizeof(total), hipMemcpyDeviceToHost, s->stream));
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    if ((long)total > s->mesh_cap) {
        if (s->mesh_v) RTPS_TRY_HIP(hipFree(s->mesh_v));
        if (s->mesh_n) RTPS_TRY_HIP(hipFree(s->mesh_n));
        s->mesh_v = s->mesh_n = nullptr;
        s->mesh_cap = 0;
        const long cap = (long)total + total / 4 + 64;
        RTPS_TRY_HIP(hipMalloc(&s->mesh_v, (size_t)cap * 9 * sizeof(float)));
        RTPS_TRY_HIP(hipMalloc(&s->mesh_n, (size_t)cap * 9 * sizeof(float)));
        s->mesh_cap = cap;
    }
    s->mesh_ntri = (int)total;
    if (total > 0)
        RTPS_TRY_HIP(launch_mc_emit(res, iso, s->cf_field, s->mc_cases, s->mc_counts, s->mc_offs, s->mc_tiles,
                                    s->mc_words, s->gp.bnd_min, s->gp.bnd_max, s->opt_frame, s->mesh_cap, s->mesh_v,
                                    s->mesh_n,
                                    s->stream));
    return RTPS_OK;
}

int copy_mesh(rtps_sph* s, int capacity, float* vertices, float* normals, int* ntris) {
    if (ntris) *ntris = s->mesh_ntri;
    const int n = std::min(capacity, s->mesh_ntri);
    if (n > 0 && vertices)
        RTPS_TRY_HIP(hipMemcpyAsync(vertices, s->mesh_v, (size_t)n * 9 * sizeof(float), hipMemcpyDeviceToHost,
                                    s->stream));
    if (n > 0 && normals)
        RTPS_TRY_HIP(hipMemcpyAsync(normals, s->mesh_n, (size_t)n * 9 * sizeof(float), hipMemcpyDeviceToHost,
                                    s->stream));
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    return RTPS_OK;
}

}  // namespace

int rtps_mc_case_table(signed char* tri, unsigned char* ntri) {
    const McTable& t = mc_table();
    if (tri) std::memcpy(tri, t.tri, sizeof(t.tri));
    if (ntri) std::memcpy(ntri, t.ntri, sizeof(t.ntri));
    return RTPS_OK;
}

int rtps_mc_geometry(signed char* corners, signed char* edges) {
    if (corners) std::memcpy(corners, kMcCorner, sizeof(kMcCorner));
    if (edges) std::memcpy(edges, kMcEdgeEnds, sizeof(kMcEdgeEnds));
    return RTPS_OK;
}

int rtps_sph_color_field(rtps_sph* s, int res, float* field) {
    if (!s) return set_error(RTPS_ERR_ARG, "rtps_sph_color_field: null handle");
    int rc = color_field(s, res);
    if (rc) return rc;
    if (field)
        RTPS_TRY_HIP(hipMemcpyAsync(field, s->cf_field, (size_t)res * res * res * sizeof(float), hipMemcpyDeviceToHost,
                                    s->stream));
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    return RTPS_OK;
}

int rtps_sph_surface(rtps_sph* s, int res, float iso, int capacity, float* vertices, float* normals, int* ntris) {
    if (!s || capacity < 0) return set_error(RTPS_ERR_ARG, "rtps_sph_surface: bad arguments");
    int rc = surface(s, res, iso);
    if (rc) return rc;
    return copy_mesh(s, capacity, vertices, normals, ntris);
}

int rtps_sph_post_process(rtps_sph* s) {
    if (!s) return set_error(RTPS_ERR_ARG, "rtps_sph_post_process: null handle");
    // SPH::postProcess (SPH.cpp:338-386): only with "use_color_field" and particles
    const bool use = s->settings.GetSettingAs<bool>("use_color_field", "0");
    if (!use || s->num_host == 0 || !s->sorted_ready) {
        s->mesh_ntri = 0;
        return RTPS_OK;
    }
    const int res = std::max(2, (int)s->settings.GetSettingAs<unsigned>("color_field_res", "2"));
    const float iso = s->settings.GetSettingAs<float>("color_field_iso", "0.0001");   // MarchingCubes.cpp isolevel
    return surface(s, res, iso);
}

int rtps_sph_mesh(rtps_sph* s, int capacity, float* vertices, float* normals, int* ntris) {
    if (!s || capacity < 0) return set_error(RTPS_ERR_ARG, "rtps_sph_mesh: bad arguments");
    return copy_mesh(s, capacity, vertices, normals, ntris);
}

int rtps_host_alloc(size_t bytes, void** host) {
    if (!host) return set_error(RTPS_ERR_ARG, "rtps_host_alloc: null argument");
    *host = nullptr;
    RTPS_TRY_HIP(hipHostMalloc(host, bytes ? bytes : 1, hipHostMallocDefault));
    return RTPS_OK;
}

int rtps_host_free(void* host) {
    if (host) RTPS_TRY_HIP(hipHostFree(host));
    return RTPS_OK;
}

int rtps_sph_read_async(rtps_sph* s, int field, void* host, int count) {
    if (!s || !host || count < 0) return set_error(RTPS_ERR_ARG, "rtps_sph_read_async: bad arguments");
    if (count == 0) return RTPS_OK;
    const void* src = nullptr;
    size_t bytes = 0;
    if (field == RTPS_FIELD_CELL_START || field == RTPS_FIELD_CELL_END) {
        if (count > s->nb_cells + 1) return set_error(RTPS_ERR_ARG, "rtps_sph_read_async: count > nb_cells + 1");
        src = field == RTPS_FIELD_CELL_START ? s->start : s->end;
        bytes = (size_t)count * 4;
    } else if (count > s->max_num) {
        return set_error(RTPS_ERR_ARG, "rtps_sph_read_async: count > max_num");
    } else if (field == RTPS_FIELD_IDX) {
        src = s->idx_sorted;
        bytes = (size_t)count * 4;
    } else {
        src = field_f4(s, field);
        bytes = (size_t)count * 16;
    }
    if (!src)
        return set_error(RTPS_ERR_ARG, "rtps_sph_read_async: only float4 fields, IDX and the cell arrays (KEYS, "
                                       "DENSITY and MASS need host-side conversion: use rtps_sph_read)");
    RTPS_TRY_HIP(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, s->stream));
    return RTPS_OK;
}

int rtps_sph_write(rtps_sph* s, int field, const void* host, int count) {
    if (!s || !host || count < 0 || count > s->max_num) return set_error(RTPS_ERR_ARG, "rtps_sph_write: bad arguments");
    if (field != RTPS_FIELD_POS && field != RTPS_FIELD_VEL && field != RTPS_FIELD_VELEVAL && field != RTPS_FIELD_COLOR)
        return set_error(RTPS_ERR_ARG, "rtps_sph_write: only POS, VEL, VELEVAL and COLOR are writable");
    if (field == RTPS_FIELD_COLOR && !s->opt_color) {   // written colours are kept with their particles
        s->opt_color = 1;
        s->graph_valid = false;
        s->timed_graph_valid = false;
    }
    RTPS_TRY_HIP(hipStreamSynchronize(s->stream));
    RTPS_TRY_HIP(hipMemcpy(field_f4(s, field), host, (size_t)count * 16, hipMemcpyHostToDevice));
    return RTPS_OK;
}

int rtps_sph_device_ptr(rtps_sph* s, int field, void** dev) {
    if (!s || !dev) return set_error(RTPS_ERR_ARG, "rtps_sph_device_ptr: null argument");
    switch (field) {
        case RTPS_FIELD_KEYS: *dev = s->keys_sorted; return RTPS_OK;
        case RTPS_FIELD_IDX: *dev = s->idx_sorted; return RTPS_OK;
        case RTPS_FIELD_CELL_START: *dev = s->start; return RTPS_OK;
        case RTPS_FIELD_CELL_END: *dev = s->end; return RTPS_OK;
        case RTPS_FIELD_DENSITY: return set_error(RTPS_ERR_ARG, "density lives in VELEVAL_S .w (packed)");
        case RTPS_FIELD_MASS:
        case RTPS_FIELD_MASS_S: return set_error(RTPS_ERR_ARG, "masses are not stored (every particle has sphp.mass)");
        default: break;
    }
    float4* p = field_f4(s, field);
    if (!p) return set_error(RTPS_ERR_ARG, "rtps_sph_device_ptr: unknown field");
    *dev = p;
    return RTPS_OK;
}

int rtps_sph_set_setting(rtps_sph* s, const char* key, const char* val) {
    if (!s || !key || !val) return set_error(RTPS_ERR_ARG, "rtps_sph_set_setting: null argument");
    if (std::strcmp(key, "integrator") == 0 && std::strcmp(val, "leapfrog") != 0 && std::strcmp(val, "euler") != 0)
        return set_error(RTPS_ERR_ARG, std::string("unknown integrator: ") + val + " (leapfrog or euler)");
    s->settings.SetSetting<std::string>(key, val);
    return RTPS_OK;
}

int rtps_sph_get_setting(rtps_sph* s, const char* key, char* buf, int buflen) {
    if (!s || !key || !buf || buflen <= 0) return set_error(RTPS_ERR_ARG, "rtps_sph_get_setting: bad arguments");
    if (!s->settings.Exists(key)) return set_error(RTPS_ERR_ARG, std::string("no setting ") + key);
    std::string v = s->settings.GetSettingAs<std::string>(key);
    std::strncpy(buf, v.c_str(), (size_t)buflen - 1);
    buf[buflen - 1] = 0;
    return RTPS_OK;
}

int rtps_sph_get_params(rtps_sph* s, rtps_sph_params* sp, rtps_grid_params* gp, rtps_grid_params* gps) {
    if (!s) return set_error(RTPS_ERR_ARG, "null handle");
    maybe_update_params(s);
    if (sp) *sp = s->sp;
    if (gp) *gp = s->gp;
    if (gps) *gps = s->gps;
    return RTPS_OK;
}

int rtps_sph_set_option(rtps_sph* s, int option, int value) {
    if (!s) return set_error(RTPS_ERR_ARG, "null handle");
    switch (option) {
        case RTPS_OPT_COLLISION: s->opt_collision = value ? 1 : 0; break;
        case RTPS_OPT_GRAPH: s->opt_graph = value ? 1 : 0; break;
        case RTPS_OPT_FORCE_OUT: s->opt_force_out = value ? 1 : 0; break;
        case RTPS_OPT_TIMERS:
            if (value < 0 || value > 3) return set_error(RTPS_ERR_ARG, "timers option must be 0, 1, 2 or 3");
            if (value != s->opt_timers) s->timed_graph_valid = false;   // the timed graphs hold the events
            s->opt_timers = value;
            break;
        case RTPS_OPT_FORCE_VARIANT:
            if (value != 0 && value != 4 && value != 5)
                return set_error(RTPS_ERR_ARG, "force loop must be 0 (direct), 4 (staged rows) or 5 (staged rows "
                                               "with accepted bits)");
            g_force_variant = value;
            break;
        case RTPS_OPT_WINDOW:
            if (value < 0 || value > 2) return set_error(RTPS_ERR_ARG, "window option must be 0, 1 or 2");
            s->opt_window = value;
            break;
        case RTPS_OPT_TILE_REMAP:
            if (value < 0 || value > 4096) return set_error(RTPS_ERR_ARG, "tile remap must be 0, 1 or a chunk of 2..4096");
            s->opt_remap = value;
            s->graph_valid = false;
            s->timed_graph_valid = false;
            break;
        case RTPS_OPT_SORT_TILE:
            if (value != 0 && !sort_config_valid(value / 100, value % 100))
                return set_error(RTPS_ERR_ARG, "sort tile must be 0 (by size) or threads * 100 + keys per thread: "
                                               "102401, 102402, 102404, 102408, 51204, 51208, 25604, 25608, 25616");
            g_sort_items = value;
            break;
        case RTPS_OPT_SURFACE_FRAME:
            if (value != 0 && value != 1) return set_error(RTPS_ERR_ARG, "surface frame must be 0 (world) or 1 (reference)");
            s->opt_frame = value;
            break;
        case RTPS_OPT_BUCKET_SORT:
            if (value != 0 && value != 1) return set_error(RTPS_ERR_ARG, "bucket sort option must be 0 or 1");
            s->opt_bucket = value;
            s->graph_valid = false;
            s->timed_graph_valid = false;
            break;
        case RTPS_OPT_SHARED_ROWS:
            if (value != 0 && value != 1) return set_error(RTPS_ERR_ARG, "shared rows option must be 0 or 1");
            s->opt_rows = value;
            break;
        case RTPS_OPT_TILE_QUEUES:
            if (value != 0 && value != 1) return set_error(RTPS_ERR_ARG, "tile queues option must be 0 or 1");
            s->opt_queues = value;
            s->graph_valid = false;
            s->timed_graph_valid = false;
            break;
        case RTPS_OPT_BLOCK_DENSITY:
        case RTPS_OPT_BLOCK_FORCE:
            if (value != 64 && value != 128 && value != 256)
                return set_error(RTPS_ERR_ARG, "threads per block must be 64, 128 or 256");
            if (option == RTPS_OPT_BLOCK_DENSITY) s->block_density = value; else s->block_force = value;
            s->force_variant = -1;
            s->tuned = true;
            break;
        case RTPS_OPT_DENSITY_STAGE:
            if (value < 0 || value > 1) return set_error(RTPS_ERR_ARG, "density stage option must be 0 or 1");
            g_density_stage = value;
            break;
        case RTPS_OPT_COLOR:
            if (value < 0 || value > 1) return set_error(RTPS_ERR_ARG, "colour option must be 0 or 1");
            s->opt_color = value;
            break;
        case RTPS_OPT_AUTOTUNE:
            if (value < 0 || value > 1) return set_error(RTPS_ERR_ARG, "autotune must be 0 or 1");
            s->opt_autotune = value;
            if (!value) apply_tune(s, Tune{0, 0, -1});
            s->tuned = false;
            break;
        default: return set_error(RTPS_ERR_ARG, "unknown option");
    }
    s->graph_valid = false;
    s->timed_graph_valid = false;
    return RTPS_OK;
}

int rtps_sph_autotune(rtps_sph* s, int measure, int* block_density, int* block_force, int* force_variant) {
    if (!s) return set_error(RTPS_ERR_ARG, "null handle");
    if (s->update_pending) return set_error(RTPS_ERR_STATE, "rtps_sph_autotune after update() without integrate()");
    maybe_update_params(s);
    if (s->num_host > 0) {
        int rc = autotune_step(s, measure != 0);
        if (rc) return rc;
    }
    if (block_density) *block_density = s->block_density;
    if (block_force) *block_force = s->block_force;
    if (force_variant) *force_variant = s->force_variant;
    return RTPS_OK;
}

int rtps_sph_stage_times(rtps_sph* s, float* avg_ms, int* counts, int n) {
    if (!s || n < 0) return set_error(RTPS_ERR_ARG, "bad arguments");
    for (int i = 0; i < n && i < ST_N; ++i) {
        if (avg_ms) avg_ms[i] = s->stage_cnt[i] ? (float)(s->stage_ms[i] / s->stage_cnt[i]) : 0.f;
        if (counts) counts[i] = (int)s->stage_cnt[i];
    }
    return RTPS_OK;
}

int rtps_sph_reset_timers(rtps_sph* s) {
    if (!s) return set_error(RTPS_ERR_ARG, "null handle");
    for (int i = 0; i < ST_N; ++i) {
        s->stage_ms[i] = 0;
        s->stage_cnt[i] = 0;
    }
    return RTPS_OK;
}

int rtps_sph_count_pairs(rtps_sph* s, long long* K, long long* M, long long* K_tested) {
    if (!s || !K || !M) return set_error(RTPS_ERR_ARG, "rtps_sph_count_pairs: null argument");
    unsigned long long* d = nullptr;
    RTPS_TRY_HIP(hipMalloc((void**)&d, 3 * sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(d, 0, 3 * sizeof(unsigned long long), s->stream);
    if (e == hipSuccess)
        e = launch_count_pairs(s->num_host, s->d_num, s->pos_s, s->start, s->end, s->sp, s->gs, win(s), d, s->stream);
    unsigned long long h[3] = {0, 0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_error(e, "rtps_sph_count_pairs");
    *K = (long long)h[0];
    *M = (long long)h[1];
    if (K_tested) *K_tested = (long long)h[2];
    return RTPS_OK;
}

int rtps_sph_stream(rtps_sph* s, void** stream) {
    if (!s || !stream) return set_error(RTPS_ERR_ARG, "null argument");
    *stream = (void*)s->stream;
    return RTPS_OK;
}

}  // extern "C"
