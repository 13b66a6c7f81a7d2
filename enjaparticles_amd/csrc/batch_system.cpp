// Batched scenes handle (rtps_batch): S independent SPH systems, each with
// its own settings (viscosity, gas constant, box, capacity ...), stepped with
// one launch per stage and one hipGraph per step.  No reference counterpart:
// every scene computes what a single rtps_sph system with the same settings
// and particles computes.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "batch_layout.h"
#include "batch_types.h"
#include "rtps_internal.h"
#include "sph_device.h"

using namespace rtps;

enum BStage { B_HASH = 0, B_SORT, B_CELL, B_PERMUTE, B_DENSITY, B_FORCE, B_N };

struct rtps_batch {
    int S = 0, T = 0, total_cells = 0;
    std::vector<RTPSSettings> settings;
    std::vector<SceneDev> scenes_h;
    std::vector<int> num_h;
    hipStream_t stream = nullptr;
    SceneDev* scenes = nullptr;
    int* d_num = nullptr;
    uint16_t* sid = nullptr;
    float4 *pos_u = nullptr, *vel_u = nullptr, *veleval_u = nullptr, *pos_s = nullptr, *vel_s = nullptr,
           *ve_s = nullptr;
    uint32_t *keys = nullptr, *idx = nullptr, *start = nullptr, *end = nullptr;
    uint32_t *keys_sorted = nullptr, *idx_sorted = nullptr;
    // work items of 256, 128 and 64 slots (the tile sizes of the pair kernels)
    BatchItem* items = nullptr;
    BatchItem* items128 = nullptr;
    BatchItem* items64 = nullptr;
    int n_items = 0, n_items128 = 0, n_items64 = 0, cus = 256;
    // launch configurations of the density and force pair kernels (the autotuner's choice)
    PairsCfg cfg_density = pairs_default_cfg(kPairsDensity), cfg_force = pairs_default_cfg(kPairsForceStep);
    int opt_autotune = 1;
    bool tuned = false;
    int* queues = nullptr;
    void* sort_tmp = nullptr;
    uint32_t* sub = nullptr;   // sub-cell starts of the windowed rows, 16 per cell
    // neighbour rows the density kernel records for the force kernel: 9 rows
    // x 3 words per slot (used while opt_rows is set)
    uint32_t* rows = nullptr;
    int opt_rows = 1;
    int shifts[4] = {}, widths[4] = {}, npasses = 0;
    int opt_graph = 1, opt_timers = 0;
    hipGraphExec_t gexec = nullptr;
    bool graph_valid = false;
    hipEvent_t ev[B_N + 1] = {};
    double stage_ms[B_N] = {};
    long stage_cnt[B_N] = {};

    ~rtps_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (gexec) (void)hipGraphExecDestroy(gexec);
        void* bufs[] = {scenes, d_num, sid, pos_u, vel_u, veleval_u, pos_s, vel_s, ve_s, keys, idx,
                        start, end, items, items128, items64, queues, sort_tmp, sub, rows};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

#define RTPS_TRY_HIP(expr)                                 \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_error(e_, #expr); \
    } while (0)

int bits_for(unsigned v) {   // number of bits to represent values < v
    int b = 0;
    while ((1u << b) < v && b < 32) ++b;
    return b;
}

void mark(rtps_batch* b, int i) {
    if (b->opt_timers) (void)hipEventRecord(b->ev[i], b->stream);
}

// Work items and persistent grid of a pair-kernel configuration.
const BatchItem* items_of(const rtps_batch* b, const PairsCfg& c) {
    return c.block == 64 ? b->items64 : (c.block == 128 ? b->items128 : b->items);
}
int nitems_of(const rtps_batch* b, const PairsCfg& c) {
    return c.block == 64 ? b->n_items64 : (c.block == 128 ? b->n_items128 : b->n_items);
}
int grid_of(const rtps_batch* b, const PairsCfg& c) {
    return std::max(1, std::min(nitems_of(b, c), b->cus * pairs_blocks_per_cu(c)));
}

// The recorded rows of a step (both launches of a step take the same choice:
// density records them exactly when the force launch reads them).
uint32_t* rows_of(const rtps_batch* b) { return b->opt_rows ? b->rows : nullptr; }

int launch_density_b(rtps_batch* b, const PairsCfg& c) {
    RTPS_TRY_HIP(launch_pairs_b(kPairsDensity, c, grid_of(b, c), items_of(b, c), nitems_of(b, c), b->queues,
                                b->scenes, b->d_num, b->pos_s, b->start, b->end, b->keys_sorted, b->ve_s, b->vel_s,
                                b->idx_sorted, b->pos_u, b->vel_u, b->veleval_u, b->stream, nullptr, nullptr, b->sub,
                                rows_of(b), (uint32_t)b->T));
    return RTPS_OK;
}

int launch_force_b(rtps_batch* b, const PairsCfg& c) {
    RTPS_TRY_HIP(launch_pairs_b(kPairsForceStep, c, grid_of(b, c), items_of(b, c), nitems_of(b, c),
                                b->queues, b->scenes, b->d_num, b->pos_s, b->start, b->end,
                                b->keys_sorted, b->ve_s, b->vel_s, b->idx_sorted, b->pos_u, b->vel_u, b->veleval_u,
                                b->stream, nullptr, nullptr, b->sub, rows_of(b), (uint32_t)b->T));
    return RTPS_OK;
}

int launch_front_b(rtps_batch* b) {
    mark(b, B_HASH);
    RTPS_TRY_HIP(launch_hash_b(b->T, b->sid, b->scenes, b->d_num, b->pos_u, b->keys, b->idx, b->start,
                               b->total_cells, b->stream));
    mark(b, B_SORT);
    RTPS_TRY_HIP(sort_pairs_passes(b->T, b->shifts, b->widths, b->npasses, b->keys, b->idx, b->sort_tmp, b->stream,
                                   &b->keys_sorted, &b->idx_sorted));
    mark(b, B_CELL);
    RTPS_TRY_HIP(launch_cellindices_b(b->T, b->keys_sorted, b->scenes, b->start, b->end, b->stream, b->sub));
    mark(b, B_PERMUTE);
    RTPS_TRY_HIP(launch_permute(b->T, nullptr, b->idx_sorted, b->pos_u, b->vel_u, b->veleval_u, b->pos_s, b->vel_s,
                                b->ve_s, b->stream));
    return RTPS_OK;
}

int launch_step(rtps_batch* b) {
    int rc = launch_front_b(b);
    if (rc) return rc;
    mark(b, B_DENSITY);
    rc = launch_density_b(b, b->cfg_density);
    if (rc) return rc;
    mark(b, B_FORCE);
    rc = launch_force_b(b, b->cfg_force);
    if (rc) return rc;
    RTPS_TRY_HIP(launch_commit_b(b->S, b->scenes, b->d_num, b->start, b->stream));
    mark(b, B_N);
    return RTPS_OK;
}

// ---------------------------------------------------------------------------
// Autotuner of the pair kernels (no reference counterpart): threads per block
// (= slots per work item, 64, 128 or 256) and register budget (4, 5 or 6 waves
// per SIMD for density, 3 or 4 for force) of the density and the force kernel, timed on the live sorted state
// of the first step; each kernel keeps its fastest configuration (a
// configuration tried later must be 3% faster).  Both kernels are idempotent
// on one sorted state (density rewrites veleval_s.w from the positions; the
// force step writes the unsorted arrays from the sorted ones), so tuning is
// an ordinary step.  Choices are cached per process by device and layout size
// class, and in the file named by RTPS_TUNE_CACHE when it is set.
std::mutex g_btune_mu;
std::map<std::string, std::pair<PairsCfg, PairsCfg>> g_btune;
bool g_btune_file_loaded = false;

std::string btune_key(const rtps_batch* b) {
    int dev = 0;
    hipDeviceProp_t prop{};
    std::string arch = "unknown";
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        arch = prop.gcnArchName;
        for (char& ch : arch)
            if (ch == ' ' || ch == ':') ch = '_';
    }
    unsigned long t2 = 1, s2 = 1;
    while (t2 < (unsigned long)b->T) t2 <<= 1;
    while (s2 < (unsigned long)b->S) s2 <<= 1;
    return "batch/" + arch + "/cu" + std::to_string(b->cus) + "/t" + std::to_string(t2) + "/s" + std::to_string(s2);
}

void btune_load_locked() {
    if (g_btune_file_loaded) return;
    g_btune_file_loaded = true;
    const char* f = std::getenv("RTPS_TUNE_CACHE");
    if (!f || !*f) return;
    std::ifstream in(f);
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream is(line);
        std::string k;
        PairsCfg d, fo;
        if (is >> k >> d.block >> d.waves >> fo.block >> fo.waves && k.rfind("batch/", 0) == 0 &&
            pairs_cfg_valid(kPairsDensity, d) && pairs_cfg_valid(kPairsForceStep, fo))
            g_btune[k] = {d, fo};
    }
}

bool btune_get(const std::string& k, PairsCfg& d, PairsCfg& f) {
    std::lock_guard<std::mutex> lk(g_btune_mu);
    btune_load_locked();
    auto it = g_btune.find(k);
    if (it == g_btune.end()) return false;
    d = it->second.first;
    f = it->second.second;
    return true;
}

void btune_put(const std::string& k, const PairsCfg& d, const PairsCfg& f) {
    std::lock_guard<std::mutex> lk(g_btune_mu);
    btune_load_locked();
    g_btune[k] = {d, f};
    const char* file = std::getenv("RTPS_TUNE_CACHE");
    if (file && *file) {
        std::ofstream out(file, std::ios::app);
        out << k << ' ' << d.block << ' ' << d.waves << ' ' << f.block << ' ' << f.waves << '\n';
    }
}

// One step with the pair-kernel configurations tuned on the way (measure =
// false: take a cached choice when there is one).
int autotune_step_b(rtps_batch* b, bool measure) {
    const std::string key = btune_key(b);
    PairsCfg d, f;
    if (!measure && btune_get(key, d, f)) {
        b->cfg_density = d;
        b->cfg_force = f;
        b->tuned = true;
        b->graph_valid = false;
        return launch_step(b);
    }
    const int timers = b->opt_timers;
    b->opt_timers = 0;
    int rc = launch_front_b(b);
    b->opt_timers = timers;
    if (rc) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    RTPS_TRY_HIP(hipEventCreateWithFlags(&e0, kTimingEvent));
    hipError_t err = hipEventCreateWithFlags(&e1, kTimingEvent);
    if (err != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_error(err, "hipEventCreate");
    }
    const std::vector<PairsCfg> cands[2] = {
        {{256, 5}, {256, 4}, {256, 6}, {128, 5}, {128, 4}, {128, 6}, {64, 5}, {64, 4}, {64, 6}},
        {{256, 4}, {256, 3}, {128, 4}, {128, 3}, {64, 4}, {64, 3}, {256, 5}, {128, 5}, {64, 5}}};
    constexpr int kReps = 3;
    constexpr float kMargin = 0.97f;
    const char* verbose_env = std::getenv("RTPS_TUNE_VERBOSE");
    const bool verbose = verbose_env && verbose_env[0] == '1';
    // ms per launch of kReps launches after one untimed one (each launch's
    // last block clears the queue counters for the next); < 0 on error
    auto time_it = [&](int which, const PairsCfg& c) -> float {
        float total = 0.f;
        for (int r = 0; r <= kReps; ++r) {
            if (hipEventRecord(e0, b->stream) != hipSuccess) return -1.f;
            if ((which == 0 ? launch_density_b(b, c) : launch_force_b(b, c)) != RTPS_OK) return -1.f;
            if (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return -1.f;
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return -1.f;
            if (r > 0) total += ms;
        }
        return total / kReps;
    };
    PairsCfg best[2];
    bool failed = false;
    for (int which = 0; which < 2 && !failed; ++which) {
        float best_ms = 1e30f;
        for (const PairsCfg& c : cands[which]) {
            const float ms = time_it(which, c);
            if (verbose)
                std::fprintf(stderr, "rtps batch autotune: %s block %d waves %d: %.4f ms\n",
                             which == 0 ? "density" : "force", c.block, c.waves, ms);
            if (ms < 0.f) { failed = true; break; }
            if (best_ms >= 1e30f || ms < kMargin * best_ms) {
                best_ms = ms;
                best[which] = c;
            }
        }
        if (which == 0 && !failed) {   // the force kernel reads the densities of the chosen density launch
            if (launch_density_b(b, best[0]) != RTPS_OK) failed = true;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (failed) return hip_error(hipGetLastError(), "batch autotune launch");
    b->cfg_density = best[0];
    b->cfg_force = best[1];
    b->tuned = true;
    b->graph_valid = false;
    btune_put(key, best[0], best[1]);
    // finish the step: the force step with the chosen configuration, then the count commit
    rc = launch_force_b(b, b->cfg_force);
    if (rc) return rc;
    RTPS_TRY_HIP(launch_commit_b(b->S, b->scenes, b->d_num, b->start, b->stream));
    return RTPS_OK;
}

int capture(rtps_batch* b) {
    if (b->gexec) {
        (void)hipGraphExecDestroy(b->gexec);
        b->gexec = nullptr;
    }
    hipGraph_t graph = nullptr;
    RTPS_TRY_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int timers = b->opt_timers;
    b->opt_timers = 0;
    int rc = launch_step(b);
    b->opt_timers = timers;
    hipError_t e = hipStreamEndCapture(b->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_error(e, "hipGraphInstantiate");
    b->graph_valid = true;
    return RTPS_OK;
}

}  // namespace

namespace rtps {

int build_batch_layout(int num_scenes, const int* n_settings, const char* const* keys, const char* const* vals,
                       BatchLayout& L) {
    if (num_scenes <= 0 || num_scenes > kMaxScenes || !n_settings)
        return set_error(RTPS_ERR_ARG, "1..256 scenes and their settings are required");
    L.settings.assign(num_scenes, RTPSSettings());
    L.scenes.assign(num_scenes, SceneDev());
    long k = 0, T = 0, cells = 0;
    int max_cells = 0;
    for (int s = 0; s < num_scenes; ++s) {
        if (n_settings[s] < 0) return set_error(RTPS_ERR_ARG, "negative settings count");
        RTPSSettings& st = L.settings[s];
        for (int q = 0; q < n_settings[s]; ++q, ++k) {
            if (!keys[k] || !vals[k]) return set_error(RTPS_ERR_ARG, "null key or value");
            st.SetSetting<std::string>(keys[k], vals[k]);
        }
        if (st.GetSettingAs<std::string>("system") != "sph")
            return set_error(RTPS_ERR_UNKNOWN_SYSTEM, "scene " + std::to_string(s) + ": system is not sph");
        const int cap = st.GetSettingAs<int>("max_num_particles");
        if (cap <= 0) return set_error(RTPS_ERR_ARG, "scene " + std::to_string(s) + ": max_num_particles <= 0");
        if (integrator_of(st) != 0)
            return set_error(RTPS_ERR_ARG, "scene " + std::to_string(s) + ": integrator " +
                                               st.GetSettingAs<std::string>("integrator") +
                                               " (batched scenes integrate with leapfrog)");
        SceneDev& d = L.scenes[s];
        std::memset(&d, 0, sizeof(d));
        rtps_grid_params gp, gps;
        derive_params(st, d.sp, gp, gps);
        if (gp.nb_cells <= 0 || gp.nb_cells >= (1 << 19))
            return set_error(RTPS_ERR_ARG, "scene " + std::to_string(s) + ": grid too large for segmented keys");
        d.g = make_grid(gp);
        d.gs = make_grid(gps);
        d.bmin = gps.bnd_min;
        d.bmax = gps.bnd_max;
        d.res_x = gps.grid_res.x;
        d.res_y = gps.grid_res.y;
        d.res_z = gps.grid_res.z;
        d.dt = st.GetSettingAs<float>("time_step");
        d.off = (int)T;
        d.cap = cap;
        d.cell_base = (int)cells;
        d.collision = 1;
        gravity_sources_of(st, d.gsrc);
        T += cap;
        cells += gp.nb_cells + 1;
        max_cells = std::max(max_cells, gp.nb_cells);
        if (T >= (1L << 31) - 1) return set_error(RTPS_ERR_ARG, "too many particle slots");
    }
    L.T = T;
    L.total_cells = cells;
    // sort plan: the low local bits (cell << 4 | xq; pads are all ones) then the scene bits
    const int low = sort_key_bits(max_cells) + kSubBits;
    int np = sort_plan(low, L.shifts, L.widths);
    if (np < 0 || np >= 4) return set_error(RTPS_ERR_ARG, "grid too large for segmented keys");
    const int sb = bits_for((unsigned)num_scenes);
    if (sb > 0) {
        L.shifts[np] = (int)kSceneShift;
        L.widths[np] = sb;
        ++np;
    }
    L.npasses = np;
    // work items: 256-slot tiles of every scene
    L.items.clear();
    L.items128.clear();
    L.items64.clear();
    L.item_first.assign(num_scenes + 1, 0);
    for (int s = 0; s < num_scenes; ++s) {
        const SceneDev& d = L.scenes[s];
        L.item_first[s] = (int)L.items.size();
        for (int f = 0; f < d.cap; f += 256) L.items.push_back(BatchItem{s, d.off + f, std::min(256, d.cap - f), 0});
        for (int f = 0; f < d.cap; f += 128) L.items128.push_back(BatchItem{s, d.off + f, std::min(128, d.cap - f), 0});
        for (int f = 0; f < d.cap; f += 64) L.items64.push_back(BatchItem{s, d.off + f, std::min(64, d.cap - f), 0});
    }
    L.item_first[num_scenes] = (int)L.items.size();
    return RTPS_OK;
}

}  // namespace rtps

extern "C" {

int rtps_batch_create(int num_scenes, const int* n_settings, const char* const* keys, const char* const* vals,
                      rtps_batch** out) {
    try {
        if (!out) return set_error(RTPS_ERR_ARG, "rtps_batch_create: null output");
        *out = nullptr;
        BatchLayout L;
        int rc = build_batch_layout(num_scenes, n_settings, keys, vals, L);
        if (rc) return rc;
        std::unique_ptr<rtps_batch> b(new rtps_batch());
        b->S = num_scenes;
        b->settings = L.settings;
        b->scenes_h = L.scenes;
        b->num_h.assign(num_scenes, 0);
        const long T = L.T, cells = L.total_cells;
        b->T = (int)T;
        b->total_cells = (int)cells;
        for (int q = 0; q < 4; ++q) {
            b->shifts[q] = L.shifts[q];
            b->widths[q] = L.widths[q];
        }
        b->npasses = L.npasses;
        const std::vector<BatchItem>& items = L.items;
        b->n_items = (int)items.size();
        b->n_items128 = (int)L.items128.size();
        b->n_items64 = (int)L.items64.size();
        int dev = 0;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
                b->cus = prop.multiProcessorCount;
        }

        std::vector<uint16_t> sid((size_t)T);
        for (int s = 0; s < num_scenes; ++s)
            std::fill(sid.begin() + b->scenes_h[s].off, sid.begin() + b->scenes_h[s].off + b->scenes_h[s].cap,
                      (uint16_t)s);

        RTPS_TRY_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        const size_t NT = (size_t)T;
        RTPS_TRY_HIP(hipMalloc((void**)&b->scenes, sizeof(SceneDev) * num_scenes));
        RTPS_TRY_HIP(hipMalloc((void**)&b->d_num, sizeof(int) * num_scenes));
        RTPS_TRY_HIP(hipMalloc((void**)&b->sid, sizeof(uint16_t) * NT));
        float4** f4[] = {&b->pos_u, &b->vel_u, &b->veleval_u, &b->pos_s, &b->vel_s, &b->ve_s};
        for (float4** p : f4) {
            RTPS_TRY_HIP(hipMalloc((void**)p, NT * sizeof(float4)));
            RTPS_TRY_HIP(hipMemsetAsync(*p, 0, NT * sizeof(float4), b->stream));
        }
        RTPS_TRY_HIP(hipMalloc((void**)&b->keys, NT * 4));
        RTPS_TRY_HIP(hipMalloc((void**)&b->idx, NT * 4));
        RTPS_TRY_HIP(hipMalloc((void**)&b->start, (size_t)cells * 4));
        RTPS_TRY_HIP(hipMalloc((void**)&b->end, (size_t)cells * 4));
        if ((size_t)cells * kSubCells * 4 <= ((size_t)2 << 30))   // at most 2 GiB; else the kernels search
            RTPS_TRY_HIP(hipMalloc((void**)&b->sub, (size_t)cells * kSubCells * 4));
        RTPS_TRY_HIP(hipMalloc((void**)&b->items, sizeof(BatchItem) * items.size()));
        RTPS_TRY_HIP(hipMalloc((void**)&b->items128, sizeof(BatchItem) * L.items128.size()));
        RTPS_TRY_HIP(hipMalloc((void**)&b->items64, sizeof(BatchItem) * L.items64.size()));
        RTPS_TRY_HIP(hipMalloc((void**)&b->queues, kTileQueueWords * sizeof(int)));
        RTPS_TRY_HIP(hipMalloc(&b->sort_tmp, sort_temp_bytes((int)T, 32)));
        RTPS_TRY_HIP(hipMalloc((void**)&b->rows, NT * 27 * 4));
        RTPS_TRY_HIP(hipMemcpyAsync(b->scenes, b->scenes_h.data(), sizeof(SceneDev) * num_scenes,
                                    hipMemcpyHostToDevice, b->stream));
        RTPS_TRY_HIP(hipMemsetAsync(b->d_num, 0, sizeof(int) * num_scenes, b->stream));
        RTPS_TRY_HIP(hipMemcpyAsync(b->sid, sid.data(), NT * sizeof(uint16_t), hipMemcpyHostToDevice, b->stream));
        RTPS_TRY_HIP(hipMemcpyAsync(b->items, items.data(), sizeof(BatchItem) * items.size(), hipMemcpyHostToDevice,
                                    b->stream));
        RTPS_TRY_HIP(hipMemcpyAsync(b->items128, L.items128.data(), sizeof(BatchItem) * L.items128.size(),
                                    hipMemcpyHostToDevice, b->stream));
        RTPS_TRY_HIP(hipMemcpyAsync(b->items64, L.items64.data(), sizeof(BatchItem) * L.items64.size(),
                                    hipMemcpyHostToDevice, b->stream));
        RTPS_TRY_HIP(hipMemsetAsync(b->queues, 0, kTileQueueWords * sizeof(int), b->stream));
        RTPS_TRY_HIP(launch_fill_u32(b->start, 666u, cells, b->stream));
        RTPS_TRY_HIP(launch_fill_u32(b->end, 666u, cells, b->stream));
        RTPS_TRY_HIP(sort_clear_temp(b->sort_tmp, (int)T, b->stream));
        b->keys_sorted = b->keys;
        b->idx_sorted = b->idx;
        for (auto& e : b->ev) RTPS_TRY_HIP(hipEventCreateWithFlags(&e, kTimingEvent));
        RTPS_TRY_HIP(hipStreamSynchronize(b->stream));
        *out = b.release();
        return RTPS_OK;
    } catch (const std::exception& e) {
        return set_error(RTPS_ERR_ARG, std::string("rtps_batch_create: ") + e.what());
    }
}

int rtps_batch_destroy(rtps_batch* b) {
    delete b;
    return RTPS_OK;
}

int rtps_batch_num_scenes(rtps_batch* b, int* num_scenes, int* total_slots) {
    if (!b) return set_error(RTPS_ERR_ARG, "null handle");
    if (num_scenes) *num_scenes = b->S;
    if (total_slots) *total_slots = b->T;
    return RTPS_OK;
}

int rtps_batch_get_num(rtps_batch* b, int scene, int* num) {
    if (!b || !num || scene < 0 || scene >= b->S) return set_error(RTPS_ERR_ARG, "rtps_batch_get_num: bad arguments");
    RTPS_TRY_HIP(hipMemcpyAsync(num, b->d_num + scene, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    RTPS_TRY_HIP(hipStreamSynchronize(b->stream));
    return RTPS_OK;
}

int rtps_batch_push_particles(rtps_batch* b, int scene, const rtps_float4* pos, const rtps_float4* vel, int n) {
    if (!b || scene < 0 || scene >= b->S || n < 0 || (n > 0 && !pos))
        return set_error(RTPS_ERR_ARG, "rtps_batch_push_particles: bad arguments");
    if (n == 0) return RTPS_OK;
    int num = 0;
    int rc = rtps_batch_get_num(b, scene, &num);
    if (rc) return rc;
    const SceneDev& d = b->scenes_h[scene];
    if ((long)num + n > d.cap)
        return set_error(RTPS_ERR_CAPACITY, "pushParticles: exceeded max nb (" + std::to_string(d.cap) +
                                                ") of particles allowed in scene " + std::to_string(scene));
    RTPS_TRY_HIP(hipMemcpyAsync(b->pos_u + d.off + num, pos, (size_t)n * 16, hipMemcpyHostToDevice, b->stream));
    if (vel)
        RTPS_TRY_HIP(hipMemcpyAsync(b->vel_u + d.off + num, vel, (size_t)n * 16, hipMemcpyHostToDevice, b->stream));
    else
        RTPS_TRY_HIP(hipMemsetAsync(b->vel_u + d.off + num, 0, (size_t)n * 16, b->stream));
    num += n;
    RTPS_TRY_HIP(hipMemcpyAsync(b->d_num + scene, &num, sizeof(int), hipMemcpyHostToDevice, b->stream));
    RTPS_TRY_HIP(hipStreamSynchronize(b->stream));
    b->num_h[scene] = num;
    return RTPS_OK;
}

int rtps_batch_step(rtps_batch* b, int nsteps) {
    if (!b || nsteps < 0) return set_error(RTPS_ERR_ARG, "rtps_batch_step: bad arguments");
    if (nsteps > 0 && b->opt_autotune && !b->tuned) {   // the first step tunes the pair kernels
        int rc = autotune_step_b(b, false);
        if (rc) return rc;
        --nsteps;
    }
    if (b->opt_timers) {
        for (int k = 0; k < nsteps; ++k) {
            int rc = launch_step(b);
            if (rc) return rc;
            RTPS_TRY_HIP(hipEventSynchronize(b->ev[B_N]));
            for (int i = 0; i < B_N; ++i) {
                float ms = 0.f;
                RTPS_TRY_HIP(hipEventElapsedTime(&ms, b->ev[i], b->ev[i + 1]));
                b->stage_ms[i] += ms;
                b->stage_cnt[i] += 1;
            }
        }
        return RTPS_OK;
    }
    if (b->opt_graph) {
        if (!b->graph_valid) {
            int rc = capture(b);
            if (rc) return rc;
        }
        for (int k = 0; k < nsteps; ++k) RTPS_TRY_HIP(hipGraphLaunch(b->gexec, b->stream));
        return RTPS_OK;
    }
    for (int k = 0; k < nsteps; ++k) {
        int rc = launch_step(b);
        if (rc) return rc;
    }
    return RTPS_OK;
}

int rtps_batch_synchronize(rtps_batch* b) {
    if (!b) return set_error(RTPS_ERR_ARG, "null handle");
    RTPS_TRY_HIP(hipStreamSynchronize(b->stream));
    return RTPS_OK;
}

int rtps_batch_read(rtps_batch* b, int scene, int field, void* host, int count) {
    if (!b || !host || scene < 0 || scene >= b->S || count < 0)
        return set_error(RTPS_ERR_ARG, "rtps_batch_read: bad arguments");
    const SceneDev& d = b->scenes_h[scene];
    if (count > d.cap) return set_error(RTPS_ERR_ARG, "rtps_batch_read: count > scene capacity");
    RTPS_TRY_HIP(hipStreamSynchronize(b->stream));
    const float4* src = nullptr;
    switch (field) {
        case RTPS_FIELD_POS: src = b->pos_u; break;
        case RTPS_FIELD_VEL: src = b->vel_u; break;
        case RTPS_FIELD_VELEVAL: src = b->veleval_u; break;
        case RTPS_FIELD_POS_S: src = b->pos_s; break;
        case RTPS_FIELD_VEL_S: src = b->vel_s; break;
        case RTPS_FIELD_VELEVAL_S: src = b->ve_s; break;
        default: break;
    }
    if (src) {
        RTPS_TRY_HIP(hipMemcpy(host, src + d.off, (size_t)count * 16, hipMemcpyDeviceToHost));
        return RTPS_OK;
    }
    if (field == RTPS_FIELD_DENSITY) {
        std::vector<float> tmp((size_t)count * 4);
        RTPS_TRY_HIP(hipMemcpy(tmp.data(), b->ve_s + d.off, (size_t)count * 16, hipMemcpyDeviceToHost));
        for (int i = 0; i < count; ++i) ((float*)host)[i] = tmp[4 * (size_t)i + 3];
        return RTPS_OK;
    }
    if (field == RTPS_FIELD_KEYS || field == RTPS_FIELD_IDX) {
        RTPS_TRY_HIP(hipMemcpy(host, (field == RTPS_FIELD_KEYS ? b->keys_sorted : b->idx_sorted) + d.off,
                               (size_t)count * 4, hipMemcpyDeviceToHost));
        uint32_t* k = (uint32_t*)host;
        if (field == RTPS_FIELD_KEYS)   // segmented reference keys: (scene << 24) | cell
            for (int i = 0; i < count; ++i) {
                const uint32_t local = k[i] & ((1u << kSceneShift) - 1u);
                k[i] = (k[i] & ~((1u << kSceneShift) - 1u)) | (local >> kSubBits);
            }
        else   // scene-local original indices
            for (int i = 0; i < count; ++i) k[i] -= (uint32_t)d.off;
        return RTPS_OK;
    }
    return set_error(RTPS_ERR_ARG, "rtps_batch_read: unsupported field");
}

int rtps_batch_set_option(rtps_batch* b, int option, int value) {
    if (!b) return set_error(RTPS_ERR_ARG, "null handle");
    switch (option) {
        case RTPS_OPT_GRAPH: b->opt_graph = value ? 1 : 0; break;
        case RTPS_OPT_TIMERS: b->opt_timers = value ? 1 : 0; break;
        case RTPS_OPT_AUTOTUNE: b->opt_autotune = value ? 1 : 0; break;
        case RTPS_OPT_SHARED_ROWS:
            if (value != 0 && value != 1) return set_error(RTPS_ERR_ARG, "shared rows option must be 0 or 1");
            b->opt_rows = value;
            break;
        case RTPS_OPT_BLOCK_DENSITY:
        case RTPS_OPT_BLOCK_FORCE: {   // value = block * 10 + waves, e.g. 2565 (256 threads, 5 waves per SIMD)
            const PairsCfg c{value / 10, value % 10};
            if (!pairs_cfg_valid(option == RTPS_OPT_BLOCK_DENSITY ? kPairsDensity : kPairsForceStep, c))
                return set_error(RTPS_ERR_ARG, "pair kernel launch: block*10 + waves with block 64, 128 or 256, waves 4..6 "
                                               "(density) or 3..4 (force)");
            (option == RTPS_OPT_BLOCK_DENSITY ? b->cfg_density : b->cfg_force) = c;
            b->tuned = true;   // fixed by the caller
            break;
        }
        case RTPS_OPT_COLLISION: {
            for (auto& d : b->scenes_h) d.collision = value ? 1 : 0;
            RTPS_TRY_HIP(hipMemcpyAsync(b->scenes, b->scenes_h.data(), sizeof(SceneDev) * b->S,
                                        hipMemcpyHostToDevice, b->stream));
            RTPS_TRY_HIP(hipStreamSynchronize(b->stream));
            break;
        }
        default: return set_error(RTPS_ERR_ARG, "unknown option");
    }
    b->graph_valid = false;
    return RTPS_OK;
}

int rtps_batch_stage_times(rtps_batch* b, float* avg_ms, int* counts, int n) {
    if (!b || n < 0) return set_error(RTPS_ERR_ARG, "bad arguments");
    for (int i = 0; i < n && i < B_N; ++i) {
        if (avg_ms) avg_ms[i] = b->stage_cnt[i] ? (float)(b->stage_ms[i] / b->stage_cnt[i]) : 0.f;
        if (counts) counts[i] = (int)b->stage_cnt[i];
    }
    return RTPS_OK;
}

int rtps_batch_reset_timers(rtps_batch* b) {
    if (!b) return set_error(RTPS_ERR_ARG, "null handle");
    for (int i = 0; i < B_N; ++i) {
        b->stage_ms[i] = 0;
        b->stage_cnt[i] = 0;
    }
    return RTPS_OK;
}

}  // extern "C"

extern "C" int rtps_batch_autotune(rtps_batch* b, int measure, int* block_density, int* waves_density,
                                   int* block_force, int* waves_force) {
    if (!b) return set_error(RTPS_ERR_ARG, "rtps_batch_autotune: null handle");
    int rc = autotune_step_b(b, measure != 0);
    if (rc) return rc;
    if (block_density) *block_density = b->cfg_density.block;
    if (waves_density) *waves_density = b->cfg_density.waves;
    if (block_force) *block_force = b->cfg_force.block;
    if (waves_force) *waves_force = b->cfg_force.waves;
    return RTPS_OK;
}
