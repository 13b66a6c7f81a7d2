// rtps_badj: forward and adjoint of many independent scenes at once (no
// reference counterpart).  Every scene runs the step of the single-system
// adjoint handle (adjoint_system.cpp) with its own settings, viscosity and
// gas constant; each stage is one launch for the whole batch (segmented keys
// (scene << 24) | (cell << 4) | xq as in the batched system), and the
// backward pass gives dL/dv0, dL/dx0 per particle and dL/dmu_s, dL/dK_s per
// scene.  Scene s holds exactly the max_num_particles of its settings, in
// slots [off_s, off_s + n_s) of the concatenated arrays.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "batch_adjoint.h"
#include "batch_layout.h"
#include "batch_types.h"
#include "rtps_internal.h"
#include "sph_device.h"

using namespace rtps;

namespace {

struct BRecord {   // one forward step's sorted state of the whole batch
    uint32_t *keys = nullptr, *idx = nullptr, *start = nullptr, *end = nullptr;
    float4 *pos_s = nullptr, *vel_s = nullptr, *ve_s = nullptr, *force_s = nullptr;
};

}  // namespace

struct rtps_badj {
    BatchLayout L;
    int S = 0, T = 0, C = 0, ck = 20, steps = 0, n_items = 0, persistent_grid = 1, force_grid = 1;
    std::vector<void*> allocs;
    SceneDev* scenes = nullptr;
    int *d_num = nullptr, *queues = nullptr, *item_first = nullptr, *left_grid = nullptr;
    uint16_t* sid = nullptr;
    BatchItem* items = nullptr;
    float4 *pos[2] = {}, *vel[2] = {}, *ve[2] = {};
    std::vector<float4*> ck_pos, ck_vel, ck_ve;
    std::vector<BRecord> rec;
    uint32_t *keys = nullptr, *idx = nullptr;
    void* sort_tmp = nullptr;
    float4* xsph = nullptr;
    float4 *lam_x[2] = {}, *lam_v[2] = {}, *lam_e[2] = {};
    float4 *ga = nullptr, *gX = nullptr, *lxp = nullptr, *gx2 = nullptr, *ge = nullptr;
    float* grho = nullptr;
    float2* partial = nullptr;
    double* acc = nullptr;
    bool have_forward = false;

    ~rtps_badj() {
        (void)hipDeviceSynchronize();
        for (void* p : allocs) (void)hipFree(p);
    }
    template <typename T>
    hipError_t alloc(T** p, size_t count) {
        hipError_t e = hipMalloc((void**)p, count * sizeof(T) + 16);
        if (e == hipSuccess) allocs.push_back(*p);
        return e;
    }
};

namespace {

#define RTPS_TRY_HIP(expr)                                 \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_error(e_, #expr); \
    } while (0)

// One forward step of every scene from state slot `in` to `out`, recording into r.
int forward_step(rtps_badj* a, int in, int out, BRecord& r, hipStream_t st) {
    const int T = a->T;
    RTPS_TRY_HIP(launch_hash_b(T, a->sid, a->scenes, a->d_num, a->pos[in], a->keys, a->idx, r.start, a->C, st));
    uint32_t *ks, *is;
    RTPS_TRY_HIP(sort_pairs_passes(T, a->L.shifts, a->L.widths, a->L.npasses, a->keys, a->idx, a->sort_tmp, st, &ks,
                                   &is));
    RTPS_TRY_HIP(hipMemcpyAsync(r.keys, ks, (size_t)T * 4, hipMemcpyDeviceToDevice, st));
    RTPS_TRY_HIP(hipMemcpyAsync(r.idx, is, (size_t)T * 4, hipMemcpyDeviceToDevice, st));
    RTPS_TRY_HIP(launch_cellindices_b(T, r.keys, a->scenes, r.start, r.end, st));
    RTPS_TRY_HIP(launch_permute(T, nullptr, r.idx, a->pos[in], a->vel[in], a->ve[in], r.pos_s, r.vel_s, r.ve_s, st));
    RTPS_TRY_HIP(launch_pairs_b(kPairsDensity, pairs_default_cfg(kPairsDensity), a->persistent_grid, a->items,
                                a->n_items, a->queues, a->scenes, a->d_num, r.pos_s, r.start, r.end, r.keys, r.ve_s,
                                r.vel_s, r.idx, nullptr, nullptr, nullptr, st));
    RTPS_TRY_HIP(launch_pairs_b(kPairsForceStore, pairs_default_cfg(kPairsForceStore), a->force_grid, a->items,
                                a->n_items, a->queues, a->scenes, a->d_num, r.pos_s, r.start, r.end,
                                r.keys, r.ve_s, r.vel_s, r.idx, nullptr, nullptr, nullptr, st, r.force_s, a->xsph));
    RTPS_TRY_HIP(launch_badj_integrate(a->n_items, a->items, a->scenes, r.pos_s, r.vel_s, r.force_s, a->xsph, r.idx,
                                       r.start, a->pos[out], a->vel[out], a->ve[out], a->left_grid, st));
    return RTPS_OK;
}

}  // namespace

extern "C" {

int rtps_badj_create(int num_scenes, const int* n_settings, const char* const* keys, const char* const* vals,
                     int checkpoint_every, int max_steps, rtps_badj** out) {
    try {
        if (!out || checkpoint_every <= 0 || max_steps <= 0)
            return set_error(RTPS_ERR_ARG, "rtps_badj_create: bad arguments");
        *out = nullptr;
        std::unique_ptr<rtps_badj> a(new rtps_badj());
        int rc = build_batch_layout(num_scenes, n_settings, keys, vals, a->L);
        if (rc) return rc;
        for (int s = 0; s < num_scenes; ++s) {
            const SceneDev& d = a->L.scenes[s];
            if (d.sp.simulation_scale != 1.0f)
                return set_error(RTPS_ERR_ARG, "scene " + std::to_string(s) + ": the adjoint supports "
                                                                               "simulation_scale = 1 only");
            for (float x : a->L.settings[s].GetSettingAs<std::vector<float>>("gravity_alphas"))
                if (x != 0.f)
                    return set_error(RTPS_ERR_ARG, "scene " + std::to_string(s) +
                                                       ": the adjoint does not support point gravity sources");
        }
        a->S = num_scenes;
        a->T = (int)a->L.T;
        a->C = (int)a->L.total_cells;
        a->ck = checkpoint_every;
        a->n_items = (int)a->L.items.size();
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
                cus = prop.multiProcessorCount;
        }
        a->persistent_grid =
            std::max(1, std::min(a->n_items, cus * pairs_blocks_per_cu(pairs_default_cfg(kPairsDensity))));
        a->force_grid = std::max(1, std::min(a->n_items, cus * pairs_blocks_per_cu(pairs_default_cfg(kPairsForceStore))));
        const size_t N = (size_t)a->T, C = (size_t)a->C;
        RTPS_TRY_HIP(a->alloc(&a->scenes, (size_t)num_scenes));
        RTPS_TRY_HIP(a->alloc(&a->d_num, (size_t)num_scenes));
        RTPS_TRY_HIP(a->alloc(&a->queues, (size_t)kTileQueueWords));
        RTPS_TRY_HIP(hipMemset(a->queues, 0, kTileQueueWords * sizeof(int)));   // each launch's last block re-clears
        RTPS_TRY_HIP(a->alloc(&a->item_first, (size_t)num_scenes + 1));
        RTPS_TRY_HIP(a->alloc(&a->left_grid, 1));
        RTPS_TRY_HIP(a->alloc(&a->sid, N));
        RTPS_TRY_HIP(a->alloc(&a->items, (size_t)a->n_items));
        for (int s = 0; s < 2; ++s) {
            RTPS_TRY_HIP(a->alloc(&a->pos[s], N));
            RTPS_TRY_HIP(a->alloc(&a->vel[s], N));
            RTPS_TRY_HIP(a->alloc(&a->ve[s], N));
            RTPS_TRY_HIP(a->alloc(&a->lam_x[s], N));
            RTPS_TRY_HIP(a->alloc(&a->lam_v[s], N));
            RTPS_TRY_HIP(a->alloc(&a->lam_e[s], N));
        }
        const int nck = (max_steps + checkpoint_every - 1) / checkpoint_every;
        a->ck_pos.resize(nck);
        a->ck_vel.resize(nck);
        a->ck_ve.resize(nck);
        for (int c = 0; c < nck; ++c) {
            RTPS_TRY_HIP(a->alloc(&a->ck_pos[c], N));
            RTPS_TRY_HIP(a->alloc(&a->ck_vel[c], N));
            RTPS_TRY_HIP(a->alloc(&a->ck_ve[c], N));
        }
        a->rec.resize(checkpoint_every);
        for (BRecord& r : a->rec) {
            RTPS_TRY_HIP(a->alloc(&r.keys, N));
            RTPS_TRY_HIP(a->alloc(&r.idx, N));
            RTPS_TRY_HIP(a->alloc(&r.start, C));
            RTPS_TRY_HIP(a->alloc(&r.end, C));
            RTPS_TRY_HIP(a->alloc(&r.pos_s, N));
            RTPS_TRY_HIP(a->alloc(&r.vel_s, N));
            RTPS_TRY_HIP(a->alloc(&r.ve_s, N));
            RTPS_TRY_HIP(a->alloc(&r.force_s, N));
        }
        RTPS_TRY_HIP(a->alloc(&a->keys, N));
        RTPS_TRY_HIP(a->alloc(&a->idx, N));
        RTPS_TRY_HIP(a->alloc((char**)&a->sort_tmp, sort_temp_bytes((int)N, 32)));
        RTPS_TRY_HIP(a->alloc(&a->xsph, N));
        RTPS_TRY_HIP(a->alloc(&a->ga, N));
        RTPS_TRY_HIP(a->alloc(&a->gX, N));
        RTPS_TRY_HIP(a->alloc(&a->lxp, N));
        RTPS_TRY_HIP(a->alloc(&a->gx2, N));
        RTPS_TRY_HIP(a->alloc(&a->ge, N));
        RTPS_TRY_HIP(a->alloc(&a->grho, N));
        RTPS_TRY_HIP(a->alloc(&a->partial, (size_t)a->n_items));
        RTPS_TRY_HIP(a->alloc(&a->acc, 2 * (size_t)num_scenes));
        std::vector<uint16_t> sid(N);
        std::vector<int> counts(num_scenes);
        for (int s = 0; s < num_scenes; ++s) {
            const SceneDev& d = a->L.scenes[s];
            std::fill(sid.begin() + d.off, sid.begin() + d.off + d.cap, (uint16_t)s);
            counts[s] = d.cap;   // every slot holds a particle
        }
        RTPS_TRY_HIP(hipMemcpy(a->sid, sid.data(), N * sizeof(uint16_t), hipMemcpyHostToDevice));
        RTPS_TRY_HIP(hipMemcpy(a->d_num, counts.data(), sizeof(int) * num_scenes, hipMemcpyHostToDevice));
        RTPS_TRY_HIP(hipMemcpy(a->items, a->L.items.data(), sizeof(BatchItem) * a->L.items.size(),
                               hipMemcpyHostToDevice));
        RTPS_TRY_HIP(hipMemcpy(a->item_first, a->L.item_first.data(), sizeof(int) * (num_scenes + 1),
                               hipMemcpyHostToDevice));
        RTPS_TRY_HIP(hipMemcpy(a->scenes, a->L.scenes.data(), sizeof(SceneDev) * num_scenes, hipMemcpyHostToDevice));
        RTPS_TRY_HIP(sort_clear_temp(a->sort_tmp, (int)N, nullptr));
        for (BRecord& r : a->rec) RTPS_TRY_HIP(launch_fill_u32(r.end, 0u, (long)C, nullptr));
        RTPS_TRY_HIP(hipDeviceSynchronize());
        *out = a.release();
        return RTPS_OK;
    } catch (const std::exception& e) {
        return set_error(RTPS_ERR_ARG, std::string("rtps_badj_create: ") + e.what());
    }
}

int rtps_badj_destroy(rtps_badj* a) {
    delete a;
    return RTPS_OK;
}

int rtps_badj_layout(rtps_badj* a, int* num_scenes, int* total, int* offsets) {
    if (!a) return set_error(RTPS_ERR_ARG, "null handle");
    if (num_scenes) *num_scenes = a->S;
    if (total) *total = a->T;
    if (offsets) {
        for (int s = 0; s < a->S; ++s) offsets[s] = a->L.scenes[s].off;
        offsets[a->S] = a->T;
    }
    return RTPS_OK;
}

int rtps_badj_forward(rtps_badj* a, const rtps_float4* pos0, const rtps_float4* vel0, int nsteps,
                      const float* viscosity, const float* gas_constant, int collision, rtps_float4* pos_T,
                      rtps_float4* vel_T, void* stream) {
    if (!a || !pos0 || !vel0 || !pos_T || !viscosity || !gas_constant || nsteps <= 0 ||
        (size_t)((nsteps + a->ck - 1) / a->ck) > a->ck_pos.size())
        return set_error(RTPS_ERR_ARG, "rtps_badj_forward: bad arguments or more steps than the handle holds");
    hipStream_t st = (hipStream_t)stream;
    for (int s = 0; s < a->S; ++s) {
        SceneDev& d = a->L.scenes[s];
        d.sp.viscosity = viscosity[s];
        d.sp.K = gas_constant[s];
        d.collision = collision ? 1 : 0;
    }
    RTPS_TRY_HIP(hipMemcpyAsync(a->scenes, a->L.scenes.data(), sizeof(SceneDev) * a->S, hipMemcpyHostToDevice, st));
    a->steps = nsteps;
    const size_t nb = (size_t)a->T * 16;
    RTPS_TRY_HIP(hipMemcpyAsync(a->pos[0], pos0, nb, hipMemcpyDeviceToDevice, st));
    RTPS_TRY_HIP(hipMemcpyAsync(a->vel[0], vel0, nb, hipMemcpyDeviceToDevice, st));
    RTPS_TRY_HIP(hipMemsetAsync(a->ve[0], 0, nb, st));   // veleval starts at zero
    RTPS_TRY_HIP(hipMemsetAsync(a->left_grid, 0, sizeof(int), st));
    int cur = 0;
    for (int n = 0; n < nsteps; ++n) {
        if (n % a->ck == 0) {
            const int c = n / a->ck;
            RTPS_TRY_HIP(hipMemcpyAsync(a->ck_pos[c], a->pos[cur], nb, hipMemcpyDeviceToDevice, st));
            RTPS_TRY_HIP(hipMemcpyAsync(a->ck_vel[c], a->vel[cur], nb, hipMemcpyDeviceToDevice, st));
            RTPS_TRY_HIP(hipMemcpyAsync(a->ck_ve[c], a->ve[cur], nb, hipMemcpyDeviceToDevice, st));
        }
        int rc = forward_step(a, cur, 1 - cur, a->rec[0], st);
        if (rc) return rc;
        cur = 1 - cur;
    }
    RTPS_TRY_HIP(hipMemcpyAsync(pos_T, a->pos[cur], nb, hipMemcpyDeviceToDevice, st));
    if (vel_T) RTPS_TRY_HIP(hipMemcpyAsync(vel_T, a->vel[cur], nb, hipMemcpyDeviceToDevice, st));
    int left = 0;
    RTPS_TRY_HIP(hipMemcpyAsync(&left, a->left_grid, sizeof(int), hipMemcpyDeviceToHost, st));
    RTPS_TRY_HIP(hipStreamSynchronize(st));
    a->have_forward = true;
    if (left) return set_error(RTPS_ERR_STATE, "particles left the grid during the forward run (no deletion path in "
                                               "the adjoint)");
    return RTPS_OK;
}

int rtps_badj_backward(rtps_badj* a, const rtps_float4* dL_dpos_T, const rtps_float4* dL_dvel_T,
                       rtps_float4* dL_dvel0, rtps_float4* dL_dpos0, double* dL_dvisc, double* dL_dgas,
                       void* stream) {
    if (!a || !dL_dpos_T || !dL_dvel0) return set_error(RTPS_ERR_ARG, "rtps_badj_backward: bad arguments");
    if (!a->have_forward) return set_error(RTPS_ERR_STATE, "rtps_badj_backward needs a forward run");
    hipStream_t st = (hipStream_t)stream;
    const size_t nb = (size_t)a->T * 16;
    int L = 0;
    RTPS_TRY_HIP(hipMemcpyAsync(a->lam_x[0], dL_dpos_T, nb, hipMemcpyDeviceToDevice, st));
    if (dL_dvel_T) RTPS_TRY_HIP(hipMemcpyAsync(a->lam_v[0], dL_dvel_T, nb, hipMemcpyDeviceToDevice, st));
    else RTPS_TRY_HIP(hipMemsetAsync(a->lam_v[0], 0, nb, st));
    RTPS_TRY_HIP(hipMemsetAsync(a->lam_e[0], 0, nb, st));
    RTPS_TRY_HIP(hipMemsetAsync(a->acc, 0, 2 * sizeof(double) * a->S, st));
    const int nck = (a->steps + a->ck - 1) / a->ck;
    for (int c = nck - 1; c >= 0; --c) {
        const int first = c * a->ck, last = std::min(a->steps, first + a->ck);
        RTPS_TRY_HIP(hipMemcpyAsync(a->pos[0], a->ck_pos[c], nb, hipMemcpyDeviceToDevice, st));
        RTPS_TRY_HIP(hipMemcpyAsync(a->vel[0], a->ck_vel[c], nb, hipMemcpyDeviceToDevice, st));
        RTPS_TRY_HIP(hipMemcpyAsync(a->ve[0], a->ck_ve[c], nb, hipMemcpyDeviceToDevice, st));
        int cur = 0;
        for (int s = first; s < last; ++s) {
            int rc = forward_step(a, cur, 1 - cur, a->rec[s - first], st);
            if (rc) return rc;
            cur = 1 - cur;
        }
        for (int s = last - 1; s >= first; --s) {
            const BRecord& r = a->rec[s - first];
            RTPS_TRY_HIP(launch_badj_b1(a->n_items, a->items, a->scenes, r.idx, r.pos_s, r.vel_s, r.force_s,
                                        a->lam_x[L], a->lam_v[L], a->lam_e[L], a->ga, a->gX, a->lam_v[1 - L], a->lxp,
                                        st));
            RTPS_TRY_HIP(launch_badj_b2(a->n_items, a->items, a->scenes, r.pos_s, r.ve_s, a->ga, a->gX, r.start,
                                        r.end, r.keys, a->gx2, a->ge, a->grho, a->partial, st));
            RTPS_TRY_HIP(launch_badj_reduce(a->S, a->item_first, a->partial, a->acc, st));
            RTPS_TRY_HIP(launch_badj_b3(a->n_items, a->items, a->scenes, r.idx, r.pos_s, a->grho, a->lxp, a->gx2,
                                        a->ge, r.start, r.end, r.keys, a->lam_x[1 - L], a->lam_e[1 - L], st));
            L = 1 - L;
        }
    }
    RTPS_TRY_HIP(hipMemcpyAsync(dL_dvel0, a->lam_v[L], nb, hipMemcpyDeviceToDevice, st));
    if (dL_dpos0) RTPS_TRY_HIP(hipMemcpyAsync(dL_dpos0, a->lam_x[L], nb, hipMemcpyDeviceToDevice, st));
    std::vector<double> acc(2 * (size_t)a->S);
    RTPS_TRY_HIP(hipMemcpyAsync(acc.data(), a->acc, sizeof(double) * acc.size(), hipMemcpyDeviceToHost, st));
    RTPS_TRY_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < a->S; ++s) {
        if (dL_dvisc) dL_dvisc[s] = acc[2 * s];
        if (dL_dgas) dL_dgas[s] = acc[2 * s + 1];
    }
    return RTPS_OK;
}

}  // extern "C"
