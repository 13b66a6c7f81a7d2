// rtps_adj: forward SPH simulation with checkpoints and its hand-written
// adjoint (no reference counterpart).
//
// forward(pos0, vel0, T steps, mu, K) runs the step of the SPH system
// (hash, sort, cellindices, permute, density, force, walls + leapfrog) from
// the given state (veleval starts at zero, SPH.cpp:491-492) and keeps the
// state every `checkpoint_every` steps.  backward(dL/dx_T, dL/dv_T) walks the
// segments from the last: it replays a segment from its checkpoint, keeping
// every step's sorted records (keys, idx, cells, sorted pos/vel/veleval +
// density, SPH force), then runs the adjoint kernels B1-B3 (adjoint.hip)
// backwards through the segment.  The neighbour structure of each step is
// the one of the forward step (held fixed).  Gradients: dL/dv0, dL/dx0 and
// dL/dmu, dL/dK (accumulated in FP64).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "adjoint.h"
#include "rtps_internal.h"
#include "sph_device.h"

using namespace rtps;

namespace {

struct Record {   // one forward step's sorted state
    uint32_t *keys = nullptr, *idx = nullptr, *start = nullptr, *end = nullptr;
    float4 *pos_s = nullptr, *vel_s = nullptr, *ve_s = nullptr, *force_s = nullptr;
};

}  // namespace

struct rtps_adj {
    RTPSSettings settings;
    rtps_sph_params sp{};
    rtps_grid_params gp{}, gps{};
    Grid g{}, gs{};
    float dt = 0.f;
    int N = 0, nb_cells = 0, key_bits = 0, ck = 20, T = 0, collision = 1;
    uint32_t pad_key = 0;
    std::vector<void*> allocs;
    // state ping-pong (original order)
    float4 *pos[2] = {}, *vel[2] = {}, *ve[2] = {};
    std::vector<float4*> ck_pos, ck_vel, ck_ve;   // checkpoints
    std::vector<Record> rec;                       // per-step records of one segment
    uint32_t *keys = nullptr, *idx = nullptr;     // hash output / sort input
    void* sort_tmp = nullptr;
    float4* xsph = nullptr;
    // adjoint temporaries
    float4 *lam_x[2] = {}, *lam_v[2] = {}, *lam_e[2] = {};
    float4 *ga = nullptr, *gX = nullptr, *lxp = nullptr, *gx2 = nullptr, *ge = nullptr;
    float* grho = nullptr;
    float2* partial = nullptr;
    double* acc = nullptr;
    int* left_grid = nullptr;
    bool have_forward = false;

    ~rtps_adj() {
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

Win win_of(const rtps_adj* a, const Record& r) {
    return Win{r.keys, a->gps.grid_res.x, a->gps.grid_res.y, a->gps.grid_res.z, 0, 0, -1};
}

// One forward step from state slot `in` to slot `out`, recording into r.
int forward_step(rtps_adj* a, int in, int out, Record& r, hipStream_t st) {
    const int n = a->N;
    RTPS_TRY_HIP(launch_hash(n, nullptr, n, a->pad_key, a->pos[in], a->keys, a->idx, a->g, r.start,
                             a->nb_cells + 1, kSubBits, st));
    uint32_t *ks, *is;
    RTPS_TRY_HIP(sort_pairs_pingpong(n, a->key_bits, a->keys, a->idx, a->sort_tmp, st, &ks, &is));
    RTPS_TRY_HIP(hipMemcpyAsync(r.keys, ks, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    RTPS_TRY_HIP(hipMemcpyAsync(r.idx, is, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    RTPS_TRY_HIP(launch_cellindices(n, nullptr, r.keys, a->nb_cells, r.start, r.end, kSubBits, st));
    RTPS_TRY_HIP(launch_permute(n, nullptr, r.idx, a->pos[in], a->vel[in], a->ve[in], r.pos_s, r.vel_s, r.ve_s, st));
    const Win w = win_of(a, r);
    RTPS_TRY_HIP(launch_density(n, nullptr, r.pos_s, r.start, r.end, a->sp, a->gs, w, nullptr, r.ve_s, st));
    RTPS_TRY_HIP(launch_force(n, nullptr, r.pos_s, nullptr, r.ve_s, true, r.start, r.end, a->sp, a->gs, w, r.force_s,
                              a->xsph, st));
    RTPS_TRY_HIP(launch_adj_integrate(n, r.pos_s, r.vel_s, r.force_s, a->xsph, r.idx, r.start, a->nb_cells, a->sp,
                                      a->gps, a->collision, a->dt, a->pos[out], a->vel[out], a->ve[out],
                                      a->left_grid, st));
    return RTPS_OK;
}

}  // namespace

extern "C" {

int rtps_adj_create(const char* const* keys, const char* const* vals, int n, int max_particles,
                    int checkpoint_every, int max_steps, rtps_adj** out) {
    try {
        if (!out || n < 0 || max_particles <= 0 || checkpoint_every <= 0 || max_steps <= 0)
            return set_error(RTPS_ERR_ARG, "rtps_adj_create: bad arguments");
        *out = nullptr;
        std::unique_ptr<rtps_adj> a(new rtps_adj());
        for (int i = 0; i < n; ++i) {
            if (!keys[i] || !vals[i]) return set_error(RTPS_ERR_ARG, "null key or value");
            a->settings.SetSetting<std::string>(keys[i], vals[i]);
        }
        derive_params(a->settings, a->sp, a->gp, a->gps);
        if (a->sp.simulation_scale != 1.0f)
            return set_error(RTPS_ERR_ARG, "the adjoint supports simulation_scale = 1 only");
        if (a->settings.GetSettingAs<std::vector<float>>("gravity_alphas").size() &&
            [&] {
                for (float x : a->settings.GetSettingAs<std::vector<float>>("gravity_alphas"))
                    if (x != 0.f) return true;
                return false;
            }())
            return set_error(RTPS_ERR_ARG, "the adjoint does not support point gravity sources");
        if (integrator_of(a->settings) != 0)
            return set_error(RTPS_ERR_ARG, "the adjoint differentiates the leapfrog step only (integrator " +
                                               a->settings.GetSettingAs<std::string>("integrator") + ")");
        a->dt = a->settings.GetSettingAs<float>("time_step");
        a->g = make_grid(a->gp);
        a->gs = make_grid(a->gps);
        a->nb_cells = a->gp.nb_cells;
        if (a->nb_cells <= 0 || a->nb_cells > (1 << 28) - 2)
            return set_error(RTPS_ERR_ARG, "grid has no cells or more than 2^28 - 2 cells (" +
                                               std::to_string(a->nb_cells) + ")");
        a->key_bits = sort_key_bits(a->nb_cells) + kSubBits;
        a->pad_key = a->key_bits >= 32 ? 0xFFFFFFFFu : (1u << a->key_bits) - 1u;
        a->N = max_particles;
        a->ck = checkpoint_every;
        const size_t N = (size_t)max_particles, C = (size_t)a->nb_cells + 1;
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
        for (Record& r : a->rec) {
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
        RTPS_TRY_HIP(a->alloc((char**)&a->sort_tmp, sort_temp_bytes((int)N, a->key_bits)));
        RTPS_TRY_HIP(a->alloc(&a->xsph, N));
        RTPS_TRY_HIP(a->alloc(&a->ga, N));
        RTPS_TRY_HIP(a->alloc(&a->gX, N));
        RTPS_TRY_HIP(a->alloc(&a->lxp, N));
        RTPS_TRY_HIP(a->alloc(&a->gx2, N));
        RTPS_TRY_HIP(a->alloc(&a->ge, N));
        RTPS_TRY_HIP(a->alloc(&a->grho, N));
        RTPS_TRY_HIP(a->alloc(&a->partial, (size_t)adjoint_blocks((int)N)));
        RTPS_TRY_HIP(a->alloc(&a->acc, 2));
        RTPS_TRY_HIP(a->alloc(&a->left_grid, 1));
        RTPS_TRY_HIP(sort_clear_temp(a->sort_tmp, (int)N, nullptr));
        for (Record& r : a->rec) RTPS_TRY_HIP(launch_fill_u32(r.end, 0u, (long)C, nullptr));
        RTPS_TRY_HIP(hipDeviceSynchronize());
        *out = a.release();
        return RTPS_OK;
    } catch (const std::exception& e) {
        return set_error(RTPS_ERR_ARG, std::string("rtps_adj_create: ") + e.what());
    }
}

int rtps_adj_destroy(rtps_adj* a) {
    delete a;
    return RTPS_OK;
}

int rtps_adj_forward(rtps_adj* a, const rtps_float4* pos0, const rtps_float4* vel0, int nsteps, float viscosity,
                     float gas_constant, int collision, rtps_float4* pos_T, rtps_float4* vel_T, void* stream) {
    if (!a || !pos0 || !vel0 || !pos_T || nsteps <= 0 ||
        (size_t)((nsteps + a->ck - 1) / a->ck) > a->ck_pos.size())
        return set_error(RTPS_ERR_ARG, "rtps_adj_forward: bad arguments or more steps than the handle holds");
    hipStream_t st = (hipStream_t)stream;
    a->sp.viscosity = viscosity;
    a->sp.K = gas_constant;
    a->collision = collision ? 1 : 0;
    a->T = nsteps;
    const size_t nb = (size_t)a->N * 16;
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

int rtps_adj_backward(rtps_adj* a, const rtps_float4* dL_dpos_T, const rtps_float4* dL_dvel_T, rtps_float4* dL_dvel0,
                      rtps_float4* dL_dpos0, double* dL_dvisc, double* dL_dgas, void* stream) {
    if (!a || !dL_dpos_T || !dL_dvel0) return set_error(RTPS_ERR_ARG, "rtps_adj_backward: bad arguments");
    if (!a->have_forward) return set_error(RTPS_ERR_STATE, "rtps_adj_backward needs a forward run");
    hipStream_t st = (hipStream_t)stream;
    const int n = a->N;
    const size_t nb = (size_t)n * 16;
    int L = 0;
    RTPS_TRY_HIP(hipMemcpyAsync(a->lam_x[0], dL_dpos_T, nb, hipMemcpyDeviceToDevice, st));
    if (dL_dvel_T) RTPS_TRY_HIP(hipMemcpyAsync(a->lam_v[0], dL_dvel_T, nb, hipMemcpyDeviceToDevice, st));
    else RTPS_TRY_HIP(hipMemsetAsync(a->lam_v[0], 0, nb, st));
    RTPS_TRY_HIP(hipMemsetAsync(a->lam_e[0], 0, nb, st));
    RTPS_TRY_HIP(hipMemsetAsync(a->acc, 0, 2 * sizeof(double), st));
    const int nck = (a->T + a->ck - 1) / a->ck;
    const int nblk = adjoint_blocks(n);
    for (int c = nck - 1; c >= 0; --c) {
        const int first = c * a->ck, last = std::min(a->T, first + a->ck);
        // replay the segment from its checkpoint, recording every step
        const size_t nbb = nb;
        RTPS_TRY_HIP(hipMemcpyAsync(a->pos[0], a->ck_pos[c], nbb, hipMemcpyDeviceToDevice, st));
        RTPS_TRY_HIP(hipMemcpyAsync(a->vel[0], a->ck_vel[c], nbb, hipMemcpyDeviceToDevice, st));
        RTPS_TRY_HIP(hipMemcpyAsync(a->ve[0], a->ck_ve[c], nbb, hipMemcpyDeviceToDevice, st));
        int cur = 0;
        for (int s = first; s < last; ++s) {
            int rc = forward_step(a, cur, 1 - cur, a->rec[s - first], st);
            if (rc) return rc;
            cur = 1 - cur;
        }
        // adjoint through the segment, last step first
        for (int s = last - 1; s >= first; --s) {
            const Record& r = a->rec[s - first];
            const Win w = win_of(a, r);
            RTPS_TRY_HIP(launch_adj_b1(n, r.idx, r.pos_s, r.vel_s, r.force_s, a->sp, a->gps, a->collision, a->dt,
                                       a->lam_x[L], a->lam_v[L], a->lam_e[L], a->ga, a->gX, a->lam_v[1 - L], a->lxp,
                                       st));
            RTPS_TRY_HIP(launch_adj_b2(n, r.pos_s, r.ve_s, a->ga, a->gX, r.start, r.end, a->sp, a->gs, w, a->gx2,
                                       a->ge, a->grho, a->partial, st));
            RTPS_TRY_HIP(launch_adj_reduce(a->partial, nblk, a->acc, st));
            RTPS_TRY_HIP(launch_adj_b3(n, r.idx, r.pos_s, a->grho, a->lxp, a->gx2, a->ge, r.start, r.end, a->sp,
                                       a->gs, w, a->lam_x[1 - L], a->lam_e[1 - L], st));
            L = 1 - L;
        }
    }
    RTPS_TRY_HIP(hipMemcpyAsync(dL_dvel0, a->lam_v[L], nb, hipMemcpyDeviceToDevice, st));
    if (dL_dpos0) RTPS_TRY_HIP(hipMemcpyAsync(dL_dpos0, a->lam_x[L], nb, hipMemcpyDeviceToDevice, st));
    double acc[2] = {0, 0};
    RTPS_TRY_HIP(hipMemcpyAsync(acc, a->acc, sizeof(acc), hipMemcpyDeviceToHost, st));
    RTPS_TRY_HIP(hipStreamSynchronize(st));
    if (dL_dvisc) *dL_dvisc = acc[0];
    if (dL_dgas) *dL_dgas = acc[1];
    return RTPS_OK;
}

}  // extern "C"
