// C ABI of the stage operators and the error channel (include/rtps_c.h).
// Every entry point validates its arguments on the host, launches on the
// given stream and reports failures through a status code plus
// rtps_last_error(); no exception leaves this file.

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <vector>

#include "rtps_internal.h"
#include "sph_device.h"

namespace rtps {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_error(hipError_t e, const char* where) {
    return set_error(RTPS_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

}  // namespace rtps

using namespace rtps;

#define RTPS_CHECK(cond, msg) \
    do {                      \
        if (!(cond)) return set_error(RTPS_ERR_ARG, msg); \
    } while (0)

#define RTPS_HIP(expr)                                   \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return hip_error(e_, #expr); \
    } while (0)

static inline hipStream_t S(void* s) { return (hipStream_t)s; }

extern "C" {

const char* rtps_last_error(void) { return g_last_error.c_str(); }

int rtps_version(void) { return 1; }

int rtps_derive_params(const char* const* keys, const char* const* vals, int n, rtps_sph_params* sphp,
                       rtps_grid_params* gp, rtps_grid_params* gps) {
    RTPS_CHECK(n >= 0 && (n == 0 || (keys && vals)) && sphp && gp && gps, "rtps_derive_params: bad arguments");
    try {
        RTPSSettings s;
        for (int i = 0; i < n; ++i) {
            RTPS_CHECK(keys[i] && vals[i], "rtps_derive_params: null key or value");
            s.SetSetting<std::string>(keys[i], vals[i]);
        }
        derive_params(s, *sphp, *gp, *gps);
        return RTPS_OK;
    } catch (const std::exception& e) {
        return set_error(RTPS_ERR_ARG, std::string("rtps_derive_params: ") + e.what());
    }
}

int rtps_sort_key_bits(int nb_cells) { return nb_cells < 0 ? RTPS_ERR_ARG : sort_key_bits(nb_cells); }

int rtps_hash(int num, int n_pad, const rtps_float4* pos_u, uint32_t* keys, uint32_t* idx, const rtps_grid_params* gp,
              void* stream) {
    RTPS_CHECK(num >= 0 && n_pad >= num && gp, "rtps_hash: bad sizes or null grid");
    RTPS_CHECK(num == 0 || (pos_u && keys && idx), "rtps_hash: null buffer");
    const int bits = sort_key_bits(gp->nb_cells);
    const uint32_t pad = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    RTPS_HIP(launch_hash(num, nullptr, n_pad, pad, (const float4*)pos_u, keys, idx, make_grid(*gp), nullptr, 0, 0,
                         S(stream)));
    return RTPS_OK;
}

size_t rtps_sort_temp_bytes(int n, int key_bits) { return n < 0 ? 0 : sort_temp_bytes(n, key_bits); }

int rtps_sort_pairs(int n, int key_bits, uint32_t* keys, uint32_t* vals, void* temp, size_t temp_bytes,
                    void* stream) {
    RTPS_CHECK(n >= 0 && key_bits >= 0 && key_bits <= 32, "rtps_sort_pairs: bad n or key_bits");
    if (n <= 1 || key_bits == 0) return RTPS_OK;
    RTPS_CHECK(keys && vals && temp, "rtps_sort_pairs: null buffer");
    RTPS_CHECK(temp_bytes >= sort_temp_bytes(n, key_bits), "rtps_sort_pairs: temp buffer too small");
    RTPS_HIP(sort_pairs_inplace(n, key_bits, keys, vals, temp, S(stream)));
    return RTPS_OK;
}

int rtps_sort_segments(int batch, int array_length, int dir, uint32_t* dst_keys, uint32_t* dst_vals,
                       const uint32_t* src_keys, const uint32_t* src_vals, void* temp, size_t temp_bytes,
                       void* stream) {
    RTPS_CHECK(batch >= 0 && array_length >= 0, "rtps_sort_segments: negative batch or array_length");
    if (array_length < 2 || batch == 0) return RTPS_OK;   // BitonicSort.cpp: arrayLength < 2 returns 0
    RTPS_CHECK((long)batch * array_length <= 0x7FFFFFFFL, "rtps_sort_segments: more than 2^31 - 1 pairs");
    RTPS_CHECK(dst_keys && dst_vals && src_keys && src_vals && temp, "rtps_sort_segments: null buffer");
    RTPS_CHECK(temp_bytes >= sort_temp_bytes(array_length, 32), "rtps_sort_segments: temp buffer too small");
    RTPS_HIP(sort_segments(batch, array_length, dir, dst_keys, dst_vals, src_keys, src_vals, temp, S(stream)));
    return RTPS_OK;
}

int rtps_cellindices(int num, const uint32_t* keys, int nb_cells, uint32_t* start, uint32_t* end, int* nc,
                     void* stream) {
    RTPS_CHECK(num >= 0 && nb_cells > 0 && start && end && (num == 0 || keys), "rtps_cellindices: bad arguments");
    RTPS_HIP(launch_fill_u32(start, kEmpty, (long)nb_cells + 1, S(stream)));
    RTPS_HIP(launch_cellindices(num, nullptr, keys, nb_cells, start, end, 0, S(stream)));
    if (nc) RTPS_HIP(launch_read_nc(start, nb_cells, nc, S(stream)));
    return RTPS_OK;
}

int rtps_permute(int num, const uint32_t* idx, const rtps_float4* pos_u, const rtps_float4* vel_u,
                 const rtps_float4* veleval_u, rtps_float4* pos_s, rtps_float4* vel_s, rtps_float4* veleval_s,
                 void* stream) {
    RTPS_CHECK(num >= 0, "rtps_permute: negative num");
    if (num == 0) return RTPS_OK;
    RTPS_CHECK(idx && pos_u && vel_u && veleval_u && pos_s && vel_s && veleval_s, "rtps_permute: null buffer");
    RTPS_HIP(launch_permute(num, nullptr, idx, (const float4*)pos_u, (const float4*)vel_u, (const float4*)veleval_u,
                            (float4*)pos_s, (float4*)vel_s, (float4*)veleval_s, S(stream)));
    return RTPS_OK;
}

int rtps_density(int num, const rtps_float4* pos_s, const uint32_t* start, const uint32_t* end,
                 const rtps_sph_params* sphp, const rtps_grid_params* gps, float* density, void* stream) {
    RTPS_CHECK(num >= 0 && sphp && gps, "rtps_density: bad arguments");
    if (num == 0) return RTPS_OK;
    RTPS_CHECK(pos_s && start && end && density, "rtps_density: null buffer");
    const Win plain{nullptr, 0.f, 0.f, 0.f, 0, 0, -1};
    RTPS_HIP(launch_density(num, nullptr, (const float4*)pos_s, start, end, *sphp, make_grid(*gps), plain, density,
                            nullptr, S(stream)));
    return RTPS_OK;
}

int rtps_force(int num, const rtps_float4* pos_s, const float* density, const rtps_float4* veleval_s,
               const uint32_t* start, const uint32_t* end, const rtps_sph_params* sphp, const rtps_grid_params* gps,
               rtps_float4* force, rtps_float4* xsph, void* stream) {
    RTPS_CHECK(num >= 0 && sphp && gps, "rtps_force: bad arguments");
    if (num == 0) return RTPS_OK;
    RTPS_CHECK(pos_s && density && veleval_s && start && end && force && xsph, "rtps_force: null buffer");
    const Win plain{nullptr, 0.f, 0.f, 0.f, 0, 0, -1};
    RTPS_HIP(launch_force(num, nullptr, (const float4*)pos_s, density, (const float4*)veleval_s, false, start, end,
                          *sphp, make_grid(*gps), plain, (float4*)force, (float4*)xsph, S(stream)));
    return RTPS_OK;
}

int rtps_gravity(int num, int nsrc, const rtps_float4* sources, const float* masses, const float* alphas,
                 const rtps_float4* pos_s, rtps_float4* force, float scale, void* stream) {
    RTPS_CHECK(num >= 0 && nsrc >= 0 && nsrc <= kMaxGravitySources, "rtps_gravity: bad sizes");
    if (num == 0 || nsrc == 0) return RTPS_OK;
    RTPS_CHECK(sources && masses && alphas && pos_s && force, "rtps_gravity: null buffer");
    GravitySources g{};
    g.n = nsrc;
    for (int i = 0; i < nsrc; ++i) {
        g.pos[i] = make_float4(sources[i].x, sources[i].y, sources[i].z, sources[i].w);
        g.mass[i] = masses[i];
        g.alpha[i] = alphas[i];
    }
    rtps_sph_params sp{};
    sp.simulation_scale = scale;
    rtps_grid_params gz{};
    RTPS_HIP(launch_post_force(num, nullptr, nullptr, 0, (const float4*)pos_s, nullptr, (float4*)force, sp, g, 0, gz,
                               S(stream)));
    return RTPS_OK;
}

int rtps_collision_wall(int num, const rtps_float4* pos_s, const rtps_float4* vel_s, rtps_float4* force,
                        const rtps_sph_params* sphp, const rtps_grid_params* gps, void* stream) {
    RTPS_CHECK(num >= 0 && sphp && gps, "rtps_collision_wall: bad arguments");
    if (num == 0) return RTPS_OK;
    RTPS_CHECK(pos_s && vel_s && force, "rtps_collision_wall: null buffer");
    GravitySources none{};
    RTPS_HIP(launch_post_force(num, nullptr, nullptr, 0, (const float4*)pos_s, (const float4*)vel_s, (float4*)force,
                               *sphp, none, 1, *gps, S(stream)));
    return RTPS_OK;
}

int rtps_collision_triangles(int num, const rtps_float4* pos_s, const rtps_float4* vel_s, rtps_float4* force,
                             const rtps_triangle* tris, int n, const rtps_sph_params* sphp, void* stream) {
    RTPS_CHECK(num >= 0 && n >= 0 && sphp && (tris || n == 0), "rtps_collision_triangles: bad arguments");
    if (num == 0 || n == 0) return RTPS_OK;
    RTPS_CHECK(pos_s && vel_s && force, "rtps_collision_triangles: null buffer");
    std::vector<TriPrep> prep = prepare_triangles(tris, n, sphp->simulation_scale);
    set_triangle_reach(prep, sphp->boundary_distance);
    const std::vector<float4> reach = triangle_reach(prep);
    TriPrep* dev = nullptr;
    RTPS_HIP(hipMalloc(&dev, prep.size() * (sizeof(TriPrep) + sizeof(float4))));
    float4* dev_reach = reinterpret_cast<float4*>(dev + prep.size());
    hipError_t e = hipMemcpyAsync(dev, prep.data(), prep.size() * sizeof(TriPrep), hipMemcpyHostToDevice, S(stream));
    if (e == hipSuccess)
        e = hipMemcpyAsync(dev_reach, reach.data(), reach.size() * sizeof(float4), hipMemcpyHostToDevice, S(stream));
    if (e == hipSuccess)
        e = launch_collision_tri(num, nullptr, nullptr, 0, (const float4*)pos_s, (const float4*)vel_s, (float4*)force,
                                 dev, dev_reach, n, *sphp, S(stream));
    if (e == hipSuccess) e = hipStreamSynchronize(S(stream));   // the triangle buffer is freed below
    (void)hipFree(dev);
    RTPS_HIP(e);
    return RTPS_OK;
}

int rtps_leapfrog(int num, const rtps_float4* pos_s, const rtps_float4* vel_s, const rtps_float4* force,
                  const rtps_float4* xsph, const uint32_t* idx, const rtps_sph_params* sphp, float dt,
                  rtps_float4* pos_u, rtps_float4* vel_u, rtps_float4* veleval_u, void* stream) {
    RTPS_CHECK(num >= 0 && sphp, "rtps_leapfrog: bad arguments");
    if (num == 0) return RTPS_OK;
    RTPS_CHECK(pos_s && vel_s && force && xsph && idx && pos_u && vel_u && veleval_u, "rtps_leapfrog: null buffer");
    RTPS_HIP(launch_leapfrog(num, nullptr, nullptr, 0, (const float4*)pos_s, (const float4*)vel_s,
                             (const float4*)force, (const float4*)xsph, idx, *sphp, dt, (float4*)pos_u,
                             (float4*)vel_u, (float4*)veleval_u, S(stream)));
    return RTPS_OK;
}

}  // extern "C"
