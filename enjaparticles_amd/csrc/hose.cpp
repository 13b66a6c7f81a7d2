// Particle hoses: Hose (rtpslib/system/common/Hose.cpp:30-128) with its disc
// emitter addDiscRandom (rtpslib/domain/IV.cpp:313-365).  Host code: a hose
// sprays a disc of particles, perpendicular to its velocity, every `em`
// calls, until total_n particles have left it.  The reference draws its
// perturbations from libc rand(); here each hose has its own SplitMix64
// stream (uniform in [0, 1) from the top 24 bits), so sprays are
// reproducible and independent of other users of rand().

#include <cmath>
#include <cstdint>
#include <new>
#include <string>
#include <vector>

#include "rtps_internal.h"

// The emitter's float arithmetic is kept unfused (one rounding per operation).
#pragma clang fp contract(off)

struct rtps_hose {
    float dt = 0.f;
    int total_n = 0, n_count = 0, em_count = 0, em = 0;
    rtps_float4 center{}, velocity{}, u{}, w{};
    float radius = 0.f, spacing = 0.f;
    uint64_t rng = 0;
};

namespace {

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Stand-in for rand() / RAND_MAX: uniform in [0, 1), 24 bits.
double uniform01(uint64_t& s) { return (double)(splitmix64(s) >> 40) * (1.0 / 16777216.0); }

float magnitude(const rtps_float4& v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

rtps_float4 normalize(const rtps_float4& a) {   // structs.cpp:53-57 (w kept)
    const float mag = magnitude(a);
    return rtps_float4{a.x / mag, a.y / mag, a.z / mag, a.w};
}

// Hose::calc_vectors (Hose.cpp:67-92): two unit vectors spanning the disc.
void calc_vectors(rtps_hose* h) {
    rtps_float4& v = h->velocity;
    if (v.y == 0.f) v.y = .0000001f;   // the reference's guard against dividing by zero
    if (v.z == 0.f) v.z = .0000001f;
    rtps_float4 u{1.f, -(v.x / v.y), 0.f, 1.f};
    const float b = 1.f;
    const float a = b * v.x / v.y;
    const float c = -b * (v.x * v.x + v.y * v.y) / (v.y * v.z);
    rtps_float4 w{a, b, c, 1.f};
    h->u = normalize(u);
    h->w = normalize(w);
}

// addDiscRandom (IV.cpp:313-365): up to num particles on a perturbed grid of
// the disc of the given radius around center, each displaced along v by a
// random multiple in [-10, 10].
std::vector<rtps_float4> add_disc_random(rtps_hose* hs, int num, const rtps_float4& center, const rtps_float4& v,
                                         const rtps_float4& u, const rtps_float4& w, float radius, float spacing) {
    const float pert = .05f * spacing;
    const float vpert = 10.f;
    const rtps_float4 umin{-radius * u.x, -radius * u.y, -radius * u.z, -radius * u.w};
    const rtps_float4 wmin{-radius * w.x, -radius * w.y, -radius * w.z, -radius * w.w};
    std::vector<rtps_float4> rvec;
    int i = 0;
    const float r2 = radius * radius;
    for (float du = 0.f; du < 2. * radius; du += spacing) {
        for (float dw = 0.f; dw < 2. * radius; dw += spacing) {
            if (i >= num) break;
            const float rrv = (float)(uniform01(hs->rng) * vpert * 2. - vpert);
            const float rru = (float)(uniform01(hs->rng) * pert * 2. - pert);
            const float rrw = (float)(uniform01(hs->rng) * pert * 2. - pert);
            du += rru;
            dw += rrw;
            rtps_float4 part;
            part.x = center.x + umin.x + u.x * du + wmin.x + w.x * dw + rrv * v.x;
            part.y = center.y + umin.y + u.y * du + wmin.y + w.y * dw + rrv * v.y;
            part.z = center.z + umin.z + u.z * du + wmin.z + w.z * dw + rrv * v.z;
            part.w = 1.0f;
            const float dx = part.x - center.x, dy = part.y - center.y, dz = part.z - center.z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 < r2) {
                rvec.push_back(part);
                ++i;
            }
        }
    }
    return rvec;
}

}  // namespace

namespace rtps {

std::vector<rtps_float4> hose_spray(rtps_hose* h) {
    // Hose::spray (Hose.cpp:111-128) with calc_em (:94-109)
    ++h->em_count;
    const float magv = magnitude(h->velocity);
    h->em = (int)(1 + h->spacing / h->dt / magv / 10);
    std::vector<rtps_float4> particles;
    if (h->em_count >= h->em && h->n_count > 0) {
        const rtps_float4 v{h->velocity.x * h->dt, h->velocity.y * h->dt, h->velocity.z * h->dt,
                            h->velocity.w * h->dt};
        particles = add_disc_random(h, h->n_count, h->center, v, h->u, h->w, h->radius, h->spacing);
        h->n_count -= (int)particles.size();
        h->em_count = 0;
    }
    return particles;
}

}  // namespace rtps

extern "C" {

int rtps_hose_create(float dt, int total_n, rtps_float4 center, rtps_float4 velocity, float radius, float spacing,
                     unsigned long long seed, rtps_hose** out) {
    if (!out || total_n < 0 || !(dt > 0.f) || !(spacing > 0.f) || !(radius >= 0.f))
        return rtps::set_error(RTPS_ERR_ARG, "rtps_hose_create: bad arguments");
    if (velocity.x == 0.f && velocity.y == 0.f && velocity.z == 0.f)
        return rtps::set_error(RTPS_ERR_ARG, "rtps_hose_create: the hose needs a nonzero velocity");
    rtps_hose* h = new (std::nothrow) rtps_hose();
    if (!h) return rtps::set_error(RTPS_ERR_ARG, "rtps_hose_create: out of memory");
    h->dt = dt;
    h->total_n = total_n;
    h->n_count = total_n;
    h->center = center;
    h->velocity = velocity;
    h->radius = radius;
    h->spacing = spacing;
    h->rng = seed;
    calc_vectors(h);
    *out = h;
    return RTPS_OK;
}

int rtps_hose_destroy(rtps_hose* h) {
    delete h;
    return RTPS_OK;
}

int rtps_hose_update(rtps_hose* h, rtps_float4 center, rtps_float4 velocity, float radius, float spacing) {
    if (!h || !(spacing > 0.f) || !(radius >= 0.f)) return rtps::set_error(RTPS_ERR_ARG, "rtps_hose_update: bad arguments");
    if (velocity.x == 0.f && velocity.y == 0.f && velocity.z == 0.f)
        return rtps::set_error(RTPS_ERR_ARG, "rtps_hose_update: the hose needs a nonzero velocity");
    h->center = center;   // Hose::update (Hose.cpp:48-56)
    h->velocity = velocity;
    h->radius = radius;
    h->spacing = spacing;
    calc_vectors(h);
    return RTPS_OK;
}

int rtps_hose_refill(rtps_hose* h, int num) {
    if (!h || num < 0) return rtps::set_error(RTPS_ERR_ARG, "rtps_hose_refill: bad arguments");
    h->total_n = num;   // Hose::refill (Hose.cpp:58-62)
    h->n_count = num;
    return RTPS_OK;
}

int rtps_hose_spray(rtps_hose* h, rtps_float4* out, int capacity, int* n) {
    if (!h || !n || capacity < 0 || (capacity > 0 && !out))
        return rtps::set_error(RTPS_ERR_ARG, "rtps_hose_spray: bad arguments");
    if (capacity < h->n_count)
        return rtps::set_error(RTPS_ERR_CAPACITY, "rtps_hose_spray: capacity below the particles left in the hose");
    try {
        std::vector<rtps_float4> p = rtps::hose_spray(h);
        for (size_t k = 0; k < p.size(); ++k) out[k] = p[k];
        *n = (int)p.size();
        return RTPS_OK;
    } catch (const std::exception& e) {
        return rtps::set_error(RTPS_ERR_ARG, std::string("rtps_hose_spray: ") + e.what());
    }
}

int rtps_hose_state(rtps_hose* h, int* n_left, int* em_count, int* em, rtps_float4* u, rtps_float4* w,
                    rtps_float4* velocity) {
    if (!h) return rtps::set_error(RTPS_ERR_ARG, "null hose");
    if (n_left) *n_left = h->n_count;
    if (em_count) *em_count = h->em_count;
    if (em) *em = h->em;
    if (u) *u = h->u;
    if (w) *w = h->w;
    if (velocity) *velocity = h->velocity;
    return RTPS_OK;
}

}  // extern "C"
