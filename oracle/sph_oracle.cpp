// CPU ORACLE for the SPH step -- TEST INFRASTRUCTURE ONLY.
//
// This file restates, on the CPU, EnjaParticles' per-timestep SPH pipeline
// (hash -> sort -> cell start/end -> permute -> [deletion] -> density ->
// force/XSPH -> gravity sources -> wall collision -> leapfrog).  It is the
// checker the HIP path is compared against; it is never the product.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
//
// Everything is templated on the arithmetic type T:
//   T = float  follows the reference's FP32 operation order line by line
//              (built with -ffp-contract=off so no multiply-add is fused);
//   T = double is the accuracy reference (same algorithm, FP64 state).
//
// Reference files followed (paths relative to the reference root):
//   hash           rtpslib/system/common/cl_src/hash.cl:50-157, cl_hash.h:33-114
//   sort           rtpslib/system/System.cpp:407-480 (ascending key/value sort)
//   cellindices    rtpslib/system/common/cl_src/cellindices.cl:12-107,
//                  rtpslib/system/common/CellIndices.cpp:59-99 (host clear, nc)
//   permute        rtpslib/system/common/cl_src/permute.cl:10-46
//   deletion path  rtpslib/system/SPH.cpp:220-250, 445-457
//   neighbours     rtpslib/system/sph/cl_src/cl_sph_neighbors.h:49-140
//   density        rtpslib/system/sph/cl_src/density.cl:44-126, cl_sph_kernels.h:29-45
//   force/XSPH     rtpslib/system/sph/cl_src/force.cl:43-186, cl_sph_kernels.h:81-127
//   gravity        rtpslib/system/common/cl_src/gravity.cl:9-36
//   collision      rtpslib/system/sph/cl_src/collision_wall.cl:31-119,
//                  rtpslib/system/sph/cl_src/cl_sph_collision.h:31-73
//   leapfrog       rtpslib/system/sph/cl_src/leapfrog.cl:30-121
//   euler          rtpslib/system/sph/cl_src/euler.cl:33-78 (integrator = euler)
//   step order     rtpslib/system/SPH.cpp:149-300, 389-441
//   parameters     rtpslib/system/SPHSettings.cpp:30-201, rtpslib/RTPSSettings.h:65-85,
//                  rtpslib/domain/Domain.cpp:38-81, rtpslib/system/System.cpp:294-329
//
// Pinning (what checks this restatement independently of its own formulas):
// closed-form KATs (isolated particle density, simple-cubic lattice density,
// two-particle force, free fall, velocity clamp, single-wall repulsion),
// brute-force O(N^2) neighbour sums, momentum antisymmetry, sort/searchsorted
// identities and FP32-vs-FP64 agreement; see tests/test_oracle.py.
// The bitonic network (oracle_sort_bitonic) is Batcher's textbook network and
// is used only to give the CPU baseline the reference's O(N log^2 N) cost.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>
#include <climits>
#include <cstddef>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint32_t u32;

// ---------------------------------------------------------------------------
// Parameter blocks: byte layout of SPHParams (rtpslib/system/SPHSettings.h:38-136)
// and GridParams (rtpslib/domain/Domain.h:144-172).
struct OF4 { float x, y, z, w; };

struct alignas(16) OSPHParams {
    float rest_density, mass, rest_distance, smoothing_distance, simulation_scale;
    float boundary_stiffness, boundary_dampening, boundary_distance, K;
    float viscosity, velocity_limit, xsph_factor;
    float friction_coef, restitution_coef, shear, attraction, spring;
    float EPSILON;
    float wpoly6_coef, wpoly6_d_coef, wpoly6_dd_coef, wspiky_coef, wspiky_d_coef;
    float wspiky_dd_coef, wvisc_coef, wvisc_d_coef, wvisc_dd_coef;
    int num, max_num;
    alignas(16) OF4 gravity;
};
static_assert(sizeof(OSPHParams) == 144, "SPHParams is 144 bytes");
static_assert(offsetof(OSPHParams, gravity) == 128, "gravity at offset 128");

struct alignas(16) OGridParams {
    OF4 grid_size, grid_min, grid_max, bnd_min, bnd_max, grid_res, grid_delta;
    int nb_cells;
    int pad_[3];
};
static_assert(sizeof(OGridParams) == 128, "GridParams is 128 bytes");

// ---------------------------------------------------------------------------
// Float -> int conversion with the hardware's saturating behaviour
// (NaN -> 0, out of range -> INT_MIN / INT_MAX), so the oracle is defined for
// every input.  The reference relies on the device's (int) cast here
// (cl_hash.h:39-42, :113).
template <typename T> static inline int sat_f2i(T v) {
    if (!(v == v)) return 0;
    if (v >= (T)2147483648.0) return INT_MAX;
    if (v <= (T)-2147483648.0) return INT_MIN;
    return (int)v;
}

template <typename T> struct V4 { T x, y, z, w; };

// ---------------------------------------------------------------------------
// Tangent-linear arithmetic: a value with kTL directional derivatives
// (forward-mode differentiation).  Running the FP64 step on TL numbers gives
// J . u exactly for kTL input directions u at once, where J is the Jacobian
// of the step at the current state with every branch (neighbour within h,
// wall contact, friction sign, velocity clamp) fixed as the values decide it:
// the derivative an adjoint of the same step computes, free of the
// truncation and branch-switching error of finite differences.  Comparisons
// and casts act on the value.  Layout: value then the kTL derivatives (as
// double[1 + kTL]).
constexpr int kTL = 3;
struct TL {
    double v, d[kTL];
    TL() = default;
    constexpr TL(double x) : v(x), d{} {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    TL& operator+=(const TL& o) { v += o.v; for (int k = 0; k < kTL; ++k) d[k] += o.d[k]; return *this; }
    TL& operator-=(const TL& o) { v -= o.v; for (int k = 0; k < kTL; ++k) d[k] -= o.d[k]; return *this; }
    TL& operator*=(const TL& o) { for (int k = 0; k < kTL; ++k) d[k] = d[k] * o.v + v * o.d[k]; v *= o.v; return *this; }
    TL& operator/=(const TL& o) {
        const double q = v / o.v;
        for (int k = 0; k < kTL; ++k) d[k] = (d[k] - q * o.d[k]) / o.v;
        v = q;
        return *this;
    }
};
static inline TL operator+(TL a, const TL& b) { return a += b; }
static inline TL operator-(TL a, const TL& b) { return a -= b; }
static inline TL operator*(TL a, const TL& b) { return a *= b; }
static inline TL operator/(TL a, const TL& b) { return a /= b; }
static inline TL operator-(const TL& a) { TL r = a; r.v = -a.v; for (int k = 0; k < kTL; ++k) r.d[k] = -a.d[k]; return r; }
static inline bool operator<(const TL& a, const TL& b) { return a.v < b.v; }
static inline bool operator>(const TL& a, const TL& b) { return a.v > b.v; }
static inline bool operator<=(const TL& a, const TL& b) { return a.v <= b.v; }
static inline bool operator>=(const TL& a, const TL& b) { return a.v >= b.v; }
static inline bool operator==(const TL& a, const TL& b) { return a.v == b.v; }
namespace std {
static inline TL sqrt(const TL& a) {
    TL r;
    r.v = std::sqrt(a.v);
    for (int k = 0; k < kTL; ++k) r.d[k] = r.v > 0 ? a.d[k] / (2 * r.v) : 0.0;
    return r;
}
}  // namespace std

// Tangents of the scalar parameters (viscosity, gas constant) of the next
// oracle_step_tl call, one per direction; zero for the other steps.
static double g_tl_dmu[kTL], g_tl_dK[kTL];
template <typename T> static inline T param_viscosity(float v) { return (T)v; }
template <typename T> static inline T param_gas_constant(float v) { return (T)v; }
template <> inline TL param_viscosity<TL>(float v) {
    TL r((double)v);
    for (int k = 0; k < kTL; ++k) r.d[k] = g_tl_dmu[k];
    return r;
}
template <> inline TL param_gas_constant<TL>(float v) {
    TL r((double)v);
    for (int k = 0; k < kTL; ++k) r.d[k] = g_tl_dK[k];
    return r;
}

template <typename T> static inline V4<T> ld(const T* a, long i) {
    return V4<T>{a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]};
}
template <typename T> static inline void st(T* a, long i, V4<T> v) {
    a[4 * i] = v.x; a[4 * i + 1] = v.y; a[4 * i + 2] = v.z; a[4 * i + 3] = v.w;
}

// calcGridCell, cl_hash.h:33-54: truncation toward zero of (p-min)*delta.
template <typename T>
static inline void grid_cell(const V4<T>& p, const OGridParams& g, int& cx, int& cy, int& cz) {
    T px = (p.x - (T)g.grid_min.x) * (T)g.grid_delta.x;
    T py = (p.y - (T)g.grid_min.y) * (T)g.grid_delta.y;
    T pz = (p.z - (T)g.grid_min.z) * (T)g.grid_delta.z;
    cx = sat_f2i(px); cy = sat_f2i(py); cz = sat_f2i(pz);
}

// calcGridHash, cl_hash.h:57-114 (wrapEdges = false).  The reference
// evaluates the linear index in float because grid_res is a float4 (:113).
template <typename T>
static inline int grid_hash(int gx, int gy, int gz, const OGridParams& g) {
    T v = ((T)gz * (T)g.grid_res.y + (T)gy) * (T)g.grid_res.x + (T)gx;
    return sat_f2i(v);
}

// ---------------------------------------------------------------------------
// a2: hash.cl:50-157
template <typename T>
static void hash_impl(int num, const T* pos_u, const OGridParams* gp, u32* keys, u32* idx) {
    const OGridParams g = *gp;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < num; ++i) {
        V4<T> p = ld(pos_u, i);
        int cx, cy, cz;
        grid_cell(p, g, cx, cy, cz);
        int h = grid_hash<T>(cx, cy, cz, g);
        h = h > g.nb_cells ? g.nb_cells : h;    // hash.cl:91
        h = h < 0 ? g.nb_cells : h;             // hash.cl:92
        keys[i] = (u32)h;
        idx[i] = (u32)i;
    }
}
extern "C" void oracle_hash_f32(int num, const float* pos_u, const OGridParams* gp, u32* keys, u32* idx) {
    hash_impl<float>(num, pos_u, gp, keys, idx);
}
extern "C" void oracle_hash_f64(int num, const double* pos_u, const OGridParams* gp, u32* keys, u32* idx) {
    hash_impl<double>(num, pos_u, gp, keys, idx);
}

// ---------------------------------------------------------------------------
// a3: ascending key/value sort.  The oracle's parity order is the STABLE one
// (equal keys keep ascending original index), which any correct sort matches
// on keys and per-cell index sets.
extern "C" void oracle_sort_stable(int num, u32* keys, u32* vals) {
    std::vector<std::pair<u32, u32>> kv((size_t)num);
    for (int i = 0; i < num; ++i) kv[i] = {keys[i], vals[i]};
    std::stable_sort(kv.begin(), kv.end(),
                     [](const std::pair<u32, u32>& a, const std::pair<u32, u32>& b) { return a.first < b.first; });
    for (int i = 0; i < num; ++i) { keys[i] = kv[i].first; vals[i] = kv[i].second; }
}

static unsigned nlpo2(unsigned x) {  // next power of two (RTPSSettings.cpp:31-40 semantics)
    x--; x |= x >> 1; x |= x >> 2; x |= x >> 4; x |= x >> 8; x |= x >> 16; return x + 1;
}

// Batcher's bitonic sorting network over L = nlpo2(num) entries, padded with
// INT_MAX keys like the reference's sort buffers (System.cpp:274-284, :414).
// Unstable; used for the CPU baseline's cost profile.
extern "C" void oracle_sort_bitonic(int num, u32* keys, u32* vals) {
    if (num <= 1) return;
    const long L = (long)nlpo2((unsigned)num);
    std::vector<u32> k((size_t)L, (u32)INT_MAX), v((size_t)L, (u32)INT_MAX);
    std::memcpy(k.data(), keys, sizeof(u32) * num);
    std::memcpy(v.data(), vals, sizeof(u32) * num);
    for (long size = 2; size <= L; size <<= 1) {
        for (long stride = size >> 1; stride > 0; stride >>= 1) {
#pragma omp parallel for schedule(static)
            for (long i = 0; i < L; ++i) {
                long l = i ^ stride;
                if (l > i) {
                    bool up = (i & size) == 0;
                    if ((k[i] > k[l]) == up) { std::swap(k[i], k[l]); std::swap(v[i], v[l]); }
                }
            }
        }
    }
    std::memcpy(keys, k.data(), sizeof(u32) * num);
    std::memcpy(vals, v.data(), sizeof(u32) * num);
}

// ---------------------------------------------------------------------------
// a4: cellindices.cl:12-107 with the host clear of start[0..C] to 0xFFFFFFFF
// (CellIndices.cpp:59-62).  end[] is not cleared (stale for empty cells, as in
// the reference).  Returns nc = (int)start[nb_cells] (CellIndices.cpp:95-99).
extern "C" int oracle_cellindices(int num, const u32* keys, int nb_cells, u32* start, u32* end) {
    const u32 nc = (u32)nb_cells;
    for (long c = 0; c <= nb_cells; ++c) start[c] = 0xFFFFFFFFu;
    for (int i = 0; i < num; ++i) {
        u32 h = keys[i];
        if (h > nc) continue;                       // cellindices.cl:32-35
        if (i == 0) start[h] = 0;
        if (i > 0) {
            u32 prev = keys[i - 1] < nc ? keys[i - 1] : nc;   // :49-50
            if (prev != h) { start[h] = (u32)i; end[prev] = (u32)i; }
        }
        if (i == num - 1) end[h] = (u32)(i + 1);
    }
    return (int)start[nb_cells];
}

// ---------------------------------------------------------------------------
// a5: permute.cl:10-46 (pos, vel, veleval; colour is an identity round trip)
template <typename T>
static void permute_impl(int num, const u32* idx, const T* pu, const T* vu, const T* eu, T* ps, T* vs, T* es) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < num; ++i) {
        long s = idx[i];
        for (int c = 0; c < 4; ++c) {
            ps[4 * (long)i + c] = pu[4 * s + c];
            vs[4 * (long)i + c] = vu[4 * s + c];
            es[4 * (long)i + c] = eu[4 * s + c];
        }
    }
}
extern "C" void oracle_permute_f32(int num, const u32* idx, const float* pu, const float* vu, const float* eu,
                        float* ps, float* vs, float* es) {
    permute_impl<float>(num, idx, pu, vu, eu, ps, vs, es);
}
extern "C" void oracle_permute_f64(int num, const u32* idx, const double* pu, const double* vu, const double* eu,
                        double* ps, double* vs, double* es) {
    permute_impl<double>(num, idx, pu, vu, eu, ps, vs, es);
}

// ---------------------------------------------------------------------------
// Neighbour iteration, cl_sph_neighbors.h:99-140 and :49-95: the 27 cells
// z,y,x in cell+-1 (no edge wrap), skipping hash >= nb_cells (unsigned) and
// empty cells, then j in [start, end).
template <typename T, typename F>
static inline void for_neighbours(const V4<T>& pi, const T* pos_s, T scale, const u32* start, const u32* end,
                                  const OGridParams& gs, F&& fn) {
    int cx, cy, cz;
    grid_cell(pi, gs, cx, cy, cz);
    for (int z = cz - 1; z <= cz + 1; ++z)
        for (int y = cy - 1; y <= cy + 1; ++y)
            for (int x = cx - 1; x <= cx + 1; ++x) {
                u32 ch = (u32)grid_hash<T>(x, y, z, gs);
                if (ch >= (u32)gs.nb_cells) continue;
                u32 s = start[ch];
                if (s == 0xFFFFFFFFu) continue;
                u32 e = end[ch];
                for (u32 j = s; j < e; ++j) {
                    V4<T> pj = ld(pos_s, j);
                    pj.x *= scale; pj.y *= scale; pj.z *= scale; pj.w *= scale;
                    fn(j, pj);
                }
            }
}

template <typename T> static inline T len3(T x, T y, T z) { return std::sqrt(x * x + y * y + z * z); }

// a7: density.cl:44-126, Wpoly6 cl_sph_kernels.h:29-45.  Self is included.
template <typename T>
static void density_impl(int num, const T* pos_s, const u32* start, const u32* end, const OSPHParams* sp,
                         const OGridParams* gps, T* density) {
    const OSPHParams p = *sp;
    const OGridParams gs = *gps;
    const T scale = p.simulation_scale, h = p.smoothing_distance;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < num; ++i) {
        V4<T> pi = ld(pos_s, i);
        pi.x *= scale; pi.y *= scale; pi.z *= scale; pi.w *= scale;
        T sum = 0;
        for_neighbours<T>(pi, pos_s, scale, start, end, gs, [&](u32, const V4<T>& pj) {
            T rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            T rlen = len3(rx, ry, rz);
            if (rlen <= h) {
                T r2 = rx * rx + ry * ry + rz * rz;
                T hr2 = h * h - r2;
                sum += hr2 * hr2 * hr2;
            }
        });
        density[i] = (T)p.wpoly6_coef * (T)p.mass * sum;
    }
}
extern "C" void oracle_density_f32(int num, const float* pos_s, const u32* start, const u32* end, const OSPHParams* sp,
                        const OGridParams* gps, float* density) {
    density_impl<float>(num, pos_s, start, end, sp, gps, density);
}
extern "C" void oracle_density_f64(int num, const double* pos_s, const u32* start, const u32* end, const OSPHParams* sp,
                        const OGridParams* gps, double* density) {
    density_impl<double>(num, pos_s, start, end, sp, gps, density);
}

// a8: force.cl:43-186 (pressure with Wspiky_dr, viscosity with Wvisc_lapl,
// XSPH with Wpoly6, all on veleval), epilogue force.cl:181-185.
template <typename T>
static void force_impl(int num, const T* pos_s, const T* density, const T* veleval, const u32* start, const u32* end,
                       const OSPHParams* sp, const OGridParams* gps, T* force, T* xsph) {
    const OSPHParams p = *sp;
    const OGridParams gs = *gps;
    const T scale = p.simulation_scale, h = p.smoothing_distance, eps = p.EPSILON;
    const T K = param_gas_constant<T>(p.K), rho0 = p.rest_density;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < num; ++i) {
        V4<T> pi = ld(pos_s, i);
        pi.x *= scale; pi.y *= scale; pi.z *= scale; pi.w *= scale;
        const T di = density[i];
        const V4<T> vi = ld(veleval, i);
        T fx = 0, fy = 0, fz = 0;
        T vx = 0, vy = 0, vz = 0, vw = 0;
        T xx = 0, xy = 0, xz = 0;
        for_neighbours<T>(pi, pos_s, scale, start, end, gs, [&](u32 j, const V4<T>& pj) {
            T rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            T rlen = len3(rx, ry, rz);
            if (rlen <= h) {
                T iej = (T)((u32)i != j);
                rlen = std::max(rlen, eps);
                T hr = h - rlen;
                T dWijdr = hr * hr / rlen;
                T dj = density[j];
                T idj = (T)1 / dj;
                T Pi = K * (di - rho0);
                T Pj = K * (dj - rho0);
                T kern = dWijdr * (Pi + Pj) * idj;
                V4<T> vj = ld(veleval, j);
                T lap = h - rlen;
                vx += (vj.x - vi.x) * lap * idj * iej;
                vy += (vj.y - vi.y) * lap * idj * iej;
                vz += (vj.z - vi.z) * lap * idj * iej;
                vw += (vj.w - vi.w) * lap * idj * iej;
                T r2 = rx * rx + ry * ry + rz * rz;
                T hr2 = h * h - r2;
                T w6 = hr2 * hr2 * hr2;
                T dd = di + dj;
                xx += w6 * (vj.x - vi.x) / dd * iej;
                xy += w6 * (vj.y - vi.y) / dd * iej;
                xz += w6 * (vj.z - vi.z) / dd * iej;
                fx += rx * kern * iej;
                fy += ry * kern * iej;
                fz += rz * kern * iej;
            }
        });
        T a = (T)p.mass * ((T)1 / di);
        T cp = (T)-0.5 * (T)p.wspiky_d_coef;
        T cv = param_viscosity<T>(p.viscosity) * (T)p.wvisc_dd_coef;
        V4<T> F{a * (cp * fx + cv * vx), a * (cp * fy + cv * vy), a * (cp * fz + cv * vz), a * (cp * (T)0 + cv * vw)};
        T cx = (T)2 * (T)p.mass * (T)p.wpoly6_coef;
        V4<T> X{cx * xx, cx * xy, cx * xz, (T)0};
        st(force, i, F);
        st(xsph, i, X);
    }
}
extern "C" void oracle_force_f32(int num, const float* pos_s, const float* density, const float* veleval, const u32* start,
                      const u32* end, const OSPHParams* sp, const OGridParams* gps, float* force, float* xsph) {
    force_impl<float>(num, pos_s, density, veleval, start, end, sp, gps, force, xsph);
}
extern "C" void oracle_force_f64(int num, const double* pos_s, const double* density, const double* veleval, const u32* start,
                      const u32* end, const OSPHParams* sp, const OGridParams* gps, double* force, double* xsph) {
    force_impl<double>(num, pos_s, density, veleval, start, end, sp, gps, force, xsph);
}

// ---------------------------------------------------------------------------
// a9: gravity.cl:9-36 (point attractors), accel += mag * fast_normalize(vect).
template <typename T>
static void gravity_impl(int num, int nsrc, const float* src, const float* smass, const float* alpha, const T* pos_s,
                         T* force, T scale) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < num; ++i) {
        V4<T> p = ld(pos_s, i);
        V4<T> a = ld(force, i);
        for (int s = 0; s < nsrc; ++s) {
            T vx = scale * ((T)src[4 * s] - p.x), vy = scale * ((T)src[4 * s + 1] - p.y),
              vz = scale * ((T)src[4 * s + 2] - p.z);
            T d2 = vx * vx + vy * vy + vz * vz;
            T inv = d2 > (T)0 ? (T)1 / std::sqrt(d2) : (T)0;   // fast_normalize; zero vector stays zero
            T mag = (T)alpha[s] * (T)smass[s];
            T cut = (T)0.5 * scale;
            cut *= cut;
            if (d2 < cut) mag /= -cut; else mag /= d2;
            a.x += mag * (vx * inv); a.y += mag * (vy * inv); a.z += mag * (vz * inv);
        }
        st(force, i, a);
    }
}
extern "C" void oracle_gravity_f32(int num, int nsrc, const float* src, const float* smass, const float* alpha,
                        const float* pos_s, float* force, float scale) {
    gravity_impl<float>(num, nsrc, src, smass, alpha, pos_s, force, scale);
}
extern "C" void oracle_gravity_f64(int num, int nsrc, const float* src, const float* smass, const float* alpha,
                        const double* pos_s, double* force, double scale) {
    gravity_impl<double>(num, nsrc, src, smass, alpha, pos_s, force, scale);
}

// ---------------------------------------------------------------------------
// a10: collision_wall.cl:31-119 with cl_sph_collision.h:31-73.  Friction
// kinetic and static limit are hard-coded to 0 (collision_wall.cl:61-62);
// v_n = v * dot(n, v) is the reference's formula (cl_sph_collision.h:54).
template <typename T>
static inline void wall(T diff, T nx, T ny, T nz, const V4<T>& v, const OSPHParams& p, V4<T>& rf, V4<T>& ff) {
    if (!(diff > (T)p.EPSILON)) return;
    T ndv = nx * v.x + ny * v.y + nz * v.z;
    T m = (T)p.boundary_stiffness * diff - (T)p.boundary_dampening * ndv;
    rf.x += m * nx; rf.y += m * ny; rf.z += m * nz;
    T vtx = v.x - v.x * ndv, vty = v.y - v.y * ndv, vtz = v.z - v.z * ndv;
    if ((vtx + vty + vtz) / (T)3 > (T)0) { ff.x += -vtx; ff.y += -vty; ff.z += -vtz; }
    else { ff.x += (T)0 * -vtx; ff.y += (T)0 * -vty; ff.z += (T)0 * -vtz; }
}

template <typename T>
static void collision_impl(int num, const T* pos_s, const T* vel_s, T* force, const OSPHParams* sp,
                           const OGridParams* gps) {
    const OSPHParams p = *sp;
    const OGridParams g = *gps;
    const T scale = p.simulation_scale;
    const T bd = (T)p.rest_distance * (T)0.5;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < num; ++i) {
        V4<T> x = ld(pos_s, i);
        x.x *= scale; x.y *= scale; x.z *= scale;
        V4<T> v = ld(vel_s, i);
        V4<T> rf{0, 0, 0, 0}, ff{0, 0, 0, 0};
        wall(bd - (x.z - (T)g.bnd_min.z), (T)0, (T)0, (T)1, v, p, rf, ff);
        wall(bd - ((T)g.bnd_max.z - x.z), (T)0, (T)0, (T)-1, v, p, rf, ff);
        wall(bd - (x.y - (T)g.bnd_min.y), (T)0, (T)1, (T)0, v, p, rf, ff);
        wall(bd - ((T)g.bnd_max.y - x.y), (T)0, (T)-1, (T)0, v, p, rf, ff);
        wall(bd - (x.x - (T)g.bnd_min.x), (T)1, (T)0, (T)0, v, p, rf, ff);
        wall(bd - ((T)g.bnd_max.x - x.x), (T)-1, (T)0, (T)0, v, p, rf, ff);
        V4<T> f = ld(force, i);
        f.x += rf.x + ff.x; f.y += rf.y + ff.y; f.z += rf.z + ff.z;
        st(force, i, f);
    }
}
extern "C" void oracle_collision_wall_f32(int num, const float* pos_s, const float* vel_s, float* force, const OSPHParams* sp,
                               const OGridParams* gps) {
    collision_impl<float>(num, pos_s, vel_s, force, sp, gps);
}
extern "C" void oracle_collision_wall_f64(int num, const double* pos_s, const double* vel_s, double* force,
                               const OSPHParams* sp, const OGridParams* gps) {
    collision_impl<double>(num, pos_s, vel_s, force, sp, gps);
}

// ---------------------------------------------------------------------------
// f4: rigid-body coupling of SPH::interact (SPH.cpp:610-663): fluid particle i
// against the boundary particles j of a rigid-body system, found through that
// system's cell start/end tables on the fluid's scaled grid with the same
// 27-cell iteration (cl_sph_neighbors.h:99-140), accepted within 2h:
//   norm = (x_j - x_i) / max(|r|, eps), stiff = s0 m,
//   F   += -stiff (2h - r) norm + s1 sqrt(stiff mn) ((vrel . norm) norm)
//   V   += w_visc (h - r)
// static  (staticrigidbody_force.cl:45-146): vrel = -e_i (fluid veleval),
//         mn = m^2 / 2m, w_visc = vrel; force_i += F / m - m / rho_i^2 mu c V
// dynamic (rigidbody_force.cl:118-193): vrel = v_j - e_i, mn = m m_j / (m + m_j),
//         w_visc = vrel - (vrel . norm) norm (tangential); force_i += F / m
//         + m / rho_i^2 mu c V
// with c = wvisc_dd_coef.  (h - r) is the reference's Wvisc_lapl, negative
// between h and 2h.  bvel / bmass are ignored for the static form.
template <typename T>
static void rigid_force_impl(int num, const T* pos_s, const T* veleval_s, const T* density, T* force,
                             const float* bpos_s, const float* bvel_s, const float* bmass_s, const u32* bstart,
                             const u32* bend, const OSPHParams* sp, const OGridParams* gps, float s0, float s1,
                             int dynamic) {
    const OSPHParams p = *sp;
    const OGridParams gs = *gps;
    const T scale = p.simulation_scale, h = p.smoothing_distance, eps = p.EPSILON, m = p.mass;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < num; ++i) {
        V4<T> pi = ld(pos_s, i);
        pi.x *= scale; pi.y *= scale; pi.z *= scale; pi.w *= scale;
        const V4<T> ei = ld(veleval_s, i);
        T fx = 0, fy = 0, fz = 0, vx = 0, vy = 0, vz = 0;
        int cx, cy, cz;
        grid_cell(pi, gs, cx, cy, cz);
        for (int z = cz - 1; z <= cz + 1; ++z)
            for (int y = cy - 1; y <= cy + 1; ++y)
                for (int x = cx - 1; x <= cx + 1; ++x) {
                    u32 ch = (u32)grid_hash<T>(x, y, z, gs);
                    if (ch >= (u32)gs.nb_cells) continue;
                    u32 st = bstart[ch];
                    if (st == 0xFFFFFFFFu) continue;
                    for (u32 j = st; j < bend[ch]; ++j) {
                        T rx = (T)bpos_s[4 * j] * scale - pi.x, ry = (T)bpos_s[4 * j + 1] * scale - pi.y,
                          rz = (T)bpos_s[4 * j + 2] * scale - pi.z;
                        T rlen = len3(rx, ry, rz);
                        if (!(rlen <= (T)2 * h)) continue;
                        rlen = std::max(rlen, eps);
                        T nx = rx / rlen, ny = ry / rlen, nz = rz / rlen;
                        T mj = dynamic ? (T)bmass_s[j] : m;
                        T mn = (m * mj) / (m + mj);
                        T stiff = (T)s0 * m;
                        T sp2 = -stiff * ((T)2 * h - rlen);
                        T ux = -ei.x, uy = -ei.y, uz = -ei.z;
                        if (dynamic) { ux += (T)bvel_s[4 * j]; uy += (T)bvel_s[4 * j + 1]; uz += (T)bvel_s[4 * j + 2]; }
                        T un = ux * nx + uy * ny + uz * nz;
                        T dmp = (T)s1 * std::sqrt(stiff * mn);
                        fx += sp2 * nx + dmp * (un * nx);
                        fy += sp2 * ny + dmp * (un * ny);
                        fz += sp2 * nz + dmp * (un * nz);
                        T lap = h - rlen;
                        if (dynamic) { ux -= un * nx; uy -= un * ny; uz -= un * nz; }
                        vx += ux * lap; vy += uy * lap; vz += uz * lap;
                    }
                }
        V4<T> f = ld(force, i);
        const T di = density[i];
        const T c = m * ((T)1 / (di * di)) * ((T)p.viscosity * (T)p.wvisc_dd_coef);
        const T sgn = dynamic ? (T)1 : (T)-1;
        f.x += fx / m + sgn * (c * vx);
        f.y += fy / m + sgn * (c * vy);
        f.z += fz / m + sgn * (c * vz);
        st(force, i, f);
    }
}
extern "C" void oracle_rigid_force_f32(int num, const float* pos_s, const float* veleval_s, const float* density,
                                       float* force, const float* bpos_s, const float* bvel_s, const float* bmass_s,
                                       const u32* bstart, const u32* bend, const OSPHParams* sp,
                                       const OGridParams* gps, float s0, float s1, int dynamic) {
    rigid_force_impl<float>(num, pos_s, veleval_s, density, force, bpos_s, bvel_s, bmass_s, bstart, bend, sp, gps,
                            s0, s1, dynamic);
}
extern "C" void oracle_rigid_force_f64(int num, const double* pos_s, const double* veleval_s, const double* density,
                                       double* force, const float* bpos_s, const float* bvel_s, const float* bmass_s,
                                       const u32* bstart, const u32* bend, const OSPHParams* sp,
                                       const OGridParams* gps, float s0, float s1, int dynamic) {
    rigid_force_impl<double>(num, pos_s, veleval_s, density, force, bpos_s, bvel_s, bmass_s, bstart, bend, sp, gps,
                             s0, s1, dynamic);
}

// Rigid-body particle sets coupled into oracle steps (set with
// oracle_set_rigid; kind 0 = dynamic, 1 = static, applied in that order as
// SPH::interact does, after the walls and before the integration).  Each set
// is hashed and sorted into the fluid's grid by the step, as the rigid-body
// system's own hash_and_sort / cellindices would (System.cpp:145-170).
namespace {
struct RigidSet {
    std::vector<float> pos, vel, mass;
    float s0 = 0, s1 = 0;
    int n = 0;
};
RigidSet g_rigid[2];
}  // namespace
extern "C" void oracle_set_rigid(int kind, int n, const float* pos, const float* vel, const float* mass, float s0,
                                 float s1) {
    RigidSet& r = g_rigid[kind ? 1 : 0];
    r.n = n > 0 ? n : 0;
    r.pos.assign(pos ? pos : (const float*)nullptr, pos ? pos + 4 * (size_t)r.n : nullptr);
    r.vel.assign((size_t)4 * r.n, 0.f);
    r.mass.assign((size_t)r.n, 0.f);
    if (vel) std::copy(vel, vel + 4 * (size_t)r.n, r.vel.begin());
    if (mass) std::copy(mass, mass + (size_t)r.n, r.mass.begin());
    r.s0 = s0;
    r.s1 = s1;
}

// Both sets' forces on the step's sorted fluid state (SPH::interact).
template <typename T>
static void interact_impl(int num, const T* pos_s, const T* veleval_s, const T* density, T* force,
                          const OSPHParams* sp, const OGridParams* gp, const OGridParams* gps) {
    for (int kind = 0; kind < 2; ++kind) {
        const RigidSet& r = g_rigid[kind];
        if (r.n == 0) continue;
        std::vector<u32> keys(r.n), idx(r.n), start(gp->nb_cells + 1), end(gp->nb_cells + 1, 666u);
        hash_impl<float>(r.n, r.pos.data(), gp, keys.data(), idx.data());
        oracle_sort_stable(r.n, keys.data(), idx.data());
        oracle_cellindices(r.n, keys.data(), gp->nb_cells, start.data(), end.data());
        std::vector<float> bp(4 * (size_t)r.n), bv(4 * (size_t)r.n), bm((size_t)r.n);
        for (int q = 0; q < r.n; ++q) {
            for (int c = 0; c < 4; ++c) {
                bp[4 * (size_t)q + c] = r.pos[4 * (size_t)idx[q] + c];
                bv[4 * (size_t)q + c] = r.vel[4 * (size_t)idx[q] + c];
            }
            bm[q] = r.mass[idx[q]];
        }
        rigid_force_impl<T>(num, pos_s, veleval_s, density, force, bp.data(), bv.data(), bm.data(), start.data(),
                            end.data(), sp, gps, r.s0, r.s1, kind == 0);
    }
}

// ---------------------------------------------------------------------------
// a11: leapfrog.cl:30-121, scattering to the original index.
template <typename T>
static void leapfrog_impl(int num, const T* pos_s, const T* vel_s, const T* force, const T* xsph, const u32* idx,
                          const OSPHParams* sp, T dt, T* pos_u, T* vel_u, T* veleval_u) {
    const OSPHParams p = *sp;
    const T scale = p.simulation_scale;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < num; ++i) {
        V4<T> x = ld(pos_s, i);
        x.x *= scale; x.y *= scale; x.z *= scale; x.w *= scale;
        V4<T> v = ld(vel_s, i);
        V4<T> f = ld(force, i);
        f.x += (T)p.gravity.x; f.y += (T)p.gravity.y; f.z += (T)p.gravity.z;
        f.w = 0;
        T speed = len3(f.x, f.y, f.z);
        if (speed > (T)p.velocity_limit) {
            T s = (T)p.velocity_limit / speed;
            f.x *= s; f.y *= s; f.z *= s;
        }
        V4<T> vn{v.x + dt * f.x, v.y + dt * f.y, v.z + dt * f.z, v.w + dt * f.w};
        V4<T> xs = ld(xsph, i);
        T xf = p.xsph_factor;
        vn.x += xf * xs.x; vn.y += xf * xs.y; vn.z += xf * xs.z; vn.w += xf * xs.w;
        V4<T> ve{(T)0.5 * (v.x + vn.x), (T)0.5 * (v.y + vn.y), (T)0.5 * (v.z + vn.z), (T)0.5 * (v.w + vn.w)};
        x.x += dt * vn.x; x.y += dt * vn.y; x.z += dt * vn.z;
        long o = idx[i];
        V4<T> out{x.x / scale, x.y / scale, x.z / scale, (T)1};
        st(vel_u, o, vn);
        st(veleval_u, o, ve);
        st(pos_u, o, out);
    }
}
extern "C" void oracle_leapfrog_f32(int num, const float* pos_s, const float* vel_s, const float* force, const float* xsph,
                         const u32* idx, const OSPHParams* sp, float dt, float* pos_u, float* vel_u, float* veleval_u) {
    leapfrog_impl<float>(num, pos_s, vel_s, force, xsph, idx, sp, dt, pos_u, vel_u, veleval_u);
}
extern "C" void oracle_leapfrog_f64(int num, const double* pos_s, const double* vel_s, const double* force, const double* xsph,
                         const u32* idx, const OSPHParams* sp, double dt, double* pos_u, double* vel_u,
                         double* veleval_u) {
    leapfrog_impl<double>(num, pos_s, vel_s, force, xsph, idx, sp, dt, pos_u, vel_u, veleval_u);
}

// ---------------------------------------------------------------------------
// a11 (alternative): euler.cl:33-78, selected by settings "integrator" =
// "euler" (SPH.cpp:89-95, 392-409).  Its quirks are restated as they are:
// gravity is a hard-coded -9.8 on z (the settings' gravity is ignored), the
// clamp is a hard-coded 600 on |f.xyz| (velocity_limit is ignored) and scales
// all four components, f.w is not cleared, there is no XSPH term and
// veleval is not written (it keeps what it held).  Scatters to idx[i].
template <typename T>
static void euler_impl(int num, const T* pos_s, const T* vel_s, const T* force, const u32* idx, const OSPHParams* sp,
                       T dt, T* pos_u, T* vel_u) {
    const T scale = sp->simulation_scale;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < num; ++i) {
        V4<T> x = ld(pos_s, i);
        x.x *= scale; x.y *= scale; x.z *= scale; x.w *= scale;
        V4<T> v = ld(vel_s, i);
        V4<T> f = ld(force, i);
        f.z += (T)-9.8f;
        T speed = len3(f.x, f.y, f.z);
        if (speed > (T)600.0f) {
            T s = (T)600.0f / speed;
            f.x *= s; f.y *= s; f.z *= s; f.w *= s;
        }
        v.x += dt * f.x; v.y += dt * f.y; v.z += dt * f.z; v.w += dt * f.w;
        x.x += dt * v.x; x.y += dt * v.y; x.z += dt * v.z;
        x.x /= scale; x.y /= scale; x.z /= scale;
        long o = idx[i];
        st(vel_u, o, v);
        st(pos_u, o, V4<T>{x.x, x.y, x.z, (T)1});
    }
}
extern "C" void oracle_euler_f32(int num, const float* pos_s, const float* vel_s, const float* force, const u32* idx,
                                 const OSPHParams* sp, float dt, float* pos_u, float* vel_u) {
    euler_impl<float>(num, pos_s, vel_s, force, idx, sp, dt, pos_u, vel_u);
}
extern "C" void oracle_euler_f64(int num, const double* pos_s, const double* vel_s, const double* force,
                                 const u32* idx, const OSPHParams* sp, double dt, double* pos_u, double* vel_u) {
    euler_impl<double>(num, pos_s, vel_s, force, idx, sp, dt, pos_u, vel_u);
}

// ---------------------------------------------------------------------------
// Candidate (K) and accepted (M) pair counts of the neighbour loop, summed
// over all particles, on the actual state.  Used for the FLOP model.
extern "C" void oracle_count_pairs(int num, const float* pos_s, const u32* start, const u32* end, const OSPHParams* sp,
                        const OGridParams* gps, long long* K_total, long long* M_total) {
    const OSPHParams p = *sp;
    const OGridParams gs = *gps;
    const float scale = p.simulation_scale, h = p.smoothing_distance;
    long long K = 0, M = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : K, M)
    for (int i = 0; i < num; ++i) {
        V4<float> pi = ld(pos_s, i);
        pi.x *= scale; pi.y *= scale; pi.z *= scale; pi.w *= scale;
        for_neighbours<float>(pi, pos_s, scale, start, end, gs, [&](u32, const V4<float>& pj) {
            float rx = pi.x - pj.x, ry = pi.y - pj.y, rz = pi.z - pj.z;
            ++K;
            if (len3(rx, ry, rz) <= h) ++M;
        });
    }
    *K_total = K;
    *M_total = M;
}

// ---------------------------------------------------------------------------
// Wall-clock seconds of each stage of the last step (steady_clock): hash,
// sort, cellindices, permute, density, force, gravity + walls, leapfrog (the
// deletion path's second hash and sort count into theirs).
static double g_stage_s[8];
using OClock = std::chrono::steady_clock;
static double secs_since(OClock::time_point& t) {
    const OClock::time_point now = OClock::now();
    const double d = std::chrono::duration<double>(now - t).count();
    t = now;
    return d;
}

// One full step (SPH::updateGPU + SPH::integrate, SPH.cpp:149-300, 389-441).
// flags bit0: wall collision on (the shipped code has the call commented out
//             at SPH.cpp:297; the dam-break workloads enable it);
//       bit1: use the bitonic network instead of the stable sort;
//       bit2: integrate with euler.cl instead of leapfrog.cl (SPH::integrate,
//             SPH.cpp:392-409).
// State arrays are (max_num x 4) in unsorted (original index) order.  The
// scratch arrays are filled with the step's sorted intermediates.  *num may
// shrink through the out-of-grid deletion path (SPH.cpp:220-250).
template <typename T>
static int step_impl(int* num_io, T* pos_u, T* vel_u, T* veleval_u, const OSPHParams* sp, const OGridParams* gp,
                     const OGridParams* gps, T dt, int flags, int nsrc, const float* src, const float* smass,
                     const float* alpha, u32* keys, u32* idx, u32* start, u32* end, T* pos_s, T* vel_s,
                     T* veleval_s, T* density, T* force, T* xsph) {
    int num = *num_io;
    for (double& v : g_stage_s) v = 0.0;
    OClock::time_point t = OClock::now();
    hash_impl<T>(num, pos_u, gp, keys, idx);
    g_stage_s[0] += secs_since(t);
    if (flags & 2) oracle_sort_bitonic(num, keys, idx); else oracle_sort_stable(num, keys, idx);
    g_stage_s[1] += secs_since(t);
    int nc = oracle_cellindices(num, keys, gp->nb_cells, start, end);
    g_stage_s[2] += secs_since(t);
    permute_impl<T>(num, idx, pos_u, vel_u, veleval_u, pos_s, vel_s, veleval_s);
    g_stage_s[3] += secs_since(t);
    int deleted = 0;
    if (nc <= num && nc >= 0) {
        deleted = num - nc;
        num = nc;
        // call_prep (SPH.cpp:445-457): sorted -> unsorted for the survivors
        std::memcpy(pos_u, pos_s, sizeof(T) * 4 * (size_t)num);
        std::memcpy(vel_u, vel_s, sizeof(T) * 4 * (size_t)num);
        std::memcpy(veleval_u, veleval_s, sizeof(T) * 4 * (size_t)num);
        (void)secs_since(t);
        hash_impl<T>(num, pos_u, gp, keys, idx);
        g_stage_s[0] += secs_since(t);
        if (flags & 2) oracle_sort_bitonic(num, keys, idx); else oracle_sort_stable(num, keys, idx);
        g_stage_s[1] += secs_since(t);
    }
    OSPHParams p = *sp;
    p.num = num;
    (void)secs_since(t);
    density_impl<T>(num, pos_s, start, end, &p, gps, density);
    g_stage_s[4] += secs_since(t);
    force_impl<T>(num, pos_s, density, veleval_s, start, end, &p, gps, force, xsph);
    g_stage_s[5] += secs_since(t);
    if (nsrc > 0) gravity_impl<T>(num, nsrc, src, smass, alpha, pos_s, force, (T)p.simulation_scale);
    if (flags & 1) collision_impl<T>(num, pos_s, vel_s, force, &p, gps);
    interact_impl<T>(num, pos_s, veleval_s, density, force, &p, gp, gps);   // no sets: nothing
    g_stage_s[6] += secs_since(t);
    if (flags & 4) euler_impl<T>(num, pos_s, vel_s, force, idx, &p, dt, pos_u, vel_u);
    else leapfrog_impl<T>(num, pos_s, vel_s, force, xsph, idx, &p, dt, pos_u, vel_u, veleval_u);
    g_stage_s[7] += secs_since(t);
    *num_io = num;
    return deleted;
}
extern "C" int oracle_step_f32(int* num, float* pos_u, float* vel_u, float* veleval_u, const OSPHParams* sp,
                    const OGridParams* gp, const OGridParams* gps, float dt, int flags, int nsrc, const float* src,
                    const float* smass, const float* alpha, u32* keys, u32* idx, u32* start, u32* end, float* pos_s,
                    float* vel_s, float* veleval_s, float* density, float* force, float* xsph) {
    return step_impl<float>(num, pos_u, vel_u, veleval_u, sp, gp, gps, dt, flags, nsrc, src, smass, alpha, keys,
                            idx, start, end, pos_s, vel_s, veleval_s, density, force, xsph);
}
extern "C" int oracle_step_f64(int* num, double* pos_u, double* vel_u, double* veleval_u, const OSPHParams* sp,
                    const OGridParams* gp, const OGridParams* gps, double dt, int flags, int nsrc, const float* src,
                    const float* smass, const float* alpha, u32* keys, u32* idx, u32* start, u32* end,
                    double* pos_s, double* vel_s, double* veleval_s, double* density, double* force, double* xsph) {
    return step_impl<double>(num, pos_u, vel_u, veleval_u, sp, gp, gps, dt, flags, nsrc, src, smass, alpha, keys,
                             idx, start, end, pos_s, vel_s, veleval_s, density, force, xsph);
}

// Tangent-linear FP64 step (see TL above): state arrays hold (1 + kTL)
// doubles per component (value, then the derivative along each of the kTL
// input directions), i.e. double[max_num][4][1 + kTL]; dmu / dK: the
// viscosity's and gas constant's tangent per direction.  The values follow
// oracle_step_f64 exactly.
extern "C" int oracle_step_tl(int* num, double* pos_u, double* vel_u, double* veleval_u, const OSPHParams* sp,
                              const OGridParams* gp, const OGridParams* gps, double dt, int flags, int nsrc,
                              const float* src, const float* smass, const float* alpha, const double* dmu,
                              const double* dK, u32* keys, u32* idx, u32* start, u32* end, double* pos_s,
                              double* vel_s, double* veleval_s, double* density, double* force, double* xsph) {
    static_assert(sizeof(TL) == (1 + kTL) * sizeof(double), "TL is value + kTL derivatives");
    for (int k = 0; k < kTL; ++k) {
        g_tl_dmu[k] = dmu ? dmu[k] : 0.0;
        g_tl_dK[k] = dK ? dK[k] : 0.0;
    }
    const int d = step_impl<TL>(num, (TL*)pos_u, (TL*)vel_u, (TL*)veleval_u, sp, gp, gps, TL(dt), flags, nsrc, src,
                                smass, alpha, keys, idx, start, end, (TL*)pos_s, (TL*)vel_s, (TL*)veleval_s,
                                (TL*)density, (TL*)force, (TL*)xsph);
    for (int k = 0; k < kTL; ++k) g_tl_dmu[k] = g_tl_dK[k] = 0.0;
    return d;
}
extern "C" int oracle_tl_directions(void) { return kTL; }

// Per-stage seconds of the last oracle step (8 values, see g_stage_s).
extern "C" void oracle_last_stage_times(double* out) {
    for (int k = 0; k < 8; ++k) out[k] = g_stage_s[k];
}

extern "C" int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// OpenMP threads of the following oracle calls (n <= 0: all host cores).
extern "C" void oracle_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
#else
    (void)n;
#endif
}


// ---------------------------------------------------------------------------
// a13: addRect (rtpslib/domain/IV.cpp:165-202) as System::addBox calls it
// (System.cpp:331-338, scale 1): z outer, y, x inner, each coordinate
// accumulated in float (`x += spacing`) while <= the box's max; at most num
// points (the inner loop stops at num; the outer loops then add nothing).
// Writes up to num float4 (w = 1) and returns the count.
extern "C" int oracle_add_rect(int num, const float* bmin, const float* bmax, float spacing, float scale,
                               float* out) {
    const float xmin = bmin[0] / scale, xmax = bmax[0] / scale;
    const float ymin = bmin[1] / scale, ymax = bmax[1] / scale;
    const float zmin = bmin[2] / scale, zmax = bmax[2] / scale;
    int i = 0;
    for (float z = zmin; z <= zmax; z += spacing)
        for (float y = ymin; y <= ymax; y += spacing)
            for (float x = xmin; x <= xmax; x += spacing) {
                if (i >= num) break;
                out[4 * i] = x; out[4 * i + 1] = y; out[4 * i + 2] = z; out[4 * i + 3] = 1.0f;
                ++i;
            }
    return i;
}

// ---------------------------------------------------------------------------
// a1: parameter replay.  SPH::calculate (SPHSettings.cpp:30-144), the settings
// string round trip (RTPSSettings.h:65-85: ostream default precision, i.e.
// "%.6g", re-parsed with istream >> float), SPH::updateParams
// (SPHSettings.cpp:149-201) and System::setupDomain / Domain::calculateCells
// (System.cpp:294-329, Domain.cpp:38-81), all in FP32 as the host does it.
//
// Build extension (documented deviation): if "mass", "rest_distance",
// "smoothing_distance" or "simulation_scale" are present in the settings
// given at construction, those values are kept instead of derived.  The
// reference never has these keys before calculate() runs, so reference-style
// settings derive identically.
namespace {
struct Settings {
    std::map<std::string, std::string> kv;
    bool has(const std::string& k) const { return kv.count(k) != 0; }
    template <typename X> X get(const std::string& k, const char* dflt = "0") const {
        auto it = kv.find(k);
        std::istringstream is(it == kv.end() ? std::string(dflt) : it->second);
        X v{};
        is >> v;
        return v;
    }
    OF4 get4(const std::string& k) const {
        auto it = kv.find(k);
        std::istringstream is(it == kv.end() ? std::string("0") : it->second);
        OF4 v{0, 0, 0, 0};
        is >> v.x >> v.y >> v.z >> v.w;
        return v;
    }
    template <typename X> void set(const std::string& k, X v) {
        std::ostringstream os;
        os << v;
        kv[k] = os.str();
    }
};
}  // namespace

extern "C" int oracle_derive_params(const char* const* keys, const char* const* vals, int n, OSPHParams* out_sphp,
                                    OGridParams* out_gp, OGridParams* out_gps) {
    Settings s;
    for (int i = 0; i < n; ++i) s.kv[keys[i]] = vals[i];
    const bool o_mass = s.has("mass"), o_rd = s.has("rest_distance"), o_h = s.has("smoothing_distance"),
               o_sc = s.has("simulation_scale");

    OF4 bmin = s.get4("domain_min"), bmax = s.get4("domain_max");
    unsigned max_num = s.get<unsigned>("max_num_particles");

    // SPH::calculate
    float rho0 = s.get<float>("rest_density");
    float mass = o_mass ? s.get<float>("mass")
                        : (float)(0.0256 / (int)(std::log((float)max_num) / std::log(2.0f)));  // float log overloads
    float VP = mass / rho0;
    float rest_distance = o_rd ? s.get<float>("rest_distance") : (float)(.87 * std::pow(mass / rho0, 1.f / 3.f));
    float smoothing = o_h ? s.get<float>("smoothing_distance") : 2.0f * rest_distance;
    float domain_vol = ((bmax.x - bmin.x) * (bmax.y - bmin.y) * (bmax.z - bmin.z)) / 2;
    float scale = o_sc ? s.get<float>("simulation_scale")
                       : std::pow(.5f * VP * max_num / domain_vol, 1.f / 3.f);
    if (!o_mass) s.set("mass", mass);
    if (!o_rd) s.set("rest_distance", rest_distance);
    if (!o_h) s.set("smoothing_distance", smoothing);
    if (!o_sc) s.set("simulation_scale", scale);
    float boundary_distance = .5f * rest_distance;
    s.set("boundary_distance", boundary_distance);
    s.set("spacing", rest_distance / scale);
    float pi = (float)M_PI;
    float h9 = std::pow(smoothing, 9.f), h6 = std::pow(smoothing, 6.f), h3 = std::pow(smoothing, 3.f);
    s.set("wpoly6", 315.f / 64.0f / pi / h9);
    s.set("wpoly6_d", -945.f / (32.0f * pi * h9));
    s.set("wpoly6_dd", -945.f / (32.0f * pi * h9));
    s.set("wspiky", 15.f / pi / h6);
    s.set("wspiky_d", -45.f / (pi * h6));
    s.set("wspiky_dd", 15. / (2. * pi * h3));
    s.set("wvisc", 15. / (2. * pi * h3));
    s.set("wvisc_d", 15. / (2. * pi * h3));
    s.set("wvisc_dd", 45. / (pi * h6));
    s.set("epsilon", 1E-6);
    s.set("num_particles", 0);

    // SPH::updateParams
    OSPHParams p;
    std::memset(&p, 0, sizeof(p));
    p.rest_density = s.get<float>("rest_density");
    p.mass = s.get<float>("mass");
    p.rest_distance = s.get<float>("rest_distance");
    p.smoothing_distance = s.get<float>("smoothing_distance");
    p.simulation_scale = s.get<float>("simulation_scale");
    p.boundary_stiffness = s.get<float>("boundary_stiffness");
    p.boundary_dampening = s.get<float>("boundary_dampening");
    p.boundary_distance = s.get<float>("boundary_distance");
    p.K = s.get<float>("gas_constant");
    p.viscosity = s.get<float>("viscosity");
    p.velocity_limit = s.get<float>("velocity_limit");
    p.xsph_factor = s.get<float>("xsph_factor");
    p.gravity = s.get4("gravity");
    p.friction_coef = s.get<float>("friction_kinetic");
    p.restitution_coef = s.get<float>("restitution");
    p.shear = s.get<float>("shear");
    p.attraction = s.get<float>("attraction");
    p.spring = s.get<float>("spring");
    p.EPSILON = s.get<float>("epsilon");
    p.wpoly6_coef = s.get<float>("wpoly6");
    p.wpoly6_d_coef = s.get<float>("wpoly6_d");
    p.wpoly6_dd_coef = s.get<float>("wpoly6_dd");
    p.wspiky_coef = s.get<float>("wspiky");
    p.wspiky_d_coef = s.get<float>("wspiky_d");
    p.wspiky_dd_coef = s.get<float>("wspiky_dd");
    p.wvisc_coef = s.get<float>("wvisc");
    p.wvisc_d_coef = s.get<float>("wvisc_d");
    p.wvisc_dd_coef = s.get<float>("wvisc_dd");
    p.num = s.get<int>("num_particles");
    p.max_num = s.get<int>("max_num_particles");

    // System::setupDomain(h/scale, scale) -> Domain::calculateCells
    float cell = p.smoothing_distance / p.simulation_scale;
    double s2 = 2. * cell;
    OF4 mn{bmin.x - (float)s2, bmin.y - (float)s2, bmin.z - (float)s2, bmin.w - 0.f};
    OF4 mx{bmax.x + (float)s2, bmax.y + (float)s2, bmax.z + (float)s2, bmax.w + 0.f};
    OF4 size{mx.x - mn.x, mx.y - mn.y, mx.z - mn.z, 0.f};
    OF4 res{std::ceil(size.x / cell), std::ceil(size.y / cell), std::ceil(size.z / cell), 0.f};
    size = OF4{res.x * cell, res.y * cell, res.z * cell, 0.f};
    OF4 delta{res.x / size.x, res.y / size.y, res.z / size.z, 0.f};
    OF4 gmax{mn.x + size.x, mn.y + size.y, mn.z + size.z, mn.w + size.w};
    OGridParams g;
    std::memset(&g, 0, sizeof(g));
    g.grid_min = mn; g.grid_max = gmax; g.bnd_min = bmin; g.bnd_max = bmax;
    g.grid_res = res; g.grid_size = size; g.grid_delta = delta;
    g.nb_cells = (int)(res.x * res.y * res.z);
    // scaled twin (System.cpp:318-325); float4 * float and float4 / float
    // (structs.h: the latter multiplies by 1/r)
    auto mul = [](OF4 a, float r) { return OF4{r * a.x, r * a.y, r * a.z, r * a.w}; };
    OGridParams gs = g;
    float sc = p.simulation_scale;
    gs.grid_min = mul(g.grid_min, sc); gs.grid_max = mul(g.grid_max, sc);
    gs.bnd_min = mul(g.bnd_min, sc); gs.bnd_max = mul(g.bnd_max, sc);
    gs.grid_size = mul(g.grid_size, sc);
    gs.grid_delta = mul(g.grid_delta, 1.f / sc);
    *out_sphp = p;
    *out_gp = g;
    *out_gps = gs;
    return 0;
}
