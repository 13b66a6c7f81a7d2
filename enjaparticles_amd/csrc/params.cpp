// Host-side parameter derivation of the SPH system.
//
// SPH::calculate (rtpslib/system/SPHSettings.cpp:30-144) writes the derived
// quantities into the settings map, SPH::updateParams (:149-201) reads every
// SPHParams field back (so derived floats carry the map's 6-significant-digit
// rounding, rtpslib/RTPSSettings.h:78-85), and System::setupDomain
// (rtpslib/system/System.cpp:294-329) builds the unscaled and scaled
// GridParams through Domain::calculateCells (rtpslib/domain/Domain.cpp:38-81).
// All arithmetic is FP32 where the reference's is.
//
// Extension: when "mass", "rest_distance", "smoothing_distance" or
// "simulation_scale" are already present in the settings, they are used as
// given instead of being derived from max_num_particles (the reference's
// settings never contain them before calculate() runs, so reference-style
// settings derive identically).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rtps_internal.h"
#include "sph_device.h"

namespace rtps {

void read_sph_params(const RTPSSettings& s, rtps_sph_params& p) {
    std::memset(&p, 0, sizeof(p));
    p.rest_density = s.GetSettingAs<float>("rest_density");
    p.mass = s.GetSettingAs<float>("mass");
    p.rest_distance = s.GetSettingAs<float>("rest_distance");
    p.smoothing_distance = s.GetSettingAs<float>("smoothing_distance");
    p.simulation_scale = s.GetSettingAs<float>("simulation_scale");
    p.boundary_stiffness = s.GetSettingAs<float>("boundary_stiffness");
    p.boundary_dampening = s.GetSettingAs<float>("boundary_dampening");
    p.boundary_distance = s.GetSettingAs<float>("boundary_distance");
    p.K = s.GetSettingAs<float>("gas_constant");
    p.viscosity = s.GetSettingAs<float>("viscosity");
    p.velocity_limit = s.GetSettingAs<float>("velocity_limit");
    p.xsph_factor = s.GetSettingAs<float>("xsph_factor");
    p.gravity = s.GetSettingAs<rtps_float4>("gravity");
    p.friction_coef = s.GetSettingAs<float>("friction_kinetic");
    p.restitution_coef = s.GetSettingAs<float>("restitution");
    p.shear = s.GetSettingAs<float>("shear");
    p.attraction = s.GetSettingAs<float>("attraction");
    p.spring = s.GetSettingAs<float>("spring");
    p.EPSILON = s.GetSettingAs<float>("epsilon");
    p.wpoly6_coef = s.GetSettingAs<float>("wpoly6");
    p.wpoly6_d_coef = s.GetSettingAs<float>("wpoly6_d");
    p.wpoly6_dd_coef = s.GetSettingAs<float>("wpoly6_dd");
    p.wspiky_coef = s.GetSettingAs<float>("wspiky");
    p.wspiky_d_coef = s.GetSettingAs<float>("wspiky_d");
    p.wspiky_dd_coef = s.GetSettingAs<float>("wspiky_dd");
    p.wvisc_coef = s.GetSettingAs<float>("wvisc");
    p.wvisc_d_coef = s.GetSettingAs<float>("wvisc_d");
    p.wvisc_dd_coef = s.GetSettingAs<float>("wvisc_dd");
    p.num = s.GetSettingAs<int>("num_particles");
    p.max_num = s.GetSettingAs<int>("max_num_particles");
}

static void calculate(RTPSSettings& s) {
    const bool has_mass = s.Exists("mass"), has_rd = s.Exists("rest_distance"),
               has_h = s.Exists("smoothing_distance"), has_scale = s.Exists("simulation_scale");
    const unsigned max_num = s.GetSettingAs<unsigned>("max_num_particles");
    const rtps_float4 dmin = s.GetSettingAs<rtps_float4>("domain_min");
    const rtps_float4 dmax = s.GetSettingAs<rtps_float4>("domain_max");

    const float rho0 = s.GetSettingAs<float>("rest_density");
    const float mass = has_mass ? s.GetSettingAs<float>("mass")
                                : (float)(0.0256 / (int)(std::log((float)max_num) / std::log(2.0f)));
    const float VP = mass / rho0;
    const float rest_distance =
        has_rd ? s.GetSettingAs<float>("rest_distance") : (float)(.87 * std::pow(mass / rho0, 1.f / 3.f));
    const float smoothing = has_h ? s.GetSettingAs<float>("smoothing_distance") : 2.0f * rest_distance;
    const float domain_vol = ((dmax.x - dmin.x) * (dmax.y - dmin.y) * (dmax.z - dmin.z)) / 2;
    const float scale = has_scale ? s.GetSettingAs<float>("simulation_scale")
                                  : std::pow(.5f * VP * max_num / domain_vol, 1.f / 3.f);
    if (!has_mass) s.SetSetting("mass", mass);
    if (!has_rd) s.SetSetting("rest_distance", rest_distance);
    if (!has_h) s.SetSetting("smoothing_distance", smoothing);
    if (!has_scale) s.SetSetting("simulation_scale", scale);

    s.SetSetting("boundary_distance", .5f * rest_distance);
    s.SetSetting("spacing", rest_distance / scale);

    const float pi = (float)M_PI;
    const float h9 = std::pow(smoothing, 9.f), h6 = std::pow(smoothing, 6.f), h3 = std::pow(smoothing, 3.f);
    s.SetSetting("wpoly6", 315.f / 64.0f / pi / h9);
    s.SetSetting("wpoly6_d", -945.f / (32.0f * pi * h9));
    s.SetSetting("wpoly6_dd", -945.f / (32.0f * pi * h9));
    s.SetSetting("wspiky", 15.f / pi / h6);
    s.SetSetting("wspiky_d", -45.f / (pi * h6));
    s.SetSetting("wspiky_dd", 15. / (2. * pi * h3));   // double expressions, as in the reference
    s.SetSetting("wvisc", 15. / (2. * pi * h3));
    s.SetSetting("wvisc_d", 15. / (2. * pi * h3));
    s.SetSetting("wvisc_dd", 45. / (pi * h6));
    s.SetSetting("epsilon", 1E-6);
    s.SetSetting("num_particles", 0);
}

// Domain::calculateCells + System::setupDomain.
static void setup_domain(const RTPSSettings& s, float cell, float sim_scale, rtps_grid_params& g,
                         rtps_grid_params& gs) {
    const rtps_float4 bmin = s.GetSettingAs<rtps_float4>("domain_min");
    const rtps_float4 bmax = s.GetSettingAs<rtps_float4>("domain_max");
    const double s2 = 2. * cell;
    const float f2 = (float)s2;
    const rtps_float4 mn{bmin.x - f2, bmin.y - f2, bmin.z - f2, bmin.w - 0.f};
    const rtps_float4 mx{bmax.x + f2, bmax.y + f2, bmax.z + f2, bmax.w + 0.f};
    rtps_float4 size{mx.x - mn.x, mx.y - mn.y, mx.z - mn.z, 0.f};
    const rtps_float4 res{std::ceil(size.x / cell), std::ceil(size.y / cell), std::ceil(size.z / cell), 0.f};
    size = rtps_float4{res.x * cell, res.y * cell, res.z * cell, 0.f};
    const rtps_float4 delta{res.x / size.x, res.y / size.y, res.z / size.z, 0.f};
    std::memset(&g, 0, sizeof(g));
    g.grid_min = mn;
    g.grid_max = rtps_float4{size.x + mn.x, size.y + mn.y, size.z + mn.z, size.w + mn.w};
    g.bnd_min = bmin;
    g.bnd_max = bmax;
    g.grid_res = res;
    g.grid_size = size;
    g.grid_delta = delta;
    g.nb_cells = (int)(res.x * res.y * res.z);
    // scaled twin: float4 * float, float4 / float = multiply by 1/r (structs.h)
    auto mul = [](const rtps_float4& a, float r) { return rtps_float4{r * a.x, r * a.y, r * a.z, r * a.w}; };
    gs = g;
    gs.grid_min = mul(g.grid_min, sim_scale);
    gs.grid_max = mul(g.grid_max, sim_scale);
    gs.bnd_min = mul(g.bnd_min, sim_scale);
    gs.bnd_max = mul(g.bnd_max, sim_scale);
    gs.grid_size = mul(g.grid_size, sim_scale);
    gs.grid_delta = mul(g.grid_delta, 1.f / sim_scale);
}

void derive_params(RTPSSettings& s, rtps_sph_params& sp, rtps_grid_params& gp, rtps_grid_params& gps) {
    calculate(s);
    read_sph_params(s, sp);
    setup_domain(s, sp.smoothing_distance / sp.simulation_scale, sp.simulation_scale, gp, gps);
}

int sort_key_bits(int nb_cells) {
    // keys are in [0, nb_cells] plus the pad key (1 << bits) - 1 > nb_cells
    unsigned v = (unsigned)nb_cells + 1u;
    int bits = 0;
    while ((1u << bits) <= v && bits < 32) ++bits;
    return bits;
}

// Point sources of gravity.cl (settings gravity_sources / gravity_mass /
// gravity_alphas, System.cpp:261-265), at most kMaxGravitySources.  With
// every alpha * mass zero the sources add exactly zero (gravity.cl:25-34), so
// n is 0 and the kernels skip them.
void gravity_sources_of(const RTPSSettings& st, GravitySources& g) {
    std::vector<rtps_float4> src = st.GetSettingAs<std::vector<rtps_float4>>("gravity_sources");
    std::vector<float> mass = st.GetSettingAs<std::vector<float>>("gravity_mass");
    std::vector<float> alpha = st.GetSettingAs<std::vector<float>>("gravity_alphas");
    std::memset(&g, 0, sizeof(g));
    const int n = (int)std::min(src.size(), (size_t)kMaxGravitySources);
    bool any = false;
    for (int i = 0; i < n; ++i) {
        g.pos[i] = make_float4(src[i].x, src[i].y, src[i].z, src[i].w);
        g.mass[i] = i < (int)mass.size() ? mass[i] : 0.f;
        g.alpha[i] = i < (int)alpha.size() ? alpha[i] : 0.f;
        any = any || g.alpha[i] * g.mass[i] != 0.f;
    }
    g.n = any ? n : 0;
}

// Settings "integrator" (SPH.cpp:89-95, 392-409): 0 = "leapfrog" (also when
// the key is absent, as in no shipped settings file), 1 = "euler", -1 = any
// other value (the reference then creates neither integrator and integrate()
// does nothing; the library rejects it).
int integrator_of(const RTPSSettings& st) {
    if (!st.Exists("integrator")) return 0;
    const std::string v = st.GetSettingAs<std::string>("integrator");
    return v == "leapfrog" ? 0 : (v == "euler" ? 1 : -1);
}

}  // namespace rtps
