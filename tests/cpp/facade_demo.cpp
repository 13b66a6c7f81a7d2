// Drives the SPH step through the C++ facade exactly as an rtpslib
// application does (settings -> generateSystemInstance -> addBox -> frame loop),
// then prints the particle count and the mean density.  Used by the tests: it
// is compiled and linked on the CPU and run on the GPU.
#include <cstdio>

#include "rtps/SPH.h"

int main() {
    rtps::RTPSSettings* s = new rtps::RTPSSettings();
    s->SetSetting<std::string>("system", "sph");
    s->SetSetting("max_num_particles", 4096);
    s->SetSetting<std::string>("domain_min", "0 0 0 0");
    s->SetSetting<std::string>("domain_max", "1 1 1 0");
    s->SetSetting("rest_density", 1000.0f);
    s->SetSetting("gas_constant", 3.0f);
    s->SetSetting("viscosity", 0.01f);
    s->SetSetting("velocity_limit", 600.0f);
    s->SetSetting("xsph_factor", 0.15f);
    s->SetSetting("boundary_stiffness", 20000.0f);
    s->SetSetting("boundary_dampening", 256.0f);
    s->SetSetting<std::string>("gravity", "0 0 -9.8 0");
    s->SetSetting<std::string>("integrator", "leapfrog");
    s->SetSetting("time_step", 0.003f);
    s->SetSetting("mass", 0.008f);
    s->SetSetting("rest_distance", 0.02f);
    s->SetSetting("smoothing_distance", 0.04f);
    s->SetSetting("simulation_scale", 1.0f);
    try {
        rtps::System* sys = rtps::RTPS::generateSystemInstance(s);
        int n = sys->addBox(4096, rtps::float4{0.01f, 0.01f, 0.01f, 1}, rtps::float4{0.31f, 0.31f, 0.31f, 1}, false);
        for (int frame = 0; frame < 10; ++frame) {
            sys->update();
            sys->interact();
            sys->integrate();
            sys->postProcess();
        }
        rtps::SPH* sph = static_cast<rtps::SPH*>(sys);
        std::vector<float> d = sph->getDensityBuffer();
        double mean = 0;
        for (float x : d) mean += x;
        mean /= d.empty() ? 1 : d.size();
        std::printf("particles %d num %d mean_density %.3f\n", n, sys->getNum(), mean);
        delete sys;
    } catch (const rtps::Error& e) {
        std::printf("rtps error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
