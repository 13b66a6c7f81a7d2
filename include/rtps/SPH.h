// rtps::SPH / rtps::RTPS -- C++ host interface of the MI355X SPH step.
//
// Header-only facade over the C ABI (include/rtps_c.h) with the reference's
// class and method names (rtpslib/RTPS.h:42-46, rtpslib/system/System.h:61-271,
// rtpslib/system/SPH.h:56-155), so rtpslib callers switch by relinking:
//
//   rtps::RTPSSettings* s = new rtps::RTPSSettings();  s->SetSetting("system", "sph"); ...
//   rtps::System* sys = rtps::RTPS::generateSystemInstance(s);   // takes ownership of s
//   sys->addBox(n, min, max, false);
//   for (;;) { sys->update(); sys->interact(); sys->integrate(); sys->postProcess(); }
//
// Differences: there is no CL context argument (the system owns one HIP
// stream), buffers are returned as host copies (getPositionBufferUnsorted()
// etc.) or as device pointers (devicePtr), and errors throw rtps::Error
// instead of being printed and swallowed (rtpslib/opencl/Kernel.cpp:68-72).
#pragma once

#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../rtps_c.h"
#include "RTPSSettings.h"

namespace rtps {

typedef rtps_float4 float4;
typedef rtps_triangle Triangle;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc < 0) throw Error(rc, rtps_last_error());
}

// Iso-surface of the colour field (rtpslib's Mesh, render/util/Mesh.h): three
// vertices of 3 floats per triangle, with per-vertex normals.
struct Mesh {
    std::vector<float> vertices, normals;
    int ntris = 0;
};

class System {
public:
    virtual ~System() {}
    virtual void update() = 0;
    virtual void interact() {}
    virtual void integrate() = 0;
    virtual void postProcess() {}
    virtual int getNum() = 0;
    virtual int addBox(int nn, float4 min, float4 max, bool scaled, float4 color = float4{1, 0, 0, 1},
                       float mass = 0.0f) = 0;
    virtual void pushParticles(const std::vector<float4>& pos, const std::vector<float4>& vel,
                               float4 color = float4{1, 0, 0, 1}, float mass = 0.0f) = 0;
    virtual RTPSSettings* getSettings() = 0;
};

class SPH : public System {
public:
    // Takes ownership of `settings` (the reference deletes it, System.cpp:120).
    explicit SPH(RTPSSettings* settings) : settings_(settings) {
        std::vector<const char*> k, v;
        for (const auto& kv : settings_->all()) {
            k.push_back(kv.first.c_str());
            v.push_back(kv.second.c_str());
        }
        check(rtps_sph_create(k.data(), v.data(), (int)k.size(), &h_));
        pushed_ = settings_->all();
        settings_->updated();
    }
    ~SPH() override { rtps_sph_destroy(h_); }
    SPH(const SPH&) = delete;
    SPH& operator=(const SPH&) = delete;

    void update() override {
        flushSettings();
        check(rtps_sph_update(h_));
    }
    // SPH::interact (SPH.cpp:610-663): forces of the loaded rigid-body sets.
    void interact() override { check(rtps_sph_interact(h_)); }
    void integrate() override { check(rtps_sph_integrate(h_)); }
    // Rigid-body particle sets (positions in world units); an empty set removes it.
    void setRigidParticles(const std::vector<float4>& pos, const std::vector<float4>& vel,
                           const std::vector<float>& mass, float spring, float dampening) {
        if (vel.size() != pos.size() || mass.size() != pos.size())
            check(RTPS_ERR_ARG);
        check(rtps_sph_set_rigid_particles(h_, pos.data(), vel.data(), mass.data(), (int)pos.size(), spring,
                                           dampening));
    }
    void setStaticBoundary(const std::vector<float4>& pos, float spring, float dampening) {
        check(rtps_sph_set_static_boundary(h_, pos.data(), (int)pos.size(), spring, dampening));
    }
    // SPH::postProcess (SPH.cpp:338-386): the colour-field surface when the
    // settings ask for it ("use_color_field", "color_field_res").
    void postProcess() override {
        flushSettings();
        check(rtps_sph_post_process(h_));
    }
    // SPH::getMCMesh: the last postProcess() surface, or nullptr.
    Mesh* getMCMesh() {
        int n = 0;
        check(rtps_sph_mesh(h_, 0, nullptr, nullptr, &n));
        if (n == 0) return nullptr;
        mesh_.ntris = n;
        mesh_.vertices.resize((size_t)n * 9);
        mesh_.normals.resize((size_t)n * 9);
        check(rtps_sph_mesh(h_, n, mesh_.vertices.data(), mesh_.normals.data(), &n));
        return &mesh_;
    }
    // n x (update + integrate) replayed from one hipGraph (no per-stage host calls).
    void step(int n) {
        flushSettings();
        check(rtps_sph_step(h_, n));
    }
    void synchronize() { check(rtps_sph_synchronize(h_)); }

    int getNum() override {
        int n = 0;
        check(rtps_sph_get_num(h_, &n));
        return n;
    }

    void pushParticles(const std::vector<float4>& pos, const std::vector<float4>& vel,
                       float4 color = float4{1, 0, 0, 1}, float = 0.0f) override {
        check(rtps_sph_push_particles_color(h_, pos.data(), vel.empty() ? nullptr : vel.data(), (int)pos.size(),
                                            color));
    }

    // System::addBox + addRect (System.cpp:331-338, IV.cpp:165-202): float-accumulated
    // lattice at the settings' spacing, z outer, x inner, at most nn particles.
    int addBox(int nn, float4 min, float4 max, bool, float4 color = float4{1, 0, 0, 1}, float = 0.0f) override {
        const float spacing = getSettingFromSystem<float>("spacing");
        std::vector<float4> rvec;
        for (float z = min.z; z <= max.z; z += spacing)
            for (float y = min.y; y <= max.y; y += spacing)
                for (float x = min.x; x <= max.x; x += spacing) {
                    if ((int)rvec.size() >= nn) break;
                    rvec.push_back(float4{x, y, z, 1.0f});
                }
        pushParticles(rvec, std::vector<float4>(), color);
        return (int)rvec.size();
    }

    // Hoses (SPH.cpp:523-557): sprayed at the start of every update with the
    // hose colour; radius in units of the particle spacing.  Mass is not stored.
    int addHose(int total_n, float4 center, float4 velocity, float radius, float4 color = float4{1, 0, 0, 1},
                float = 0.0f) {
        flushSettings();
        int index = -1;
        check(rtps_sph_add_hose(h_, total_n, center, velocity, radius, color, &index));
        return index;
    }
    void updateHose(int index, float4 center, float4 velocity, float radius, float4 color = float4{1, 0, 0, 1}) {
        flushSettings();
        check(rtps_sph_update_hose(h_, index, center, velocity, radius, color));
    }
    void refillHose(int index, int refill) { check(rtps_sph_refill_hose(h_, index, refill)); }
    // SPH::loadTriangles: collision triangles (world units), used with walls.
    void loadTriangles(std::vector<Triangle>& triangles) {
        check(rtps_sph_load_triangles(h_, triangles.data(), (int)triangles.size()));
    }
    // System::testDelete (System.cpp:366-376)
    void testDelete() { check(rtps_sph_test_delete(h_)); }

    RTPSSettings* getSettings() override { return settings_.get(); }

    // Buffer access (System.h:168-177, SPH.h:71): host copies of num entries.
    std::vector<float4> getPositionBufferUnsorted() { return readF4(RTPS_FIELD_POS); }
    std::vector<float4> getVelocityBufferUnsorted() { return readF4(RTPS_FIELD_VEL); }
    std::vector<float4> getPositionBuffer() { return readF4(RTPS_FIELD_POS_S); }
    std::vector<float4> getVelocityBuffer() { return readF4(RTPS_FIELD_VEL_S); }
    std::vector<float4> getColorBufferUnsorted() { return readF4(RTPS_FIELD_COLOR); }
    std::vector<float4> getColorBuffer() { return readF4(RTPS_FIELD_COLOR_S); }
    std::vector<float> getMassBuffer() {
        std::vector<float> out(getNum());
        if (!out.empty()) check(rtps_sph_read(h_, RTPS_FIELD_MASS_S, out.data(), (int)out.size()));
        return out;
    }
    std::vector<float> getDensityBuffer() {
        std::vector<float> out(getNum());
        if (!out.empty()) check(rtps_sph_read(h_, RTPS_FIELD_DENSITY, out.data(), (int)out.size()));
        return out;
    }
    std::vector<unsigned int> getCellStartBuffer() { return readCells(RTPS_FIELD_CELL_START); }
    std::vector<unsigned int> getCellEndBuffer() { return readCells(RTPS_FIELD_CELL_END); }
    void* devicePtr(int field) {
        void* p = nullptr;
        check(rtps_sph_device_ptr(h_, field, &p));
        return p;
    }

    void setOption(int option, int value) { check(rtps_sph_set_option(h_, option, value)); }

    // SPH::printTimers (SPH.cpp:664-672): per-stage average GPU ms.
    std::vector<float> stageTimes() {
        std::vector<float> ms(RTPS_NUM_STAGES);
        check(rtps_sph_stage_times(h_, ms.data(), nullptr, RTPS_NUM_STAGES));
        return ms;
    }

    rtps_sph* handle() { return h_; }

private:
    std::unique_ptr<RTPSSettings> settings_;
    rtps_sph* h_ = nullptr;
    Mesh mesh_;
    std::map<std::string, std::string> pushed_;

    // Settings written through getSettings() since the last step reach the
    // handle here; the handle re-reads its parameters at the next update
    // (SPH.cpp:154).
    void flushSettings() {
        if (!settings_->hasChanged()) return;
        for (const auto& kv : settings_->all()) {
            auto it = pushed_.find(kv.first);
            if (it == pushed_.end() || it->second != kv.second) {
                check(rtps_sph_set_setting(h_, kv.first.c_str(), kv.second.c_str()));
                pushed_[kv.first] = kv.second;
            }
        }
        settings_->updated();
    }

    template <typename T>
    T getSettingFromSystem(const char* key) {
        char buf[128];
        check(rtps_sph_get_setting(h_, key, buf, (int)sizeof(buf)));
        std::istringstream is(buf);
        T v{};
        is >> v;
        return v;
    }

    std::vector<float4> readF4(int field) {
        std::vector<float4> out(getNum());
        if (!out.empty()) check(rtps_sph_read(h_, field, out.data(), (int)out.size()));
        return out;
    }
    std::vector<unsigned int> readCells(int field) {
        rtps_grid_params gp;
        check(rtps_sph_get_params(h_, nullptr, &gp, nullptr));
        std::vector<unsigned int> out((size_t)gp.nb_cells + 1);
        check(rtps_sph_read(h_, field, out.data(), (int)out.size()));
        return out;
    }
};

class RTPS {
public:
    // RTPS::generateSystemInstance (RTPS.cpp:37-67): NULL for an unknown system.
    static System* generateSystemInstance(RTPSSettings* settings) {
        if (settings->GetSettingAs<std::string>("system") != "sph") {
            delete settings;
            return nullptr;
        }
        return new SPH(settings);
    }
};

}  // namespace rtps
