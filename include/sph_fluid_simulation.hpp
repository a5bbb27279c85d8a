// sph_fluid_simulation.hpp -- C++ drop-in for Physics::Fluid::FluidSimulation
// (Allkams/Fluid-Simulation-3D, engine/physics/physicsWorld.h:28-155) and a
// FluidSimBase-shaped backend (projects/Simulation/code/simulations/
// fluidSimBase.h:5-13), header-only over the C ABI of sph_amd.h.
//
// Same method names, argument meanings and error behaviour as the reference:
// getters return zero for an out-of-range index (physicsWorld.cc:149-182),
// setters take effect at the next Update, the public `positions` and
// `OutPositions` vectors hold the state after each Update (the reference's
// render hand-off, fluidSimCPU.cc:104), and positions written into
// `positions` between Updates are the ones the next Update starts from, as
// stage A reads them in the reference (physicsWorld.cc:47).  Vector types are layout-compatible
// with glm::vec3 / glm::vec4 (three / four packed floats); define
// SPH_AMD_VEC3 / SPH_AMD_VEC4 to use the application's own (e.g. glm) types.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "sph_amd.h"

namespace sph_amd {
struct vec3 {
    float x, y, z;
};
struct vec4 {
    float x, y, z, w;
};
}  // namespace sph_amd

#ifndef SPH_AMD_VEC3
#define SPH_AMD_VEC3 sph_amd::vec3
#endif
#ifndef SPH_AMD_VEC4
#define SPH_AMD_VEC4 sph_amd::vec4
#endif

namespace Physics {
namespace Fluid {

class FluidSimulation {
  public:
    using vec3 = SPH_AMD_VEC3;
    using vec4 = SPH_AMD_VEC4;

    // physicsWorld.h:32 -- one instance per process, as in the reference
    static FluidSimulation &getInstance()
    {
        static FluidSimulation instance;
        return instance;
    }

    // physicsWorld.h:34 / cc:39-111: one Update on the GPU, then the host
    // copies of positions, velocities and densities are refreshed.  Positions
    // the caller wrote into `positions` since the last refresh go to the
    // device first (only positions; velocities keep the device's values).
    void Update(float deltatime)
    {
        if (!h_) return;
        push_positions();
        check(sph_update(h_, deltatime));
        pull();
    }

    // physicsWorld.h:36 / cc:112-147 (Centre is unused there too)
    void InitializeData(int particleAmmount, vec3 Centre = {0, 0, 0})
    {
        (void)Centre;
        if (particleAmmount <= 0) return;
        if (h_ && n_ != (uint32_t)particleAmmount) {
            sph_destroy(h_);
            h_ = nullptr;
        }
        if (!h_) {
            check(sph_create(&p_, (uint32_t)particleAmmount, 0, &h_));
            n_ = (uint32_t)particleAmmount;
            check(sph_set_timing(h_, 1));
        }
        check(sph_initialize_data(h_));
        pull();
    }

    // physicsWorld.h:38-43 (cc:149-182)
    vec3 getPosition(uint32_t i) { return i < n_ ? vec3{pos4_[4 * i], pos4_[4 * i + 1], pos4_[4 * i + 2]} : vec3{0, 0, 0}; }
    vec3 getVelocity(uint32_t i) { return i < n_ ? vec3{vel4_[4 * i], vel4_[4 * i + 1], vel4_[4 * i + 2]} : vec3{0, 0, 0}; }
    float getDensity(uint32_t i) { return i < n_ ? dens_[2 * i] : 0.0f; }
    float getNearDensity(uint32_t i) { return i < n_ ? dens_[2 * i + 1] : 0.0f; }
    float getSpeed(uint32_t i)
    {
        if (i >= n_) return 0.0f;
        const float x = vel4_[4 * i], y = vel4_[4 * i + 1], z = vel4_[4 * i + 2];
        return std::sqrt(x * x + y * y + z * z);
    }
    float getSpeedNormalzied(uint32_t i)  // (sic) physicsWorld.h:43
    {
        if (i >= n_) return 0.0f;
        return std::min(std::max(getSpeed(i), 0.0f), 1.5f) / 1.5f;
    }

    // physicsWorld.h:45-50: stage times of the last Update in ms, measured
    // with device events around each kernel.  Two of the reference's six
    // stages have no kernel of their own:
    //  - gravity (stage A, cc:41-50) reads the time of the separate predict
    //    kernel, which runs only when the previous Update did not record this
    //    one's buckets (the first Update, or after new state, settings or dt);
    //    on every other Update stage A runs inside the previous Update's
    //    viscosity kernel and this one's sort, so it reads 0;
    //  - position + collision (stage F, cc:80-110) runs in the epilogue of the
    //    viscosity kernel, so it reads 0 and getElapsedTimeViscosity includes it.
    double getElapsedTimeGravity() { return stage(0); }
    double getElapsedTimeSpatial() { return stage(1); }
    double getElapsedTimeDensity() { return stage(2); }
    double getElapsedTimePressure() { return stage(3); }
    double getElapsedTimeViscosity() { return stage(4); }
    double getElapsedTimePosNColl() { return stage(5); }

    // physicsWorld.h:52-77
    void setSimulationTime(float time) { simTime_ = time; }
    float getSimulationTime() { return simTime_; }
    void setGravity(bool status) { p_.gravity_on = status ? 1 : 0; push(); }
    bool getGravityStatus() { return p_.gravity_on != 0; }
    void setInteractionRadius(float value) { p_.smoothing_radius = value; push(); }
    float getInteractionRadius() { return p_.smoothing_radius; }
    void setDensityTarget(float value) { p_.target_density = value; push(); }
    float getDensityTarget() { return p_.target_density; }
    void setPressureMultiplier(float value) { p_.pressure_mult = value; push(); }
    float getPressureMultiplier() { return p_.pressure_mult; }
    void setNearPressureMultiplier(float value) { p_.near_pressure_mult = value; push(); }
    float getNearPressureMultiplier() { return p_.near_pressure_mult; }
    void setViscosityStrength(float value) { p_.viscosity = value; push(); }
    float getViscosityStrength() { return p_.viscosity; }
    void setGravityScale(float value) { p_.gravity_scale = value; push(); }
    float getGravityScale() { return p_.gravity_scale; }
    void setBound(const vec3 &value)
    {
        p_.bounds[0] = value.x;
        p_.bounds[1] = value.y;
        p_.bounds[2] = value.z;
        push();
    }
    vec3 getBounds() { return vec3{p_.bounds[0], p_.bounds[1], p_.bounds[2]}; }

    // physicsWorld.h:79-80
    std::vector<vec3> positions;
    std::vector<vec4> OutPositions;

    // Beyond the reference: the C handle, for the batch/serving entry points.
    sph_handle handle() const { return h_; }

    FluidSimulation(const FluidSimulation &) = delete;
    FluidSimulation &operator=(const FluidSimulation &) = delete;

  private:
    FluidSimulation() { sph_default_params(&p_); }
    ~FluidSimulation()
    {
        if (h_) sph_destroy(h_);
    }

    static void check(int rc)
    {
        if (rc != SPH_OK) throw std::runtime_error(std::string("sph_amd: ") + sph_last_error());
    }
    void push()
    {
        if (h_) check(sph_set_params(h_, &p_));
    }
    void pull()
    {
        pos4_.resize(4 * (size_t)n_);
        vel4_.resize(4 * (size_t)n_);
        dens_.resize(2 * (size_t)n_);
        check(sph_download(h_, pos4_.data(), vel4_.data(), dens_.data()));
        positions.resize(n_);
        OutPositions.resize(n_);
        for (uint32_t i = 0; i < n_; ++i) {
            positions[i] = vec3{pos4_[4 * i], pos4_[4 * i + 1], pos4_[4 * i + 2]};
            OutPositions[i] = vec4{pos4_[4 * i], pos4_[4 * i + 1], pos4_[4 * i + 2], pos4_[4 * i + 3]};
        }
    }
    // The caller's writes to `positions` (h:79) since the last pull, to the device.
    void push_positions()
    {
        if (positions.size() != n_ || pos4_.size() != 4 * (size_t)n_) return;
        bool changed = false;
        for (uint32_t i = 0; i < n_ && !changed; ++i)
            changed = positions[i].x != pos4_[4 * i] || positions[i].y != pos4_[4 * i + 1] ||
                      positions[i].z != pos4_[4 * i + 2];
        if (!changed) return;
        for (uint32_t i = 0; i < n_; ++i) {
            pos4_[4 * i] = positions[i].x;
            pos4_[4 * i + 1] = positions[i].y;
            pos4_[4 * i + 2] = positions[i].z;
        }
        check(sph_upload(h_, pos4_.data(), nullptr));
    }
    double stage(int q)
    {
        double ms[6] = {0, 0, 0, 0, 0, 0};
        if (h_) check(sph_stage_ms(h_, ms));
        return ms[q];
    }

    sph_handle h_ = nullptr;
    uint32_t n_ = 0;
    sph_params p_{};
    float simTime_ = 0.0f;
    std::vector<float> pos4_, vel4_, dens_;
};

}  // namespace Fluid
}  // namespace Physics

namespace sph_amd {

// The FluidSimBase backend slot (fluidSimBase.h:5-13) without its render
// half: a FluidSimHIP for createSimulation (gameApp.cc:87-92) derives from the
// application's FluidSimBase and forwards to these.
class FluidSimHIPCore {
  public:
    using vec4 = Physics::Fluid::FluidSimulation::vec4;
    void initialize(int particleAmount)  // fluidSimCPU.cc:9-33
    {
        n_ = particleAmount;
        Physics::Fluid::FluidSimulation::getInstance().InitializeData(particleAmount);
        if (colors_on_) set_colors(stops_);
    }
    void update(float dt) { Physics::Fluid::FluidSimulation::getInstance().Update(dt); }  // :35-40
    void reset() { initialize(n_); }                                                      // :42-46
    void cleanup() {}                                                                     // :48-51
    // the SSBO upload source of render(): OutPositions (fluidSimCPU.cc:104)
    const std::vector<vec4> &out_positions() const
    {
        return Physics::Fluid::FluidSimulation::getInstance().OutPositions;
    }
    // FluidSimCPU::updateColors (fluidSimCPU.cc:100-125), computed on the GPU
    // inside every Update from then on: stops = Color1..Color4 (RGBA each).
    void set_colors(const float stops[16])
    {
        for (int c = 0; c < 16; ++c) stops_[c] = stops[c];
        colors_on_ = true;
        sph_handle h = Physics::Fluid::FluidSimulation::getInstance().handle();
        if (h && sph_set_colors(h, stops_, 1) != SPH_OK)
            throw std::runtime_error(std::string("sph_amd: ") + sph_last_error());
    }
    // The colours of the last Update (the render() colour buffer, fluidSimCPU.cc:80-83).
    const std::vector<vec4> &colors()
    {
        sph_handle h = Physics::Fluid::FluidSimulation::getInstance().handle();
        colors_.resize((size_t)n_);
        if (!h || !colors_on_ || n_ <= 0) return colors_;
        if (sph_download_colors(h, reinterpret_cast<float *>(colors_.data())) != SPH_OK)
            throw std::runtime_error(std::string("sph_amd: ") + sph_last_error());
        return colors_;
    }

  private:
    int n_ = 0;
    bool colors_on_ = false;
    float stops_[16] = {0.0f, 0.75f, 1.0f, 1.0f, 0.0f, 1.0f, 0.0f, 1.0f,
                        1.0f, 1.0f, 0.0f, 1.0f, 1.0f, 0.0f, 0.0f, 1.0f};  // fluidSimCPU.h:26-29
    std::vector<vec4> colors_;
};

}  // namespace sph_amd
