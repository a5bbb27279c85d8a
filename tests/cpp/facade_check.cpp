// Drives the C++ drop-in (include/sph_fluid_simulation.hpp) the way the
// reference application drives Physics::Fluid::FluidSimulation
// (gameApp.cc:112-113, 231, 363-411): InitializeData, settings, Update per
// frame, getters; and the FluidSimBase-shaped backend with render colours.
// Writes the final positions, velocities, densities and colours (float32) to
// argv[1] for the GPU test to compare with the C-ABI path and the oracle.
#include <stdio.h>

#include "sph_fluid_simulation.hpp"

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: facade_check OUT\n");
        return 2;
    }
    auto &sim = Physics::Fluid::FluidSimulation::getInstance();
    // the FluidSimBase-shaped backend: initialize = InitializeData(particleAmount)
    // (fluidSimCPU.cc:9-33, gameApp.cc:41), colours on with FluidSimCPU's stops
    sph_amd::FluidSimHIPCore core;
    core.initialize(10000);
    const float stops[16] = {0.0f, 0.75f, 1.0f, 1.0f, 0.0f, 1.0f, 0.0f, 1.0f,
                             1.0f, 1.0f, 0.0f, 1.0f, 1.0f, 0.0f, 0.0f, 1.0f};
    core.set_colors(stops);
    sim.setGravity(true);
    sim.setGravityScale(9.81f);
    sim.setViscosityStrength(0.3f);
    sim.setBound({6.0f, 6.0f, 6.0f});
    for (int frame = 0; frame < 5; ++frame) sim.Update(1.0f / 60.0f);
    // out-of-range getters return zero (physicsWorld.cc:151-181)
    if (sim.getDensity(10000) != 0.0f || sim.getPosition(99999).x != 0.0f) return 3;
    if (sim.positions.size() != 10000 || sim.OutPositions[5].w != 0.34f) return 4;
    if (!(sim.getElapsedTimeDensity() > 0.0)) return 5;
    const auto &col = core.colors();
    if (col.size() != 10000 || core.out_positions().size() != 10000) return 7;
    FILE *f = fopen(argv[1], "wb");
    if (!f) return 6;
    auto dump = [&](const std::vector<sph_amd::FluidSimHIPCore::vec4> &c) {
        for (uint32_t i = 0; i < 10000; ++i) {
            const auto p = sim.getPosition(i), v = sim.getVelocity(i);
            const float rec[12] = {p.x, p.y, p.z, v.x, v.y, v.z, sim.getDensity(i), sim.getNearDensity(i),
                                   c[i].x, c[i].y, c[i].z, c[i].w};
            fwrite(rec, sizeof(float), 12, f);
        }
    };
    dump(col);
    // the public positions vector is an input too (physicsWorld.h:79): lift
    // the fluid by 0.05 and step once more from there
    for (auto &p : sim.positions) p.y += 0.05f;
    sim.Update(1.0f / 60.0f);
    dump(core.colors());
    fclose(f);
    printf("facade ok: speed[0]=%g normalized=%g\n", sim.getSpeed(0), sim.getSpeedNormalzied(0));
    return 0;
}
