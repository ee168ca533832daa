// facade_demo.cpp -- a reference-style host program written against the C++
// facade (include/sph2s/Solver.hpp) only: the single-scale scene of
// SPH2S_FINAL_SINGLE_SCALE/main.cpp:172-251 scaled down (a fluid box, a
// canvas tank and a pillar as boundary particles, MakeConfiguration), then
// Solver::Advance(dt) once per step.  Writes NumParticles and the state's
// positions (sph_export_positions_aos, 3 floats per particle, state order) to
// the file named by argv[2] after argv[1] steps.  Built by build() into
// tests/_bin/; the GPU tests run it and compare with the Python path.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "sph2s/Solver.hpp"

using namespace sph2s;

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: facade_demo <steps> <out.bin>\n");
        return 2;
    }
    const int steps = std::atoi(argv[1]);
    try {
        // The test-sized tank of the walled scenes (1,377 fluid particles, a
        // 3-layer canvas tank and pillar, the reference's constants).
        Grid fluidGrid = Grid::MakeGrid(make_float3(0.015f, 0.015f, 0.015f),
                                        make_float3(0.075f, 0.135f, 0.075f), 0.0075f);
        ParticleData* fluid = ParticleData::CreateParticleBox(fluidGrid);
        ParticleData* tank = ParticleData::CreateParticleBoxCanvas(
            Grid::MakeGrid(make_float3(0.0f, 0.0f, 0.0f), make_float3(0.2f, 0.2f, 0.09f), 0.0075f),
            3);
        ParticleData* pillar = ParticleData::CreateParticleBoxCanvas(
            Grid::MakeGrid(make_float3(0.12f, 0.0f, 0.03f), make_float3(0.135f, 0.2f, 0.045f),
                           0.0075f),
            3);
        ParticleData* walls = ParticleData::Union(tank, pillar);
        SolverConfiguration c = SolverConfiguration::MakeConfiguration(
            make_float3(-0.05f, -0.05f, -0.05f), make_float3(0.25f, 0.25f, 0.14f),
            Grid::ComputeVolume(fluidGrid), fluid->MaxParticles, 35, 1000.0f, 30.0f, 5.0f,
            88.1472f, 0.8f);
        Solver* s = new Solver(fluid, walls, &c);
        s->Bind();
        for (int t = 0; t < steps; ++t) s->Advance(0.001f);
        s->Synchronize();
        std::vector<float> xyz(3 * static_cast<size_t>(fluid->NumParticles));
        SafeCall(sph_export_positions_aos(s->handle(), xyz.data(), 0));
        std::FILE* f = std::fopen(argv[2], "wb");
        if (!f) return 3;
        const unsigned n = fluid->NumParticles;
        std::fwrite(&n, sizeof(n), 1, f);
        std::fwrite(xyz.data(), sizeof(float), xyz.size(), f);
        std::fclose(f);
        std::printf("facade_demo: %u particles, %u walls, %d steps\n", n, walls->NumParticles,
                    steps);
        delete s;
        delete fluid;
        delete tank;
        delete pillar;
        delete walls;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "facade_demo: %s\n", e.what());
        return 1;
    }
    return 0;
}
