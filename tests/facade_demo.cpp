// facade_demo.cpp -- a reference-style host program written against the C++
// facade (include/sph2s/Solver.hpp) only: the single-scale scene of
// SPH2S_FINAL_SINGLE_SCALE/main.cpp:172-251 scaled down (a fluid box, a
// canvas tank and a pillar as boundary particles, MakeConfiguration), then
// Solver::Advance(dt) once per step.  Writes NumParticles and the state's
// positions (sph_export_positions_aos, 3 floats per particle, state order) to
// the file named by argv[2] after argv[1] steps.  "two <steps> <out>" runs
// the two-scale form (TwoScaleSolver) instead.  Built by build() into
// tests/_bin/; the GPU tests run it and compare with the Python path.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sph2s/Solver.hpp"

using namespace sph2s;

// Two-scale mode: a lattice box of fluid, MakeConfiguration, a split region
// over one corner of it, TwoScaleSolver::Advance once per step; writes the
// (low, high) counts and both sets' positions.
static int two_scale(int steps, const char* path) {
    Grid fluidGrid = Grid::MakeGrid(make_float3(0.02f, 0.02f, 0.02f),
                                    make_float3(0.2f, 0.2f, 0.2f), 0.01f);
    ParticleData* low = ParticleData::CreateParticleBox(fluidGrid);
    ParticleData* high = new ParticleData(8 * low->MaxParticles);
    SolverConfiguration c = SolverConfiguration::MakeConfiguration(
        make_float3(0.0f, 0.0f, 0.0f), make_float3(0.4f, 0.4f, 0.4f),
        Grid::ComputeVolume(fluidGrid), low->MaxParticles, 35, 1000.0f, 30.0f, 5.0f, 88.1472f,
        0.0f);
    TwoScaleSolver* s = new TwoScaleSolver(low, high, nullptr, &c,
                                           make_float3(0.02f, 0.02f, 0.02f),
                                           make_float3(0.1f, 0.1f, 0.1f));
    s->Bind();
    for (int t = 0; t < steps; ++t) s->Advance(0.001f);
    s->Synchronize();
    const unsigned n[2] = {low->NumParticles, high->NumParticles};
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return 3;
    std::fwrite(n, sizeof(unsigned), 2, f);
    for (int r = 0; r < 2; ++r) {
        std::vector<float> xyz(3 * static_cast<size_t>(n[r]));
        if (n[r])
            SafeCall(sph2_download(s->handle(), r, SPH_FIELD_POSITION, xyz.data(),
                                   xyz.size() * sizeof(float)));
        std::fwrite(xyz.data(), sizeof(float), xyz.size(), f);
    }
    std::fclose(f);
    std::printf("facade_demo two-scale: %u low, %u high, %d steps\n", n[0], n[1], steps);
    delete s;
    delete low;
    delete high;
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4 && std::string(argv[1]) == "two") {
        try {
            return two_scale(std::atoi(argv[2]), argv[3]);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "facade_demo: %s\n", e.what());
            return 1;
        }
    }
    if (argc != 3) {
        std::fprintf(stderr, "usage: facade_demo <steps> <out.bin> | two <steps> <out.bin>\n");
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
