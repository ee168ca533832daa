// facade_demo.cpp -- a reference-style host program written against the C++
// facade (include/sph2s/Solver.hpp) only: the single-scale scene of
// SPH2S_FINAL_SINGLE_SCALE/main.cpp:172-251 scaled down (a fluid box, a
// canvas tank and a pillar as boundary particles, MakeConfiguration), then
// Solver::Advance(dt) once per step.  After argv[1] steps it writes
// NumParticles and the renderer's buffers as the reference's renderer reads
// them -- fluid->dPositions (3 floats per particle) and fluid->dColorValues,
// copied from the device -- to the file named by argv[2].
//   "two <steps> <out>"  the region-driven TwoScaleSolver;
//   "ref2 <steps> <out>" the reference's two-scale program shape
//                        (sph2s::two_scale names, TWO_SCALE/main.cpp:150-254):
//                        blended coupling, splits requested by id.
// Built by build() into tests/_bin/; the GPU tests run it and compare with
// the Python path.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sph2s/Solver.hpp"

using namespace sph2s;

// hipMemcpy device to host (tests/facade_copy.cpp).
void copy_from_device(void* dst, const void* src, std::size_t bytes);

// Copies n elements of a device buffer the renderer would bind.
template <class T>
static std::vector<T> from_device(const T* d, size_t n) {
    std::vector<T> h(n);
    copy_from_device(h.data(), d, n * sizeof(T));
    return h;
}

// NumParticles, then dPositions (3n floats), dColorValues (n) and dStates (n)
// of each set.
static void write_sets(std::FILE* f, const std::vector<ParticleData*>& sets) {
    for (ParticleData* p : sets) std::fwrite(&p->NumParticles, sizeof(unsigned), 1, f);
    for (ParticleData* p : sets) {
        const size_t n = p->NumParticles;
        const std::vector<float> xyz = from_device(p->dPositions, 3 * n);
        const std::vector<float> col = from_device(p->dColorValues, n);
        const std::vector<int> st = from_device(p->dStates, n);
        std::fwrite(xyz.data(), sizeof(float), xyz.size(), f);
        std::fwrite(col.data(), sizeof(float), col.size(), f);
        std::fwrite(st.data(), sizeof(int), st.size(), f);
    }
}

// Two-scale mode: a lattice box of fluid, MakeConfiguration, a split region
// over one corner of it, TwoScaleSolver::Advance once per step; writes the
// (low, high) counts and both sets' positions.
static int region_two_scale(int steps, const char* path) {
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
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return 3;
    write_sets(f, {low, high});
    std::fclose(f);
    std::printf("facade_demo two-scale: %u low, %u high, %d steps\n", low->NumParticles,
                high->NumParticles, steps);
    delete s;
    delete low;
    delete high;
    return 0;
}

// The reference's two-scale program shape (TWO_SCALE/main.cpp:150-254):
// CreateParticleBox, an empty high-resolution ParticleData, canvas walls, the
// two-scale MakeConfiguration with its blend increment, Solver(fluid,
// fluidHigh, boundary, &config).  After the first step every 7th low particle
// id is marked for splitting (the renderer's mark), then the rest run.
static int reference_two_scale(int steps, const char* path) {
    namespace ts = sph2s::two_scale;
    ts::Grid fluidGrid = ts::Grid::MakeGrid(make_float3(0.02f, 0.02f, 0.02f),
                                            make_float3(0.14f, 0.2f, 0.14f), 0.01f);
    ts::ParticleData* low = ts::ParticleData::CreateParticleBox(fluidGrid);
    ts::ParticleData* high = new ts::ParticleData(low->MaxParticles * 8);
    high->NumParticles = 0;
    ts::ParticleData* walls = ts::ParticleData::CreateParticleBoxCanvas(
        ts::Grid::MakeGrid(make_float3(0.0f, 0.0f, 0.0f), make_float3(0.3f, 0.3f, 0.3f), 0.01f),
        2);
    ts::SolverConfiguration c = ts::SolverConfiguration::MakeConfiguration(
        make_float3(-0.05f, -0.05f, -0.05f), make_float3(0.35f, 0.35f, 0.35f),
        ts::Grid::ComputeVolume(fluidGrid), low->MaxParticles, 35, 1000.0f, 30.0f, 5.0f,
        88.1472f, 0.8f, 0.0125f);
    ts::Solver* s = new ts::Solver(low, high, walls, &c);
    s->Bind();
    for (int t = 0; t < steps; ++t) {
        s->Advance(0.001f);
        if (t == 0) {
            std::vector<unsigned int> ids;
            for (unsigned i = 0; i < low->MaxParticles; i += 7) ids.push_back(i);
            s->MarkSplit(ids);
        }
    }
    s->Synchronize();
    std::FILE* f = std::fopen(path, "wb");
    if (!f) return 3;
    write_sets(f, {low, high});
    std::fclose(f);
    std::printf("facade_demo reference two-scale: %u low, %u high, %d steps\n",
                low->NumParticles, high->NumParticles, steps);
    delete s;
    delete low;
    delete high;
    delete walls;
    return 0;
}

int main(int argc, char** argv) {
    if (argc == 4 && (std::string(argv[1]) == "two" || std::string(argv[1]) == "ref2")) {
        try {
            return std::string(argv[1]) == "two" ? region_two_scale(std::atoi(argv[2]), argv[3])
                                                 : reference_two_scale(std::atoi(argv[2]), argv[3]);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "facade_demo: %s\n", e.what());
            return 1;
        }
    }
    if (argc != 3) {
        std::fprintf(stderr, "usage: facade_demo <steps> <out.bin> | two|ref2 <steps> <out.bin>\n");
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
        std::FILE* f = std::fopen(argv[2], "wb");
        if (!f) return 3;
        write_sets(f, {fluid});
        std::fclose(f);
        const unsigned n = fluid->NumParticles;
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
