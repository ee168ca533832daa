// Solver.hpp -- header-only C++ facade over the C ABI of sph2s.h with the
// reference's class names, so a caller of the reference
// (SPH2S_FINAL_SINGLE_SCALE/Solver.h, ParticleData.h, Grid.h) keeps its code:
//
//   Grid g = Grid::MakeGrid(origin, end, spacing);                 // Grid.h:16-35
//   SolverConfiguration c = SolverConfiguration::MakeConfiguration(...); // Solver.h:15-48
//   ParticleData* fluid = ParticleData::CreateParticleBox(g);      // ParticleData.cpp:141-176
//   Solver* s = new Solver(fluid, nullptr, &c);                     // Solver.cu:1003-1061
//   s->Advance(0.001f);                                             // Solver.cu:1082-1096
//   fluid->dPositions / fluid->NumParticles                         // what the renderer reads
//
// Failures throw std::runtime_error with the library's message, like
// CUDA::SafeCall (SPH2S_FINAL_TWO_SCALE/cuda.h:115-129).  Differences from the
// reference: there is no GL interop (Map/Unmap are no-ops), and dPositions
// points at float4 (x, y, z, id) records (sph_export_positions_aos gives the
// packed 3-float layout).  Boundary particles (CreateParticleBoxCanvas, Union)
// become the solver's static walls (sph_set_boundary).
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../sph2s.h"

namespace sph2s {

inline void SafeCall(int rc) {
    if (rc != SPH_OK) throw std::runtime_error(std::string("sph2s: ") + sph_last_error());
}

struct float3 {
    float x, y, z;
};

inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }

struct Grid : sph_grid {
    static Grid MakeGrid(float3 origin, float3 end, float spacing) {
        Grid g;
        const float o[3] = {origin.x, origin.y, origin.z};
        const float e[3] = {end.x, end.y, end.z};
        SafeCall(sph_make_grid(o, e, spacing, &g));
        return g;
    }
    static unsigned int ComputeNumGridCells(const Grid& g) { return sph_grid_num_cells(&g); }
    static float ComputeVolume(const Grid& g) {
        return static_cast<float>(g.dims[0] * g.dims[1] * g.dims[2]) * g.spacing * g.spacing *
               g.spacing;
    }
};

struct SolverConfiguration : sph_config {
    static SolverConfiguration MakeConfiguration(float3 origin, float3 end, float volume,
                                                 unsigned int numParticles,
                                                 unsigned int avgParticles, float restDensity,
                                                 float bulkModulus, float viscosity,
                                                 float speedSound, float tensionCoefficient,
                                                 float /*epsilon: XSPH, disabled upstream*/ = 0.0f) {
        SolverConfiguration c;
        const float o[3] = {origin.x, origin.y, origin.z};
        const float e[3] = {end.x, end.y, end.z};
        SafeCall(sph_make_configuration(o, e, volume, numParticles, avgParticles, restDensity,
                                        bulkModulus, viscosity, speedSound, tensionCoefficient,
                                        &c));
        return c;
    }
};

class ParticleData {
public:
    explicit ParticleData(unsigned int maxParticles)
        : NumParticles(0), MaxParticles(maxParticles) {}

    void Map() {}
    void Unmap() {}

    static ParticleData* CreateParticleBox(const Grid& g) {
        auto* p = new ParticleData(Grid::ComputeNumGridCells(g));
        p->hostPositions.reserve(3 * static_cast<size_t>(p->MaxParticles));
        for (int k = 0; k < g.dims[2]; ++k)
            for (int j = 0; j < g.dims[1]; ++j)
                for (int i = 0; i < g.dims[0]; ++i) {
                    p->hostPositions.push_back(g.origin[0] + i * g.spacing);
                    p->hostPositions.push_back(g.origin[1] + j * g.spacing);
                    p->hostPositions.push_back(g.origin[2] + k * g.spacing);
                }
        p->NumParticles = p->MaxParticles;
        return p;
    }

    // The shell of `border` lattice layers around the grid's box, x fastest
    // (ParticleData.cpp:178-209).
    static ParticleData* CreateParticleBoxCanvas(const Grid& g, int border) {
        const long nx = g.dims[0], ny = g.dims[1], nz = g.dims[2], b = border;
        const unsigned int total = static_cast<unsigned int>((nx + 2 * b) * (ny + 2 * b) * (nz + 2 * b) -
                                                             nx * ny * nz);
        auto* p = new ParticleData(total);
        p->hostPositions.reserve(3 * static_cast<size_t>(total));
        for (long k = -b; k < nz + b; ++k)
            for (long j = -b; j < ny + b; ++j)
                for (long i = -b; i < nx + b; ++i) {
                    if (i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz) continue;
                    p->hostPositions.push_back(g.origin[0] + i * g.spacing);
                    p->hostPositions.push_back(g.origin[1] + j * g.spacing);
                    p->hostPositions.push_back(g.origin[2] + k * g.spacing);
                }
        p->NumParticles = total;
        return p;
    }

    static ParticleData* FromArrays(const std::vector<float>& pos_xyz,
                                    const std::vector<float>& vel_xyz, unsigned int maxParticles) {
        auto* p = new ParticleData(maxParticles);
        p->hostPositions = pos_xyz;
        p->hostVelocities = vel_xyz;
        p->NumParticles = static_cast<unsigned int>(pos_xyz.size() / 3);
        return p;
    }

    static ParticleData* Union(const ParticleData* a, const ParticleData* b) {
        if (!a || !b) return nullptr;
        auto* p = new ParticleData(a->MaxParticles + b->MaxParticles);
        p->hostPositions = a->hostPositions;
        p->hostPositions.insert(p->hostPositions.end(), b->hostPositions.begin(),
                                b->hostPositions.end());
        p->NumParticles = a->NumParticles + b->NumParticles;
        return p;
    }

    unsigned int NumParticles;
    unsigned int MaxParticles;
    float* dPositions = nullptr;    // float4 (x, y, z, id) per particle, device
    float* dDensities = nullptr;    // float per particle, device
    std::vector<float> hostPositions, hostVelocities;
};

class Solver {
public:
    Solver(ParticleData* fluidData, ParticleData* boundaryData,
           const SolverConfiguration* configuration)
        : mFluid(fluidData) {
        const float* vel = fluidData->hostVelocities.empty() ? nullptr
                                                             : fluidData->hostVelocities.data();
        SafeCall(sph_create(configuration, fluidData->hostPositions.data(), vel,
                            fluidData->NumParticles, fluidData->MaxParticles, &mHandle));
        if (boundaryData && boundaryData->NumParticles) {
            const int rc = sph_set_boundary(mHandle, boundaryData->hostPositions.data(),
                                            boundaryData->NumParticles);
            if (rc != SPH_OK) {
                sph_destroy(mHandle);
                SafeCall(rc);
            }
        }
        refresh();
    }
    ~Solver() { sph_destroy(mHandle); }
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    void Bind() const {}
    void Advance(float dt) {
        SafeCall(sph_advance(mHandle, dt, 1));
        refresh();
    }
    void Synchronize() { SafeCall(sph_synchronize(mHandle)); }
    sph_solver* handle() { return mHandle; }

private:
    void refresh() {
        sph_views v;
        SafeCall(sph_get_views(mHandle, &v));
        mFluid->dPositions = static_cast<float*>(v.pos4);
        mFluid->dDensities = v.density;
        mFluid->NumParticles = v.n;
    }
    ParticleData* mFluid;
    sph_solver* mHandle = nullptr;
};

// The two-scale Solver (SPH2S_FINAL_TWO_SCALE/Solver.h:120-130): fluidData is
// the low-resolution set (h0, m0), fluidDataHigh the high-resolution set (h0/2,
// m0/8), which starts empty and is filled by splits (its MaxParticles is the
// capacity), boundaryData the static walls.  The reference triggers a split
// from the renderer's surface marks (left commented out upstream,
// TWO_SCALE/Solver.cu:1077-1082); here low-resolution particles inside
// [regionLo, regionHi] split, and setOptions / sph2_mark_split select the
// reference's blended coupling and per-particle trigger.  Advance reads both
// counts back into NumParticles, as the reference's per-step count copy does
// (TWO_SCALE/Solver.cu:1633-1638), so it synchronises once per step.
class TwoScaleSolver {
public:
    TwoScaleSolver(ParticleData* fluidData, ParticleData* fluidDataHigh,
                   ParticleData* boundaryData, const SolverConfiguration* configuration,
                   float3 regionLo, float3 regionHi)
        : mLow(fluidData), mHigh(fluidDataHigh) {
        const float lo[3] = {regionLo.x, regionLo.y, regionLo.z};
        const float hi[3] = {regionHi.x, regionHi.y, regionHi.z};
        const float* vel = fluidData->hostVelocities.empty() ? nullptr
                                                             : fluidData->hostVelocities.data();
        SafeCall(sph2_create(configuration, lo, hi, fluidData->hostPositions.data(), vel,
                             fluidData->NumParticles, fluidData->MaxParticles,
                             fluidDataHigh->MaxParticles, &mHandle));
        if (boundaryData && boundaryData->NumParticles) {
            const int rc = sph2_set_boundary(mHandle, boundaryData->hostPositions.data(),
                                             boundaryData->NumParticles);
            if (rc != SPH_OK) {
                sph2_destroy(mHandle);
                SafeCall(rc);
            }
        }
        refresh();
    }
    ~TwoScaleSolver() { sph2_destroy(mHandle); }
    TwoScaleSolver(const TwoScaleSolver&) = delete;
    TwoScaleSolver& operator=(const TwoScaleSolver&) = delete;

    void Bind() const {}
    void Advance(float dt) {
        SafeCall(sph2_advance(mHandle, dt, 1));
        refresh();
    }
    void setOptions(const sph2_options& o) { SafeCall(sph2_set_options(mHandle, &o)); }
    void Synchronize() { SafeCall(sph2_synchronize(mHandle)); }
    sph_two_scale* handle() { return mHandle; }

private:
    void refresh() {
        uint32_t c[4];
        SafeCall(sph2_counts(mHandle, c));
        mLow->NumParticles = c[0];
        mHigh->NumParticles = c[1];
    }
    ParticleData* mLow;
    ParticleData* mHigh;
    sph_two_scale* mHandle = nullptr;
};

}  // namespace sph2s
