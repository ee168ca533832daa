// Solver.hpp -- header-only C++ facade over the C ABI of sph2s.h with the
// reference's class names, so a caller of the reference
// (SPH2S_FINAL_SINGLE_SCALE/Solver.h, ParticleData.h, Grid.h) keeps its code:
//
//   Grid g = Grid::MakeGrid(origin, end, spacing);                 // Grid.h:16-35
//   SolverConfiguration c = SolverConfiguration::MakeConfiguration(...); // Solver.h:15-48
//   ParticleData* fluid = ParticleData::CreateParticleBox(g);      // ParticleData.cpp:141-176
//   Solver* s = new Solver(fluid, nullptr, &c);                     // Solver.cu:1003-1061
//   s->Advance(0.001f);                                             // Solver.cu:1082-1096
//   fluid->dPositions / dColorValues / NumParticles                 // what the renderer reads
//
// A two-scale caller (SPH2S_FINAL_TWO_SCALE/Solver.h, main.cpp:228-254) uses
// the names of namespace sph2s::two_scale instead: its SolverConfiguration
// carries EffectiveRadius[2], FluidParticleMass[2] and BlendIncrement, and its
// Solver(fluid, fluidHigh, boundary, &config) runs the reference's blended
// coupling.  TwoScaleSolver is the region-driven split/merge form.
//
// ParticleData's device buffers are the renderer's: after every Advance,
// dPositions holds 3 floats per particle (x, y, z), dColorValues 1 float and
// dStates 1 int (ParticleData.h:39-45, Renderer.cpp:38-46), in the solver's
// current (sorted) order, for NumParticles particles.  The solver owns them
// (sph_enable_render_buffers); they stay valid while the solver lives.
//
// Failures throw std::runtime_error with the library's message, like
// CUDA::SafeCall (SPH2S_FINAL_TWO_SCALE/cuda.h:115-129); that includes the
// device-side errors of a step (non-finite state, two-scale capacity, a sort
// wait that gave up), which Synchronize (and the two-scale Advance, which
// synchronises for the counts) report.  There is no GL interop: Map/Unmap are
// no-ops.  Boundary particles (CreateParticleBoxCanvas, Union) become the
// solver's static walls (sph_set_boundary).
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
    float* dPositions = nullptr;    // device, 3 floats (x, y, z) per particle
    float* dColorValues = nullptr;  // device, 1 float per particle
    int* dStates = nullptr;         // device, 1 int per particle (two-scale state)
    float* dDensities = nullptr;    // device, 1 float per particle (last step)
    std::vector<float> hostPositions, hostVelocities;
};

namespace detail {
inline void bind(ParticleData* p, const sph_render_buffers& r) {
    p->dPositions = r.positions;
    p->dColorValues = r.color_values;
    p->dStates = r.states;
    p->NumParticles = r.n;
}
}  // namespace detail

class Solver {
public:
    Solver(ParticleData* fluidData, ParticleData* boundaryData,
           const SolverConfiguration* configuration)
        : mFluid(fluidData) {
        const float* vel = fluidData->hostVelocities.empty() ? nullptr
                                                             : fluidData->hostVelocities.data();
        SafeCall(sph_create(configuration, fluidData->hostPositions.data(), vel,
                            fluidData->NumParticles, fluidData->MaxParticles, &mHandle));
        int rc = SPH_OK;
        if (boundaryData && boundaryData->NumParticles)
            rc = sph_set_boundary(mHandle, boundaryData->hostPositions.data(),
                                  boundaryData->NumParticles);
        // The single-scale reference never writes dColorValues: they stay 0
        // unless SetColorValues(true) turns the colour pass on.
        if (rc == SPH_OK) rc = sph_enable_render_buffers(mHandle, 0u);
        if (rc != SPH_OK) {
            sph_destroy(mHandle);
            SafeCall(rc);
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
    // Waits for the solver and throws on a device-side error of its steps.
    void Synchronize() { SafeCall(sph_synchronize(mHandle)); }
    // Colour values of the two-scale reference's force pass in dColorValues
    // (the colour pass, SPH_DEBUG_COLOR) every step.
    void SetColorValues(bool on) {
        SafeCall(sph_enable_render_buffers(mHandle, on ? SPH_RENDER_COLORS : 0u));
        refresh();
    }
    sph_solver* handle() { return mHandle; }

private:
    void refresh() {
        sph_views v;
        SafeCall(sph_get_views(mHandle, &v));
        sph_render_buffers r;
        SafeCall(sph_get_render_buffers(mHandle, &r));
        detail::bind(mFluid, r);
        mFluid->dDensities = v.density;
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
        mHandle = create(fluidData, fluidDataHigh, boundaryData, configuration, lo, hi, 0u);
        refresh();
    }
    ~TwoScaleSolver() { sph2_destroy(mHandle); }
    TwoScaleSolver(const TwoScaleSolver&) = delete;
    TwoScaleSolver& operator=(const TwoScaleSolver&) = delete;

    void Bind() const {}
    // One step; reads both counts back (and with them any device-side error
    // of the step, which throws), as the reference's per-step count copy does.
    void Advance(float dt) {
        SafeCall(sph2_advance(mHandle, dt, 1));
        refresh();
    }
    void setOptions(const sph2_options& o) {
        SafeCall(sph2_set_options(mHandle, &o));
        refresh();
    }
    void SetColorValues(bool on) {
        SafeCall(sph2_enable_render_buffers(mHandle, on ? SPH_RENDER_COLORS : 0u));
        refresh();
    }
    void Synchronize() { SafeCall(sph2_synchronize(mHandle)); }
    sph_two_scale* handle() { return mHandle; }

    // sph2_create + walls + renderer buffers, for both two-scale facades.  The
    // high set is filled by splits only: a fluidDataHigh with particles is
    // rejected (its particles would need group ids and tags).
    static sph_two_scale* create(ParticleData* low, ParticleData* high, ParticleData* walls,
                                 const sph_config* cfg, const float lo[3], const float hi[3],
                                 uint32_t renderFlags) {
        if (!low || !high) throw std::runtime_error("sph2s: fluid data is null");
        if (high->NumParticles != 0 || !high->hostPositions.empty())
            throw std::runtime_error(
                "sph2s: the high-resolution ParticleData must start empty (NumParticles = 0); "
                "its particles come from splits");
        const float* vel = low->hostVelocities.empty() ? nullptr : low->hostVelocities.data();
        sph_two_scale* h = nullptr;
        SafeCall(sph2_create(cfg, lo, hi, low->hostPositions.data(), vel, low->NumParticles,
                             low->MaxParticles, high->MaxParticles, &h));
        int rc = SPH_OK;
        if (walls && walls->NumParticles)
            rc = sph2_set_boundary(h, walls->hostPositions.data(), walls->NumParticles);
        if (rc == SPH_OK) rc = sph2_enable_render_buffers(h, renderFlags);
        if (rc != SPH_OK) {
            sph2_destroy(h);
            SafeCall(rc);
        }
        return h;
    }
    // Both sets' counts and renderer buffers; throws on a device-side error.
    static void refresh(sph_two_scale* h, ParticleData* low, ParticleData* high) {
        SafeCall(sph2_synchronize(h));
        sph_render_buffers r;
        SafeCall(sph2_get_render_buffers(h, 0, &r));
        detail::bind(low, r);
        SafeCall(sph2_get_render_buffers(h, 1, &r));
        detail::bind(high, r);
    }

private:
    void refresh() { refresh(mHandle, mLow, mHigh); }
    ParticleData* mLow;
    ParticleData* mHigh;
    sph_two_scale* mHandle = nullptr;
};

// The names of the two-scale reference build (SPH2S_FINAL_TWO_SCALE/Solver.h),
// which declares its own SolverConfiguration and Solver.
namespace two_scale {

using sph2s::Grid;
using sph2s::ParticleData;
using sph2s::float3;
using sph2s::make_float3;

// SolverConfiguration of TWO_SCALE/Solver.h:14-60: both resolutions' radius
// and mass, the blend increment and one grid per resolution.  The library
// configuration of the low set (h0, m0) is kept beside the reference members.
struct SolverConfiguration {
    static SolverConfiguration MakeConfiguration(float3 origin, float3 end, float volume,
                                                 unsigned int numParticles,
                                                 unsigned int avgParticles, float restDensity,
                                                 float bulkModulus, float viscosity,
                                                 float speedSound, float tensionCoefficient,
                                                 float blendIncrement) {
        SolverConfiguration c;
        const float o[3] = {origin.x, origin.y, origin.z};
        const float e[3] = {end.x, end.y, end.z};
        SafeCall(sph_make_configuration(o, e, volume, numParticles, avgParticles, restDensity,
                                        bulkModulus, viscosity, speedSound, tensionCoefficient,
                                        &c.Low));
        c.EffectiveRadius[0] = c.Low.h;
        c.EffectiveRadius[1] = 0.5f * c.Low.h;               // Solver.h:36
        c.FluidParticleMass[0] = c.Low.mass;
        c.FluidParticleMass[1] = 1.0f / 8.0f * c.Low.mass;   // Solver.h:39
        c.BoundaryParticleMass = c.Low.boundary_mass;
        c.RestDensity = restDensity;
        c.Grid[0] = sph2s::Grid::MakeGrid(origin, end, c.EffectiveRadius[0]);
        c.Grid[1] = sph2s::Grid::MakeGrid(origin, end, c.EffectiveRadius[1]);
        c.BulkModulus = bulkModulus;
        c.Viscosity = viscosity;
        c.SpeedSound = speedSound;
        c.TensionCoefficient = tensionCoefficient;
        c.BlendIncrement = blendIncrement;
        return c;
    }

    float EffectiveRadius[2];
    float FluidParticleMass[2];
    float BoundaryParticleMass;
    float RestDensity;
    float BulkModulus;
    float Viscosity;
    float SpeedSound;
    float TensionCoefficient;
    float BlendIncrement;
    sph2s::Grid Grid[2];
    sph_config Low;  // the library's low-set configuration (h0, m0, grid 0)
};

// Solver(fluid, fluidHigh, boundary, &config) of TWO_SCALE/Solver.h:120-130
// with the reference's step (TWO_SCALE/Solver.cu:1568-1614): the blended
// coupling with every neighbour term weighted by lambda_j, the fade timeline
// with config.BlendIncrement per step, the velocity cap 0.8 h / dt
// (Solver.cu:1212-1223) and children at r / sqrt(3) along the diagonals
// (insertHighResParticles, Solver.cu:541-586).  As in the reference, a split
// happens only where a particle is marked (MarkSplit, or DetectSurface, the
// view-independent form of the renderer's surface marks that the reference
// left commented out at Solver.cu:1077-1082) and nothing merges.  Colour
// values and state words are written every step, as the reference's passes
// write dColorValues and dStates.
class Solver {
public:
    Solver(ParticleData* fluidData, ParticleData* fluidDataHigh, ParticleData* boundaryData,
           const SolverConfiguration* configuration)
        : mLow(fluidData), mHigh(fluidDataHigh) {
        // An empty split region (lo > hi): no particle is inside it.
        const float lo[3] = {1.0f, 1.0f, 1.0f}, hi[3] = {0.0f, 0.0f, 0.0f};
        mHandle = TwoScaleSolver::create(fluidData, fluidDataHigh, boundaryData,
                                         &configuration->Low, lo, hi, SPH_RENDER_COLORS);
        const sph2_options o{SPH2_COUPLING_BLENDED, SPH2_SPLIT_REFERENCE, 0u,
                             configuration->BlendIncrement, 0.8f};
        const int rc = sph2_set_options(mHandle, &o);
        if (rc != SPH_OK) {
            sph2_destroy(mHandle);
            SafeCall(rc);
        }
        TwoScaleSolver::refresh(mHandle, mLow, mHigh);
    }
    ~Solver() { sph2_destroy(mHandle); }
    Solver(const Solver&) = delete;
    Solver& operator=(const Solver&) = delete;

    void Bind() const {}
    void Advance(float dt) {
        SafeCall(sph2_advance(mHandle, dt, 1));
        TwoScaleSolver::refresh(mHandle, mLow, mHigh);
    }
    // Request the split of the low-resolution particles with these ids (the
    // renderer's mark, state | 3); they split in the next step.
    void MarkSplit(const std::vector<unsigned int>& ids) {
        SafeCall(sph2_mark_split(mHandle, ids.data(), static_cast<uint32_t>(ids.size())));
    }
    // Mark the free surface of the last step (colour-field gradient |n| h0 >
    // threshold) for splitting.
    void DetectSurface(float threshold) {
        SafeCall(sph2_detect_surface(mHandle, threshold, 1));
    }
    void Synchronize() { SafeCall(sph2_synchronize(mHandle)); }
    sph_two_scale* handle() { return mHandle; }

private:
    ParticleData* mLow;
    ParticleData* mHigh;
    sph_two_scale* mHandle = nullptr;
};

}  // namespace two_scale

}  // namespace sph2s
