// walls.hpp -- boundary-particle walls shared by the single-scale and
// two-scale solvers (walls.hip).
#pragma once

#include "neighbors.hpp"

namespace sph {

// A static wall set sorted on one grid: positions in cell order, the cells'
// [start, end) ranges, a per-cell flag "a wall particle is within reach of
// this cell", and the constants of the boundary acceleration.
struct Walls {
    uint32_t n = 0, num_cells = 0;
    float4* pos = nullptr;
    uint32_t* cs = nullptr;
    uint32_t* ce = nullptr;
    uint32_t* near = nullptr;
    void* arena = nullptr;
    GridView grid{};      // the grid the walls are sorted on, search radius h
    float h = 0.0f;
    float coef = 0.0f;    // 0.02 c^2
    float mratio = 0.0f;  // m_b / (m + m_b)
};

// Uploads, hashes and sorts n wall particles (AoS xyz) on `grid`; n = 0
// leaves no walls.  Synchronises `st`.
int walls_create(Walls* w, const sph_grid& grid, float h, float boundary_mass, float fluid_mass,
                 float speed_sound, const float* pos_xyz, uint32_t n, hipStream_t st);
void walls_destroy(Walls* w);
// bacc[i] = boundary acceleration of particle i (positions tpos, count
// *n_dev <= cap).
int launch_wall_accel(const Walls& w, const float4* tpos, const uint32_t* n_dev, uint32_t cap,
                      float4* bacc, hipStream_t st);

}  // namespace sph
