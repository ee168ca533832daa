// neighbors.hpp -- device-side neighbour enumeration shared by the density,
// force and adjoint kernels.
//
// The stencil bounds are the reference's: cells clamp((int)((x -/+ h - o)/D))
// per axis (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:559-593).  Within one grid row
// (cy, cz) the cells cx are consecutive hashes, so the selected cells of a row
// form one contiguous run [b, e) of sorted particles.  Rows and cells that
// cannot hold a particle closer than h are skipped: a row whose y/z distance
// to the particle is at least h, and, inside a row, the cells outside
// x +/- sqrt(h^2 - d_yz^2).  The test uses h (1 + 1e-3), far above the
// rounding of positions and cell bounds, so no particle with r < h is ever
// skipped and the neighbour set {j : r_ij^2 < h^2} is unchanged; only the
// number of candidates tested drops.
#pragma once

#include "internal.hpp"

namespace sph {

struct GridView {
    float ox, oy, oz, spacing, inv_spacing;
    int dx, dy, dz;
    float h;            // search radius
    float prune_h2;     // (h (1 + 1e-3))^2
    uint32_t cell_offset;  // first global cell of this grid (batched scenes share one cell array)
};

__host__ __device__ inline GridView make_grid_view(float ox, float oy, float oz, float spacing,
                                                   int dx, int dy, int dz, float h) {
    GridView g;
    g.ox = ox;
    g.oy = oy;
    g.oz = oz;
    g.spacing = spacing;
    g.inv_spacing = 1.0f / spacing;
    g.dx = dx;
    g.dy = dy;
    g.dz = dz;
    g.h = h;
    const float ph = h * 1.001f;
    g.prune_h2 = ph * ph;
    g.cell_offset = 0u;
    return g;
}

// Offset of a coordinate from the grid origin, clamped into the grid's extent
// [0, dim * sp], for the pruning tests below.  A particle outside the grid is
// hashed into a boundary cell (clamped cell coordinates) although it lies
// beyond that cell's slab, so its distances to the slabs, measured from where
// it is, would overstate the distance to its own cell's particles and to
// others outside the grid (such a particle's density once lost even its self
// term).  Measured from the nearest point of the grid they are lower bounds
// again, so the pruning stays conservative.  Inside the grid the value is
// unchanged, and so is every result.  A candidate outside the grid is always
// beyond its slab on the side away from the grid, so for a target inside the
// grid the slab distance already bounds its distance from below.
__device__ __forceinline__ float grid_offset(float x, float o, float sp, int dim) {
    return fminf(fmaxf(x - o, 0.0f), static_cast<float>(dim) * sp);
}

// Distance from f to the interval [c*sp, (c+1)*sp].
__device__ __forceinline__ float slab_gap(float f, int c, float sp) {
    const float lo = static_cast<float>(c) * sp;
    const float hi = lo + sp;
    return fmaxf(0.0f, fmaxf(lo - f, f - hi));
}

// Contiguous particle run [b, e) covering cells lo..hi of a row starting at
// hash `row`.  The starts and ends of three cells are loaded together (a
// stencil row of cells of edge h spans at most three; indices past hi repeat
// hi, which min/max absorb).  Empty cells hold SPH_EMPTY_CELL as start, so
// the run's start is the minimum start and its end the maximum end over the
// non-empty cells.  Cells with a start always have an end <= n; min() keeps
// the loop bound in range regardless.
__device__ __forceinline__ void row_run(const uint32_t* __restrict__ cs,
                                        const uint32_t* __restrict__ ce, uint32_t row, int lo,
                                        int hi, uint32_t n, uint32_t& b, uint32_t& e) {
    b = kEmptyCell;
    e = 0u;
    for (int cx = lo; cx <= hi; cx += 3) {
        const uint32_t i0 = row + cx;
        const uint32_t i1 = row + min(cx + 1, hi);
        const uint32_t i2 = row + min(cx + 2, hi);
        const uint32_t s0 = cs[i0], s1 = cs[i1], s2 = cs[i2];
        const uint32_t e0 = ce[i0], e1 = ce[i1], e2 = ce[i2];
        b = min(b, min(s0, min(s1, s2)));
        e = max(e, max(s0 != kEmptyCell ? e0 : 0u,
                       max(s1 != kEmptyCell ? e1 : 0u, s2 != kEmptyCell ? e2 : 0u)));
    }
    e = min(e, n);
}

// Calls f(b, e) for every non-empty candidate run around (x, y, z).
template <class F>
__device__ __forceinline__ void for_each_run(const GridView& g, const uint32_t* __restrict__ cs,
                                             const uint32_t* __restrict__ ce, uint32_t n, float x,
                                             float y, float z, F&& f) {
    const int sx = cell_coord_off(x, -g.h, g.ox, g.spacing, g.dx);
    const int sy = cell_coord_off(y, -g.h, g.oy, g.spacing, g.dy);
    const int sz = cell_coord_off(z, -g.h, g.oz, g.spacing, g.dz);
    const int ex = cell_coord_off(x, g.h, g.ox, g.spacing, g.dx);
    const int ey = cell_coord_off(y, g.h, g.oy, g.spacing, g.dy);
    const int ez = cell_coord_off(z, g.h, g.oz, g.spacing, g.dz);
    const float fx = grid_offset(x, g.ox, g.spacing, g.dx);
    const float fy = grid_offset(y, g.oy, g.spacing, g.dy);
    const float fz = grid_offset(z, g.oz, g.spacing, g.dz);
    for (int cz = sz; cz <= ez; ++cz) {
        const float gz = slab_gap(fz, cz, g.spacing);
        for (int cy = sy; cy <= ey; ++cy) {
            const float gy = slab_gap(fy, cy, g.spacing);
            const float d2 = gy * gy + gz * gz;
            if (d2 >= g.prune_h2) continue;
            const float w = __builtin_sqrtf(g.prune_h2 - d2);
            const int lo = max(sx, static_cast<int>(floorf((fx - w) * g.inv_spacing)));
            const int hi = min(ex, static_cast<int>(floorf((fx + w) * g.inv_spacing)));
            if (lo > hi) continue;
            uint32_t b, e;
            row_run(cs, ce, g.cell_offset + static_cast<uint32_t>(g.dx) * (cy + g.dy * cz), lo, hi,
                    n, b, e);
            if (b < e) f(b, e);
        }
    }
}

}  // namespace sph

namespace sph {

// Per-scene constants of a batch: the scene's step parameters and grid, whose
// cells start at g.cell_offset in the batch's global cell array.
struct SceneParams {
    StepParams p;
    GridView g;
};

// Batched-scene passes (step.hip / radix_sort.hip): cap.n is the launch
// capacity, cap.num_cells the batch's total cell count, cap.h/spacing the grid
// used when a pass needs no per-scene constants.
int hash_hist_cells_batched(const StepParams& cap, const uint32_t* n_dev, const float4* pos,
                            const float4* vel, const SceneParams* scenes, uint32_t* keys,
                            uint32_t key_bits, void* temp, size_t temp_bytes,
                            uint32_t* cell_start, uint32_t* cell_end, hipStream_t stream);
int launch_batch_pass(int pass, const StepParams& cap, const StepBuffers& b,
                      const SceneParams* scenes, bool kahan, bool tension, float dt,
                      uint32_t key_bits, bool write_accel, hipStream_t stream);

// Autotuner (step.hip): times the launch variants of the density and force
// passes on the current state (scenes == NULL: a single scene with p's
// constants) and returns the fastest tune word; ms_out (may be NULL) gets the
// default and tuned milliseconds of density and force.  pos/vel are restored.
bool tune_word_valid(uint32_t tune, bool batched);
int tune_neighbour_passes(StepParams p, const StepBuffers& b, const SceneParams* scenes,
                          bool kahan, bool tension, float dt, uint32_t key_bits, uint32_t reps,
                          hipStream_t stream, uint32_t* tune_out, float* ms_out);

// Per-particle view of the reference stencil: own cell, stencil bounds and
// position relative to the grid origin (clamped into the grid, grid_offset).
struct Probe {
    float fx, fy, fz;
    int cx, cy, cz;
    int sx, sy, sz, ex, ey, ez;
};

__device__ __forceinline__ Probe make_probe(const GridView& g, float x, float y, float z) {
    Probe pr;
    pr.sx = cell_coord_off(x, -g.h, g.ox, g.spacing, g.dx);
    pr.sy = cell_coord_off(y, -g.h, g.oy, g.spacing, g.dy);
    pr.sz = cell_coord_off(z, -g.h, g.oz, g.spacing, g.dz);
    pr.ex = cell_coord_off(x, g.h, g.ox, g.spacing, g.dx);
    pr.ey = cell_coord_off(y, g.h, g.oy, g.spacing, g.dy);
    pr.ez = cell_coord_off(z, g.h, g.oz, g.spacing, g.dz);
    pr.cx = cell_coord(x, g.ox, g.spacing, g.dx);
    pr.cy = cell_coord(y, g.oy, g.spacing, g.dy);
    pr.cz = cell_coord(z, g.oz, g.spacing, g.dz);
    pr.fx = grid_offset(x, g.ox, g.spacing, g.dx);
    pr.fy = grid_offset(y, g.oy, g.spacing, g.dy);
    pr.fz = grid_offset(z, g.oz, g.spacing, g.dz);
    return pr;
}

// Candidate run [b, e) of grid row (ry, rz) for the probe: empty when the row
// is outside the reference stencil or farther than h; otherwise the cells of
// the row that can hold a particle closer than h (see the file comment).
__device__ __forceinline__ void probe_row_run(const GridView& g, const Probe& pr,
                                              const uint32_t* __restrict__ cs,
                                              const uint32_t* __restrict__ ce, uint32_t n,
                                              int ry, int rz, uint32_t& b, uint32_t& e) {
    b = kEmptyCell;
    e = 0u;
    if (ry < pr.sy || ry > pr.ey || rz < pr.sz || rz > pr.ez) return;
    const float gy = slab_gap(pr.fy, ry, g.spacing);
    const float gz = slab_gap(pr.fz, rz, g.spacing);
    const float d2 = gy * gy + gz * gz;
    if (d2 >= g.prune_h2) return;
    const float w = __builtin_amdgcn_sqrtf(g.prune_h2 - d2);
    const int lo = max(pr.sx, static_cast<int>(floorf((pr.fx - w) * g.inv_spacing)));
    const int hi = min(pr.ex, static_cast<int>(floorf((pr.fx + w) * g.inv_spacing)));
    if (lo > hi) return;
    row_run(cs, ce, g.cell_offset + static_cast<uint32_t>(g.dx) * (ry + g.dy * rz), lo, hi, n, b,
            e);
    if (b >= e) {
        b = kEmptyCell;
        e = 0u;
    }
}

// x-bins: each cell's edge is cut into kXBins bins; the internal order keeps
// a cell's particles sorted by x, so each bin is a contiguous range of it.
// t is the position along x in cell units relative to the cell's low face.
// A cell's table holds the bins' cumulative counts, one byte each, in
// kXBinQuads 16-byte words (kXBinBytes per cell).  Cells of more than
// kXBinMax particles are left unbinned (walked whole): every byte of their
// table holds kXBinUnbinned, which no count reaches.  Against 8 bins the
// 32 cut the candidates per particle by about 6% (uniform positions: 12.7
// against 13.5 h^3 per target).
constexpr int kXBins = 32;
constexpr int kXBinQuads = kXBins / 16;
constexpr size_t kXBinBytes = 16 * kXBinQuads;
constexpr uint32_t kXBinMax = 0xFEu;
constexpr uint32_t kXBinUnbinned = 0xFFu;
__device__ __forceinline__ int xbin_of(float t) {
    return min(kXBins - 1, max(0, static_cast<int>(floorf(t * static_cast<float>(kXBins)))));
}

// Particles of cell `cell` in bins 0..k (kXBinUnbinned for an unbinned
// cell); reads the one 16-byte word that holds count k.
__device__ __forceinline__ uint32_t xbin_count(const uint4* __restrict__ bins, uint32_t cell,
                                               int k) {
    const uint4 v = bins[cell * kXBinQuads + static_cast<uint32_t>(k >> 4)];
    const int q = k & 15;
    const uint32_t w = (q < 4) ? v.x : (q < 8) ? v.y : (q < 12) ? v.z : v.w;
    return (w >> (8 * (q & 3))) & 0xFFu;
}

// probe_row_run in the internal x-order: the run additionally starts at the
// first x-bin of its first cell, and ends after the last x-bin of its last
// cell, that the x-window [x - w, x + w] reaches.  w carries the pruning
// margin h (1 + 1e-3), far above the rounding of bin positions, so no
// particle with r < h is dropped.
__device__ __forceinline__ void probe_row_run_x(const GridView& g, const Probe& pr,
                                                const uint32_t* __restrict__ cs,
                                                const uint32_t* __restrict__ ce,
                                                const uint4* __restrict__ bins, uint32_t n,
                                                int ry, int rz, uint32_t& b, uint32_t& e) {
    b = kEmptyCell;
    e = 0u;
    if (ry < pr.sy || ry > pr.ey || rz < pr.sz || rz > pr.ez) return;
    const float gy = slab_gap(pr.fy, ry, g.spacing);
    const float gz = slab_gap(pr.fz, rz, g.spacing);
    const float d2 = gy * gy + gz * gz;
    if (d2 >= g.prune_h2) return;
    const float w = __builtin_amdgcn_sqrtf(g.prune_h2 - d2);
    const float xl = (pr.fx - w) * g.inv_spacing, xh = (pr.fx + w) * g.inv_spacing;
    const int lo = max(pr.sx, static_cast<int>(floorf(xl)));
    const int hi = min(pr.ex, static_cast<int>(floorf(xh)));
    if (lo > hi) return;
    const uint32_t row = g.cell_offset + static_cast<uint32_t>(g.dx) * (ry + g.dy * rz);
    row_run(cs, ce, row, lo, hi, n, b, e);
    if (b >= e) {
        b = kEmptyCell;
        e = 0u;
        return;
    }
    const uint32_t s_lo = cs[row + lo], s_hi = cs[row + hi];
    if (s_lo != kEmptyCell) {  // the run starts in cell lo
        const int k = xbin_of(xl - static_cast<float>(lo));
        if (k > 0) {
            const uint32_t c = xbin_count(bins, row + lo, k - 1);
            if (c != kXBinUnbinned) b = s_lo + c;
        }
    }
    if (s_hi != kEmptyCell) {  // the run ends in cell hi
        const int k = xbin_of(xh - static_cast<float>(hi));
        const uint32_t c = xbin_count(bins, row + hi, k);
        if (c != kXBinUnbinned) e = min(e, s_hi + c);
    }
    if (b >= e) {
        b = kEmptyCell;
        e = 0u;
    }
}

template <bool kX>
__device__ __forceinline__ void probe_row(const GridView& g, const Probe& pr,
                                          const uint32_t* __restrict__ cs,
                                          const uint32_t* __restrict__ ce,
                                          const uint4* __restrict__ bins, uint32_t n, int ry,
                                          int rz, uint32_t& b, uint32_t& e) {
    if (kX)
        probe_row_run_x(g, pr, cs, ce, bins, n, ry, rz, b, e);
    else
        probe_row_run(g, pr, cs, ce, n, ry, rz, b, e);
}

// Block-wide [min b, max e) over the lanes' runs (empty runs contribute
// b = SPH_EMPTY_CELL, e = 0).  red must hold 2 * (blockDim / 64) words.
template <int kBlockThreads>
__device__ __forceinline__ void block_run_union(uint32_t b, uint32_t e, uint32_t* red,
                                                uint32_t& ub, uint32_t& ue) {
    // Wave-wide min / max (DPP-based device-library reductions, every lane
    // active), then across the block's waves through LDS.
    b = __ockl_wfred_min_u32(b);
    e = __ockl_wfred_max_u32(e);
    constexpr int kW = kBlockThreads / 64;
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wave] = b;
        red[kW + wave] = e;
    }
    __syncthreads();
    ub = red[0];
    ue = red[kW];
#pragma unroll
    for (int w = 1; w < kW; ++w) {
        ub = min(ub, red[w]);
        ue = max(ue, red[kW + w]);
    }
}

}  // namespace sph
