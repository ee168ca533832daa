// walls.hip -- boundary-particle walls: static wall particles, hashed and
// sorted once on a fluid grid, and the per-step boundary acceleration of every
// fluid particle.
//
// Reference: SPH2S_FINAL_SINGLE_SCALE/Solver.cu -- evaluateBoundaryWeight
// (:169-196), computeBoundaryForceCell (:311-344), the boundary stencil walk
// and a_i = f_i / rho_i + b_i + g (:687-710), and the one-time pre-sort in the
// constructor (:1003-1059: computeHashs, thrust::sort_by_key,
// reorderComputeCellStartEndBoundaryD :426-476).  The two-scale solver does the
// same on its low-resolution grid with h0 and m0 for both sets
// (SPH2S_FINAL_TWO_SCALE/Solver.cu:454-499, :1007-1044).
//
//   b_i = sum_{j : r_ij < h} B(r_ij) m_b / (m + m_b) x_ij / r_ij
//   B(r) = 0.02 c^2 / r * { 2/3               q < 2/3
//                           2 q - 3/2 q^2     q < 1
//                           1/2 (2 - q)^2     q < 2 },   q = 2 r / h
//
// The wall set is static, so each grid cell also gets a flag telling whether a
// wall particle lies within reach of any point of it; a fluid particle in an
// unflagged cell (most of them) writes b = 0 without walking the stencil.  A
// pair at r = 0, where the reference divides by zero, is skipped.
#include <algorithm>
#include <cmath>
#include <vector>

#include "neighbors.hpp"
#include "walls.hpp"

namespace sph {
namespace {

constexpr int kWallBlock = 256;

inline dim3 wall_blocks(uint32_t n) { return dim3(std::max(1u, (n + kWallBlock - 1) / kWallBlock)); }

__global__ void k_wall_pack(const float* __restrict__ xyz, uint32_t n, GridView g,
                            float4* __restrict__ pos, uint32_t* __restrict__ keys) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    pos[i] = make_float4(x, y, z, 0.0f);
    keys[i] = cell_hash(x, y, z, g.ox, g.oy, g.oz, g.spacing, g.dx, g.dy, g.dz);
}

// Gather into sorted order and record each cell's [start, end)
// (reorderComputeCellStartEndBoundaryD, Solver.cu:426-476).
__global__ void k_wall_reorder(const uint32_t* __restrict__ skeys, const uint32_t* __restrict__ perm,
                               const float4* __restrict__ pos, uint32_t n,
                               float4* __restrict__ spos, uint32_t* __restrict__ cs,
                               uint32_t* __restrict__ ce) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    spos[i] = pos[perm[i]];
    const uint32_t k = skeys[i];
    if (i == 0 || skeys[i - 1] != k) cs[k] = i;
    if (i == n - 1 || skeys[i + 1] != k) ce[k] = i + 1;
}

__global__ void k_fill_u32(uint32_t* __restrict__ w, uint32_t count, uint32_t value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        w[i] = value;
}

// Flags every cell within `reach` cells (per axis) of a wall particle's cell.
__global__ void k_wall_near(const float4* __restrict__ spos, uint32_t n, GridView g, int reach,
                            uint32_t* __restrict__ near) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = spos[i];
    const int cx = cell_coord(p.x, g.ox, g.spacing, g.dx);
    const int cy = cell_coord(p.y, g.oy, g.spacing, g.dy);
    const int cz = cell_coord(p.z, g.oz, g.spacing, g.dz);
    for (int z = max(0, cz - reach); z <= min(g.dz - 1, cz + reach); ++z)
        for (int y = max(0, cy - reach); y <= min(g.dy - 1, cy + reach); ++y)
            for (int x = max(0, cx - reach); x <= min(g.dx - 1, cx + reach); ++x)
                near[x + g.dx * (y + g.dy * z)] = 1u;
}

// evaluateBoundaryWeight (Solver.cu:169-196); r > 0, r < h.
__device__ __forceinline__ float wall_weight(float r, float h, float coef) {
    const float q = 2.0f * r / h;
    const float c = coef / r;
    if (q < 2.0f / 3.0f) return c * (2.0f / 3.0f);
    if (q < 1.0f) return c * (2.0f * q - 1.5f * q * q);
    if (q < 2.0f) {
        const float a = 2.0f - q;
        return c * 0.5f * a * a;
    }
    return 0.0f;
}

__global__ __launch_bounds__(kWallBlock) void k_wall_accel(
    const float4* __restrict__ tpos, const uint32_t* __restrict__ n_dev, GridView g,
    const float4* __restrict__ wpos, const uint32_t* __restrict__ wcs,
    const uint32_t* __restrict__ wce, uint32_t wn, const uint32_t* __restrict__ near, float coef,
    float mratio, float4* __restrict__ bacc) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    const float4 x = tpos[i];
    float bx = 0.f, by = 0.f, bz = 0.f;
    if (near[cell_hash(x.x, x.y, x.z, g.ox, g.oy, g.oz, g.spacing, g.dx, g.dy, g.dz)]) {
        const float h2 = g.h * g.h;
        for_each_run(g, wcs, wce, wn, x.x, x.y, x.z, [&](uint32_t b, uint32_t e) {
            for (uint32_t j = b; j < e; ++j) {
                const float4 w = wpos[j];
                const float dx = x.x - w.x, dy = x.y - w.y, dz = x.z - w.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < h2 && r2 > 0.0f) {
                    const float r = __fsqrt_rn(r2);
                    const float s = wall_weight(r, g.h, coef) * mratio / r;
                    bx += s * dx;
                    by += s * dy;
                    bz += s * dz;
                }
            }
        });
    }
    bacc[i] = make_float4(bx, by, bz, 0.0f);
}

size_t al256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

}  // namespace

void walls_destroy(Walls* w) {
    if (!w) return;
    if (w->arena) (void)hipFree(w->arena);
    *w = Walls{};
}

int walls_create(Walls* w, const sph_grid& grid, float h, float boundary_mass, float fluid_mass,
                 float speed_sound, const float* pos_xyz, uint32_t n, hipStream_t st) {
    walls_destroy(w);
    if (n == 0) return SPH_OK;
    if (!pos_xyz) return set_error(SPH_ERR_INVALID_ARGUMENT, "boundary positions are null");
    if (!(boundary_mass > 0.0f) || !(speed_sound > 0.0f) || !std::isfinite(speed_sound))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "boundary: mass and speed of sound must be > 0");
    const uint32_t C = sph_grid_num_cells(&grid);
    const size_t temp = radix_sort_temp_bytes(n);
    // wall arena: spos, cs, ce, near | staging: xyz, pos, keys, skeys, perm, sort temp
    const size_t keep = al256(16ull * n) + 3 * al256(4ull * C);
    const size_t stage = al256(12ull * n) + al256(16ull * n) + 3 * al256(4ull * n) + al256(temp);
    void* arena = nullptr;
    SPH_HIP(hipMalloc(&arena, keep + stage));
    char* a = static_cast<char*>(arena);
    w->arena = arena;
    w->n = n;
    w->num_cells = C;
    w->pos = reinterpret_cast<float4*>(a);
    a += al256(16ull * n);
    w->cs = reinterpret_cast<uint32_t*>(a);
    a += al256(4ull * C);
    w->ce = reinterpret_cast<uint32_t*>(a);
    a += al256(4ull * C);
    w->near = reinterpret_cast<uint32_t*>(a);
    a += al256(4ull * C);
    float* xyz = reinterpret_cast<float*>(a);
    a += al256(12ull * n);
    float4* upos = reinterpret_cast<float4*>(a);
    a += al256(16ull * n);
    uint32_t* keys = reinterpret_cast<uint32_t*>(a);
    a += al256(4ull * n);
    uint32_t* skeys = reinterpret_cast<uint32_t*>(a);
    a += al256(4ull * n);
    uint32_t* perm = reinterpret_cast<uint32_t*>(a);
    a += al256(4ull * n);
    void* sort_temp = a;

    w->h = h;
    w->coef = 0.02f * speed_sound * speed_sound;
    w->mratio = boundary_mass / (fluid_mass + boundary_mass);
    w->grid = make_grid_view(grid.origin[0], grid.origin[1], grid.origin[2], grid.spacing,
                             grid.dims[0], grid.dims[1], grid.dims[2], h);
    // Stencil bounds reach ceil(h / spacing) cells from the own cell, one more
    // for rounding of the bounds.
    const int reach = static_cast<int>(std::ceil(h / grid.spacing)) + 1;
    const uint32_t key_bits = key_bits_for_cells(C);

    const dim3 fill_grid(std::min<uint32_t>((C + kWallBlock - 1) / kWallBlock, 4096u));
    int rc = SPH_OK;
    auto fail = [&](int code) {
        walls_destroy(w);
        return code;
    };
    hipError_t e = hipMemcpyAsync(xyz, pos_xyz, 12ull * n, hipMemcpyHostToDevice, st);
    // The sort's error word is sticky (the sort never clears it) and the
    // arena is fresh: clear it so that the check below sees this sort only.
    if (e == hipSuccess) e = hipMemsetAsync(radix_sort_error_word(sort_temp), 0, 4, st);
    if (e != hipSuccess) return fail(hip_error(e, "hipMemcpyAsync", __FILE__, __LINE__));
    hipLaunchKernelGGL(k_wall_pack, wall_blocks(n), dim3(kWallBlock), 0, st, xyz, n, w->grid, upos,
                       keys);
    hipLaunchKernelGGL(k_fill_u32, fill_grid, dim3(kWallBlock), 0, st, w->cs, C, kEmptyCell);
    hipLaunchKernelGGL(k_fill_u32, fill_grid, dim3(kWallBlock), 0, st, w->ce, C, kEmptyCell);
    hipLaunchKernelGGL(k_fill_u32, fill_grid, dim3(kWallBlock), 0, st, w->near, C, 0u);
    rc = radix_sort_pairs(keys, nullptr, skeys, perm, n, key_bits, sort_temp, temp, st);
    if (rc != SPH_OK) return fail(rc);
    hipLaunchKernelGGL(k_wall_reorder, wall_blocks(n), dim3(kWallBlock), 0, st, skeys, perm, upos,
                       n, w->pos, w->cs, w->ce);
    hipLaunchKernelGGL(k_wall_near, wall_blocks(n), dim3(kWallBlock), 0, st, w->pos, n, w->grid,
                       reach, w->near);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(hip_error(e, "wall setup", __FILE__, __LINE__));
    uint32_t sort_err = 0;
    e = hipMemcpyAsync(&sort_err, radix_sort_error_word(sort_temp), 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return fail(hip_error(e, "hipMemcpy", __FILE__, __LINE__));
    if (sort_err) return fail(set_error(SPH_ERR_HIP, "boundary sort: look-back wait gave up"));
    return SPH_OK;
}

int launch_wall_accel(const Walls& w, const float4* tpos, const uint32_t* n_dev, uint32_t cap,
                      float4* bacc, hipStream_t st) {
    if (!w.n || cap == 0) return SPH_OK;
    hipLaunchKernelGGL(k_wall_accel, wall_blocks(cap), dim3(kWallBlock), 0, st, tpos, n_dev, w.grid,
                       w.pos, w.cs, w.ce, w.n, w.near, w.coef, w.mratio, bacc);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

}  // namespace sph
