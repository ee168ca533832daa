// diagnostics.hip -- the renderer-facing side of the solvers: the Mexican-hat
// energy colour value, the export of the renderer's buffers, and the surface
// marks that request splits.
//
// Reference: SPH2S_FINAL_TWO_SCALE/Solver.cu -- evaluateMexicanHat3D
// (:206-213), the accumulation of psi, velW, xc and the mass in
// computeAccelerationCell (:376-387) starting from the particle itself
// (:897-907), and the colour value written to dColorValues (:1050-1075);
// ParticleData's renderer buffers dPositions / dColorValues / dStates
// (TWO_SCALE/ParticleData.h:39-45, drawn by Renderer.cpp:38-46); the
// screen-space surface pass getSurfaceIDs (SSFRenderer.cu:196-218, launched
// at :735-743), which sets state bit 8 on the particles whose index the
// surface index map holds, and the split request on such particles that the
// reference leaves commented out (Solver.cu:1077-1082).
#include <algorithm>
#include <cmath>

#include "diagnostics.hpp"

namespace sph {
namespace {

constexpr int kDiagBlock = 256;

inline dim3 diag_blocks(uint32_t n) { return dim3(std::max(1u, (n + kDiagBlock - 1) / kDiagBlock)); }

// 2 / (sqrt(3) pi^(1/4)), the reference's MEXICAN_HAT_C.
constexpr float kMexicanHatC = 0.8673250705840776f;

// evaluateMexicanHat3D (Solver.cu:206-213): C (|u|^2 - 3) exp(-|u|^2 / 2).
__device__ __forceinline__ float mexican_hat(float x, float y, float z) {
    const float s = x * x + y * y + z * z;
    return kMexicanHatC * (s - 3.0f) * expf(-0.5f * s);
}

// Colour value of every particle of one set from the step's sorted state:
//   psi_j = hat(x_ij / h) over the particle itself and its own-set neighbours
//   with 0 < r < h (no blend weights); velW = sum psi_j v_j, psiSum = sum psi_j,
//   x_c = mass-weighted mean position;
//   ene = min(|velW|^2 / (psiSum^2 h), 300);
//   colour = 0.3 where |x_i - x_c| < 0.0025, else ene / 300.
__global__ __launch_bounds__(kDiagBlock) void k_energy_color(
    const float4* __restrict__ tpos, const float4* __restrict__ tvel,
    const uint32_t* __restrict__ cs, const uint32_t* __restrict__ ce,
    const uint32_t* __restrict__ n_dev, GridView g, float mass, float* __restrict__ color) {
    const uint32_t n = *n_dev;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 xi = tpos[i];
    const float4 vi = tvel[i];
    const float h = g.h, inv_h = 1.0f / g.h, h2 = g.h * g.h;
    const float psi0 = mexican_hat(0.0f, 0.0f, 0.0f);
    float psi_sum = psi0;
    float wx = vi.x * psi0, wy = vi.y * psi0, wz = vi.z * psi0;
    float msum = mass;
    float cx = mass * xi.x, cy = mass * xi.y, cz = mass * xi.z;
    for_each_run(g, cs, ce, n, xi.x, xi.y, xi.z, [&](uint32_t b, uint32_t e) {
        for (uint32_t j = b; j < e; ++j) {
            const float4 xj = tpos[j];
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 0.0f && r2 < h2) {
                const float4 vj = tvel[j];
                const float psi = mexican_hat(dx * inv_h, dy * inv_h, dz * inv_h);
                wx += vj.x * psi;
                wy += vj.y * psi;
                wz += vj.z * psi;
                psi_sum += psi;
                cx += mass * xj.x;
                cy += mass * xj.y;
                cz += mass * xj.z;
                msum += mass;
            }
        }
    });
    cx /= msum;
    cy /= msum;
    cz /= msum;
    const float ddx = xi.x - cx, ddy = xi.y - cy, ddz = xi.z - cz;
    const float dist_cm = __fsqrt_rn(ddx * ddx + ddy * ddy + ddz * ddz);
    constexpr float kMaxEne = 300.0f;
    float ene = (wx * wx + wy * wy + wz * wz) / (psi_sum * psi_sum * h);
    ene = fminf(ene, kMaxEne);
    color[i] = (dist_cm < 0.0025f) ? 0.3f : ene / kMaxEne;
}

__global__ void k_export_render(const float4* __restrict__ pos, const float2* __restrict__ blend,
                                const float* __restrict__ color_in, const uint32_t* __restrict__ n_dev,
                                uint32_t default_state, float* __restrict__ xyz,
                                float* __restrict__ color, int32_t* __restrict__ state) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    if (xyz) {
        const float4 p = pos[i];
        xyz[3 * i] = p.x;
        xyz[3 * i + 1] = p.y;
        xyz[3 * i + 2] = p.z;
    }
    if (color) color[i] = color_in[i];
    if (state) state[i] = static_cast<int32_t>(blend ? __float_as_uint(blend[i].y) : default_state);
}

// Spiky-kernel gradient magnitude over r: c45 (h - r)^2 / r for 0 < r < h.
__device__ __forceinline__ float spiky_over_r(float r, float h, float c45) {
    const float a = h - r;
    return c45 * a * a / r;
}

// Colour-field gradient of the low set from the step's sorted state of both
// sets:  n_i = sum_{L, 0<r<h0} m0 B_j grad W_h0 + sum_{H, 0<r<h0} m1 B_j
// 1/2 (grad W_h0 + grad W_h1), B = 1/rho.  A low particle with |n_i| h0 above
// the threshold is at the free surface: its id is flagged.
__global__ __launch_bounds__(kDiagBlock) void k_surface_low(
    const float4* __restrict__ tpos_l, const float2* __restrict__ aux_l,
    const uint32_t* __restrict__ cs_l, const uint32_t* __restrict__ ce_l,
    const uint32_t* __restrict__ n_l, GridView gl, const float4* __restrict__ tpos_h,
    const float2* __restrict__ aux_h, const uint32_t* __restrict__ cs_h,
    const uint32_t* __restrict__ ce_h, const uint32_t* __restrict__ n_h, GridView gh,
    SurfaceParams sp, uint32_t* __restrict__ flags, uint32_t num_ids) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nl = *n_l, nh = *n_h;
    if (i >= nl) return;
    const float4 xi = tpos_l[i];
    const float h0 = sp.h0, h1 = sp.h1, h02 = h0 * h0;
    float nx = 0.f, ny = 0.f, nz = 0.f;
    for_each_run(gl, cs_l, ce_l, nl, xi.x, xi.y, xi.z, [&](uint32_t b, uint32_t e) {
        for (uint32_t j = b; j < e; ++j) {
            const float4 xj = tpos_l[j];
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 0.0f && r2 < h02) {
                const float r = __fsqrt_rn(r2);
                const float w = sp.m0 * aux_l[j].y * spiky_over_r(r, h0, sp.c45_0);
                nx -= w * dx;
                ny -= w * dy;
                nz -= w * dz;
            }
        }
    });
    if (nh > 0)
        for_each_run(gh, cs_h, ce_h, nh, xi.x, xi.y, xi.z, [&](uint32_t b, uint32_t e) {
            for (uint32_t j = b; j < e; ++j) {
                const float4 xj = tpos_h[j];
                const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > 0.0f && r2 < h02) {
                    const float r = __fsqrt_rn(r2);
                    const float g = spiky_over_r(r, h0, sp.c45_0) +
                                    ((r < h1) ? spiky_over_r(r, h1, sp.c45_1) : 0.0f);
                    const float w = sp.m1 * aux_h[j].y * 0.5f * g;
                    nx -= w * dx;
                    ny -= w * dy;
                    nz -= w * dz;
                }
            }
        });
    const float mag = __fsqrt_rn(nx * nx + ny * ny + nz * nz) * h0;
    const uint32_t id = __float_as_uint(xi.w);
    if (mag > sp.threshold && id < num_ids) flags[id] = 1u;
}

// Surface marks on the compacted low set: bit 8 from the flags (cleared after
// use); with `split`, a settled particle that carries bit 8 and nothing else
// (state 8) also gets mode 3, the reference's commented-out request.
__global__ void k_apply_surface_flags(const float4* __restrict__ pos,
                                      const uint32_t* __restrict__ n_dev,
                                      uint32_t* __restrict__ flags, uint32_t num_ids,
                                      uint32_t split, float2* __restrict__ blend) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    const uint32_t id = __float_as_uint(pos[i].w);
    if (id >= num_ids || !flags[id]) return;
    flags[id] = 0u;
    const float2 b = blend[i];
    uint32_t st = __float_as_uint(b.y) | 8u;
    if (split && st == 8u) st |= 3u;
    blend[i] = make_float2(b.x, __uint_as_float(st));
}

// getSurfaceIDs (SSFRenderer.cu:196-218): every non-zero entry v of the index
// map names low-set slot v - 1, which gets bit 8 (and the split request under
// `split`, as above).
__global__ void k_mark_surface_map(const uint32_t* __restrict__ map, uint32_t count,
                                   const uint32_t* __restrict__ n_dev, uint32_t split,
                                   float2* __restrict__ blend) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint32_t v = map[k];
    if (v == 0u || v > *n_dev) return;
    // Several map entries may name one slot; the update is idempotent.
    const float2 b = blend[v - 1u];
    uint32_t st = __float_as_uint(b.y) | 8u;
    if (split && st == 8u) st |= 3u;
    blend[v - 1u] = make_float2(b.x, __uint_as_float(st));
}

}  // namespace

int launch_energy_color(const float4* tpos, const float4* tvel, const uint32_t* cs,
                        const uint32_t* ce, const uint32_t* n_dev, uint32_t cap, const GridView& g,
                        float mass, float* color, hipStream_t st) {
    if (cap == 0) return SPH_OK;
    hipLaunchKernelGGL(k_energy_color, diag_blocks(cap), dim3(kDiagBlock), 0, st, tpos, tvel, cs,
                       ce, n_dev, g, mass, color);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

int launch_export_render(const float4* pos, const float2* blend, const float* color_in,
                         const uint32_t* n_dev, uint32_t cap, uint32_t default_state, float* xyz,
                         float* color, int32_t* state, hipStream_t st) {
    if (cap == 0) return SPH_OK;
    hipLaunchKernelGGL(k_export_render, diag_blocks(cap), dim3(kDiagBlock), 0, st, pos, blend,
                       color_in, n_dev, default_state, xyz, color, state);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

int launch_surface_detect(const float4* tpos_l, const float2* aux_l, const uint32_t* cs_l,
                          const uint32_t* ce_l, const uint32_t* n_l, uint32_t cap_l,
                          const GridView& gl, const float4* tpos_h, const float2* aux_h,
                          const uint32_t* cs_h, const uint32_t* ce_h, const uint32_t* n_h,
                          const GridView& gh, const SurfaceParams& sp, uint32_t* flags,
                          uint32_t num_ids, hipStream_t st) {
    if (cap_l == 0) return SPH_OK;
    hipLaunchKernelGGL(k_surface_low, diag_blocks(cap_l), dim3(kDiagBlock), 0, st, tpos_l, aux_l,
                       cs_l, ce_l, n_l, gl, tpos_h, aux_h, cs_h, ce_h, n_h, gh, sp, flags, num_ids);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

int launch_apply_surface(const float4* pos, const uint32_t* n_dev, uint32_t cap, uint32_t* flags,
                         uint32_t num_ids, uint32_t split, float2* blend, hipStream_t st) {
    if (cap == 0) return SPH_OK;
    hipLaunchKernelGGL(k_apply_surface_flags, diag_blocks(cap), dim3(kDiagBlock), 0, st, pos, n_dev,
                       flags, num_ids, split, blend);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

int launch_mark_surface_map(const uint32_t* map, uint32_t count, const uint32_t* n_dev,
                            uint32_t split, float2* blend, hipStream_t st) {
    if (count == 0) return SPH_OK;
    hipLaunchKernelGGL(k_mark_surface_map, diag_blocks(count), dim3(kDiagBlock), 0, st, map, count,
                       n_dev, split, blend);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

}  // namespace sph
