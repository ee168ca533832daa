// step.hip -- the single-scale SPH step on gfx950.
//
// Pass order of Solver::Advance (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1082-1198):
//   k_hash_reset   cell hash per particle + empty-cell reset of the ranges
//                  (computeHashs :401-424 and the two cudaMemsets :1118-1127)
//   radix sort     stable (hash, slot) sort          (thrust, :1110-1115)
//   k_reorder      gather of x, v into sorted slots and cell start/end
//                  (reorderAndComputeCellStartEndD :478-534)
//   k_density      Poly6 density                    (:536-599, :198-229)
//   k_force        spiky pressure + viscosity Laplacian (+ tension) fused with
//                  symplectic Euler and the wall clamp (:601-711, :713-755)
//
// State layout: float4 pos = (x, y, z, persistent id bits), float4 vel =
// (vx, vy, vz, 0), one 16-byte load per particle.  Within one row (cy, cz)
// of the grid the cells cx are consecutive hashes, so a stencil row of up to
// three cells is one contiguous run of sorted particles.
#include "internal.hpp"

namespace sph {
namespace {

constexpr int kBlock = 256;

// computeCoordinatesOff (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:20-38), FP32:
// clamp((int)((x + off - o) / spacing), 0, D-1).  Division is IEEE (the
// library is built with -fhip-fp32-correctly-rounded-divide-sqrt) so hashes
// and stencil bounds are bit-identical to the FP32 reference arithmetic.
__device__ __forceinline__ int cell_coord_off(float x, float off, float o, float sp, int dim) {
    float t = __fadd_rn(x, off);
    t = __fsub_rn(t, o);
    t = __fdiv_rn(t, sp);
    const int c = static_cast<int>(t);
    return max(0, min(c, dim - 1));
}

__device__ __forceinline__ int cell_coord(float x, float o, float sp, int dim) {
    float t = __fsub_rn(x, o);
    t = __fdiv_rn(t, sp);
    const int c = static_cast<int>(t);
    return max(0, min(c, dim - 1));
}

__global__ __launch_bounds__(kBlock) void k_hash_reset(const float4* __restrict__ pos,
                                                       StepParams p,
                                                       uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ cell_start,
                                                       uint32_t* __restrict__ cell_end) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < p.n) {
        const float4 x = pos[i];
        const int cx = cell_coord(x.x, p.ox, p.spacing, p.dx);
        const int cy = cell_coord(x.y, p.oy, p.spacing, p.dy);
        const int cz = cell_coord(x.z, p.oz, p.spacing, p.dz);
        keys[i] = static_cast<uint32_t>(cx + p.dx * (cy + p.dy * cz));
    }
    if (i < p.num_cells) {
        cell_start[i] = kEmptyCell;
        cell_end[i] = kEmptyCell;
    }
}

// Boundary detection compares each sorted key with its predecessor read
// straight from global memory (no block barrier, so no divergent
// __syncthreads as in the reference at Solver.cu:494-518).
__global__ __launch_bounds__(kBlock) void k_reorder(const uint32_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ perm,
                                                    const float4* __restrict__ pos,
                                                    const float4* __restrict__ vel,
                                                    float4* __restrict__ tpos,
                                                    float4* __restrict__ tvel,
                                                    uint32_t* __restrict__ cell_start,
                                                    uint32_t* __restrict__ cell_end, uint32_t n,
                                                    uint32_t num_cells) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n) return;
    // Keys are clamped cell hashes (< num_cells) and perm is a permutation of
    // [0, n); the min() guards keep every access in bounds regardless.
    const uint32_t k = min(skeys[s], num_cells - 1);
    if (s == 0) {
        cell_start[k] = 0;
    } else {
        const uint32_t kp = min(skeys[s - 1], num_cells - 1);
        if (k != kp) {
            cell_start[k] = s;
            cell_end[kp] = s;
        }
    }
    if (s == n - 1) cell_end[k] = n;
    const uint32_t j = min(perm[s], n - 1);
    tpos[s] = pos[j];
    tvel[s] = vel[j];
}

struct Stencil {
    int sx, sy, sz, ex, ey, ez;
};

__device__ __forceinline__ Stencil stencil_of(const StepParams& p, float x, float y, float z) {
    Stencil st;
    st.sx = cell_coord_off(x, -p.h, p.ox, p.spacing, p.dx);
    st.sy = cell_coord_off(y, -p.h, p.oy, p.spacing, p.dy);
    st.sz = cell_coord_off(z, -p.h, p.oz, p.spacing, p.dz);
    st.ex = cell_coord_off(x, p.h, p.ox, p.spacing, p.dx);
    st.ey = cell_coord_off(y, p.h, p.oy, p.spacing, p.dy);
    st.ez = cell_coord_off(z, p.h, p.oz, p.spacing, p.dz);
    return st;
}

// Contiguous particle run [b, e) covering cells sx..ex of row (cy, cz).
// Cells with a start always have an end <= n; the final min() keeps the loop
// in bounds regardless.
__device__ __forceinline__ void row_range(const uint32_t* __restrict__ cs,
                                          const uint32_t* __restrict__ ce, uint32_t row,
                                          int sx, int ex, uint32_t n, uint32_t& b, uint32_t& e) {
    b = kEmptyCell;
    e = 0u;
    for (int cx = sx; cx <= ex; ++cx) {
        const uint32_t st = cs[row + cx];
        if (st != kEmptyCell) {
            b = min(b, st);
            e = max(e, ce[row + cx]);
        }
    }
    e = min(e, n);
}

__global__ __launch_bounds__(kBlock) void k_density(const float4* __restrict__ tpos,
                                                    const uint32_t* __restrict__ cs,
                                                    const uint32_t* __restrict__ ce,
                                                    StepParams p, float* __restrict__ rho) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= p.n) return;
    const float4 xi = tpos[s];
    const Stencil st = stencil_of(p, xi.x, xi.y, xi.z);
    const bool kahan = (p.flags & SPH_FLAG_KAHAN) != 0u;
    float sum = 0.0f, comp = 0.0f;
    for (int cz = st.sz; cz <= st.ez; ++cz) {
        for (int cy = st.sy; cy <= st.ey; ++cy) {
            const uint32_t row = static_cast<uint32_t>(p.dx) * (cy + p.dy * cz);
            uint32_t b, e;
            row_range(cs, ce, row, st.sx, st.ex, p.n, b, e);
            for (uint32_t j = b; j < e; ++j) {
                const float4 xj = tpos[j];
                const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < p.h2) {
                    const float a = p.h2 - r2;
                    const float w = a * a * a;
                    if (kahan) {
                        const float y = w - comp;
                        const float t = sum + y;
                        comp = (t - sum) - y;
                        sum = t;
                    } else {
                        sum += w;
                    }
                }
            }
        }
    }
    rho[s] = p.c6m * sum;
}

// a_i = m c45 sum (A_i + A_j)(h-r)^2/r x_ij            (pressure, A = p/rho^2)
//     + (mu m c45 / rho_i) sum (v_j - v_i)(h-r)/rho_j  (viscosity)
//     - (sigma m c6 / rho_i) sum (h^2-r^2)^3 x_ij      (tension)
//     + g
// which is f_i/rho_i of Solver.cu:231-309 with rho_i factored out.
template <bool kTension>
__global__ __launch_bounds__(kBlock) void k_force(const float4* __restrict__ tpos,
                                                  const float4* __restrict__ tvel,
                                                  const float* __restrict__ rho,
                                                  const uint32_t* __restrict__ cs,
                                                  const uint32_t* __restrict__ ce, StepParams p,
                                                  float dt, float4* __restrict__ pos,
                                                  float4* __restrict__ vel,
                                                  float4* __restrict__ accel) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= p.n) return;
    const float4 xi = tpos[s];
    const float4 vi = tvel[s];
    const float rhoi = rho[s];
    const float inv_rhoi = __builtin_amdgcn_rcpf(rhoi);
    const float Ai = p.stiffness * (rhoi - p.rest_density) * inv_rhoi * inv_rhoi;
    const Stencil st = stencil_of(p, xi.x, xi.y, xi.z);
    float Px = 0.f, Py = 0.f, Pz = 0.f;
    float Vx = 0.f, Vy = 0.f, Vz = 0.f;
    float Tx = 0.f, Ty = 0.f, Tz = 0.f;
    for (int cz = st.sz; cz <= st.ez; ++cz) {
        for (int cy = st.sy; cy <= st.ey; ++cy) {
            const uint32_t row = static_cast<uint32_t>(p.dx) * (cy + p.dy * cz);
            uint32_t b, e;
            row_range(cs, ce, row, st.sx, st.ex, p.n, b, e);
            for (uint32_t j = b; j < e; ++j) {
                const float4 xj = tpos[j];
                const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < p.h2 && r2 > 0.0f) {
                    const float4 vj = tvel[j];
                    const float rhoj = rho[j];
                    const float inv_rhoj = __builtin_amdgcn_rcpf(rhoj);
                    const float Aj = p.stiffness * (rhoj - p.rest_density) * inv_rhoj * inv_rhoj;
                    const float inv_r = __builtin_amdgcn_rsqf(r2);
                    const float r = r2 * inv_r;
                    const float hr = p.h - r;
                    const float tp = (Ai + Aj) * hr * hr * inv_r;
                    Px += tp * dx;
                    Py += tp * dy;
                    Pz += tp * dz;
                    const float tv = hr * inv_rhoj;
                    Vx += tv * (vj.x - vi.x);
                    Vy += tv * (vj.y - vi.y);
                    Vz += tv * (vj.z - vi.z);
                    if (kTension) {
                        const float a = p.h2 - r2;
                        const float w = a * a * a;
                        Tx += w * dx;
                        Ty += w * dy;
                        Tz += w * dz;
                    }
                }
            }
        }
    }
    const float vc = p.visc * inv_rhoi;
    float ax = p.c45m * Px + vc * Vx + p.gx;
    float ay = p.c45m * Py + vc * Vy + p.gy;
    float az = p.c45m * Pz + vc * Vz + p.gz;
    if (kTension) {
        const float tc = p.tens * inv_rhoi;
        ax -= tc * Tx;
        ay -= tc * Ty;
        az -= tc * Tz;
    }
    // integrateD (Solver.cu:713-755): v += dt a; x += dt v.
    float vx = vi.x + dt * ax, vy = vi.y + dt * ay, vz = vi.z + dt * az;
    float x = xi.x + dt * vx, y = xi.y + dt * vy, z = xi.z + dt * vz;
    if (p.flags & SPH_FLAG_WALL_CLAMP) {
        const float e = p.restitution;
        if (x < p.lox) { x = p.lox; vx = -e * vx; }
        if (x > p.hix) { x = p.hix; vx = -e * vx; }
        if (y < p.loy) { y = p.loy; vy = -e * vy; }
        if (y > p.hiy) { y = p.hiy; vy = -e * vy; }
        if (z < p.loz) { z = p.loz; vz = -e * vz; }
        if (z > p.hiz) { z = p.hiz; vz = -e * vz; }
    }
    pos[s] = make_float4(x, y, z, xi.w);
    vel[s] = make_float4(vx, vy, vz, 0.0f);
    if (accel) accel[s] = make_float4(ax, ay, az, 0.0f);
}

__global__ void k_pack_state(const float* __restrict__ pxyz, const float* __restrict__ vxyz,
                             uint32_t n, float4* __restrict__ pos, float4* __restrict__ vel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos[i] = make_float4(pxyz[3 * i], pxyz[3 * i + 1], pxyz[3 * i + 2], __uint_as_float(i));
    if (vxyz)
        vel[i] = make_float4(vxyz[3 * i], vxyz[3 * i + 1], vxyz[3 * i + 2], 0.0f);
    else
        vel[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void k_export_aos(const float4* __restrict__ src, uint32_t n, float* __restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 v = src[i];
    dst[3 * i] = v.x;
    dst[3 * i + 1] = v.y;
    dst[3 * i + 2] = v.z;
}

inline dim3 blocks_for(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

}  // namespace

int launch_pass(int pass, const StepParams& p, const StepBuffers& b, float dt,
                uint32_t key_bits, bool write_accel, hipStream_t stream) {
    if (p.n == 0) return SPH_OK;
    switch (pass) {
        case kPassHash: {
            const uint32_t reset_n = p.n > p.num_cells ? p.n : p.num_cells;
            hipLaunchKernelGGL(k_hash_reset, blocks_for(reset_n), dim3(kBlock), 0, stream, b.pos,
                               p, b.keys, b.cell_start, b.cell_end);
            break;
        }
        case kPassSort: {
            const int rc = radix_sort_pairs(b.keys, nullptr, b.skeys, b.perm, p.n, key_bits,
                                            b.sort_temp, b.sort_temp_bytes, stream);
            if (rc != SPH_OK) return rc;
            break;
        }
        case kPassReorder:
            hipLaunchKernelGGL(k_reorder, blocks_for(p.n), dim3(kBlock), 0, stream, b.skeys,
                               b.perm, b.pos, b.vel, b.tpos, b.tvel, b.cell_start, b.cell_end,
                               p.n, p.num_cells);
            break;
        case kPassDensity:
            hipLaunchKernelGGL(k_density, blocks_for(p.n), dim3(kBlock), 0, stream, b.tpos,
                               b.cell_start, b.cell_end, p, b.rho);
            break;
        case kPassForce: {
            float4* acc = write_accel ? b.accel : nullptr;
            if (p.tens != 0.0f)
                hipLaunchKernelGGL(k_force<true>, blocks_for(p.n), dim3(kBlock), 0, stream,
                                   b.tpos, b.tvel, b.rho, b.cell_start, b.cell_end, p, dt, b.pos,
                                   b.vel, acc);
            else
                hipLaunchKernelGGL(k_force<false>, blocks_for(p.n), dim3(kBlock), 0, stream,
                                   b.tpos, b.tvel, b.rho, b.cell_start, b.cell_end, p, dt, b.pos,
                                   b.vel, acc);
            break;
        }
        default:
            return set_error(SPH_ERR_INVALID_ARGUMENT, "unknown pass");
    }
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

int launch_step(const StepParams& p, const StepBuffers& b, float dt, uint32_t key_bits,
                bool write_accel, hipStream_t stream) {
    for (int pass = 0; pass < kNumPasses; ++pass) {
        const int rc = launch_pass(pass, p, b, dt, key_bits, write_accel, stream);
        if (rc != SPH_OK) return rc;
    }
    return SPH_OK;
}

int launch_pack_state(const float* pxyz, const float* vxyz, uint32_t n, float4* pos,
                      float4* vel, hipStream_t stream) {
    if (n == 0) return SPH_OK;
    hipLaunchKernelGGL(k_pack_state, blocks_for(n), dim3(kBlock), 0, stream, pxyz, vxyz, n, pos,
                       vel);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

int launch_export_aos(const float4* src, uint32_t n, float* dst, hipStream_t stream) {
    if (n == 0) return SPH_OK;
    hipLaunchKernelGGL(k_export_aos, blocks_for(n), dim3(kBlock), 0, stream, src, n, dst);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

}  // namespace sph
