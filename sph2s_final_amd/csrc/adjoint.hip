// adjoint.hip -- differentiable batched rollout: T steps of the batched
// single-scale step and the reverse-mode adjoint of a loss on the final
// positions with respect to the initial positions and velocities and every
// scene's stiffness and viscosity.  The reference has no differentiation; the
// forward step it differentiates is Solver::Advance
// (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1082-1198) with the wall clamp.
//
// Forward form per step (sorted slot order, scene constants per particle):
//   rho_i = c6m sum_j (h^2 - r^2)^3                       (r < h, self included)
//   A_i = k (rho_i - rho0) / rho_i^2,  B_i = 1 / rho_i
//   a_i = C45 sum_j (A_i + A_j) F_ij x_ij + Vc B_i sum_j (h - r) B_j (v_j - v_i) + g
//         F = (h - r)^2 / r, C45 = m 45/(pi h^6), Vc = mu C45    (0 < r < h)
//   v' = v + dt a,  x' = x + dt v',  then the wall clamp.
// Adjoint per step, topology frozen (the pair sets of the forward step):
//   k_adj_integrate  clamp and symplectic Euler: abar = dL/da, direct dL/dx, dL/dv
//   k_adj_force      dL/dA_i = C45 sum_j F x_ij.(abar_i - abar_j)
//                    dL/dB_i = Vc sum_j (h - r) B_j (v_j - v_i).(abar_i - abar_j)
//                    dL/drho_i = dL/dA_i k (2 rho0 - rho_i) B_i^3 - dL/dB_i B_i^2
//                    dL/dx_i += C45 sum_j (A_i + A_j) [F dA + F'/r (x_ij.dA) x_ij]
//                               - Vc B_i sum_j B_j (v_j - v_i).dA x_ij / r
//                    dL/dv_i += Vc B_i sum_j (h - r) B_j (abar_j - abar_i)
//                    dL/dk += dL/dA_i (rho_i - rho0) B_i^2,
//                    dL/dmu += C45 B_i sum_j (h - r) B_j (v_j - v_i).abar_i
//                    (dA = abar_i - abar_j, F'/r = -(h - r)(h + r)/r^3)
//   k_adj_density    dL/dx_i += -6 c6m sum_j (h^2 - r^2)^2 (dL/drho_i + dL/drho_j) x_ij,
//                    written at the step's input slot perm[i].
// The pair sums use that j is a neighbour of i iff i is one of j's.
//
// The adjoint passes run in the forward step's internal x-order (cells sorted
// by x inside, stencil rows trimmed with the cells' x-bins), like the forward
// neighbour passes; the adjoint of the step's outputs is gathered from the
// sorted slots and the result scattered to the step's input slots
// perm[i2r[q]].
//
// Memory: the state is checkpointed every K steps; the backward sweep
// recomputes each K-step segment from its checkpoint, keeping every step's
// x-ordered state, accelerations, densities, A/B factors, permutations, cell
// ranges and x-bin tables (68 B per particle + 24 B per cell per step), then
// runs the adjoint passes over the segment in reverse.  The forward writes the last segment's records
// itself, so that segment is never recomputed.  Parameter gradients accumulate per slot in
// FP64 (a slot always belongs to the same scene) and are reduced per scene in
// a fixed order, so the result is deterministic.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "batch.hpp"
#include "pairwalk.hpp"

using namespace sph;

namespace {

constexpr int kBlock = 256;
inline dim3 blocks_for(uint32_t n) { return dim3((n + kBlock - 1) / kBlock); }

__device__ __forceinline__ uint32_t scene_of(const uint32_t* __restrict__ off, uint32_t n_scenes,
                                             uint32_t i) {
    uint32_t lo = 0, hi = n_scenes;  // off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// State from AoS device arrays in global id order: pos.w = id, vel.w = scene.
__global__ void k_pack_batch(const float* __restrict__ p3, const float* __restrict__ v3,
                             const uint32_t* __restrict__ off, uint32_t n_scenes, uint32_t n,
                             float4* __restrict__ pos, float4* __restrict__ vel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t sc = scene_of(off, n_scenes, i);
    pos[i] = make_float4(p3[3 * i], p3[3 * i + 1], p3[3 * i + 2], __uint_as_float(i));
    vel[i] = v3 ? make_float4(v3[3 * i], v3[3 * i + 1], v3[3 * i + 2], __uint_as_float(sc))
                : make_float4(0.f, 0.f, 0.f, __uint_as_float(sc));
}

// out[id] = pos[s].xyz (ids are a permutation of [0, n)).
__global__ void k_export_by_id(const float4* __restrict__ pos, uint32_t n, float* __restrict__ out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const float4 x = pos[s];
    const uint32_t id = min(__float_as_uint(x.w), n - 1);
    out[3 * id] = x.x;
    out[3 * id + 1] = x.y;
    out[3 * id + 2] = x.z;
}

// Adjoint seed in the final state's slot order: gx[s] = g[id(s)], gv = 0.
__global__ void k_seed(const float4* __restrict__ pos, const float* __restrict__ g, uint32_t n,
                       float4* __restrict__ gx, float4* __restrict__ gv) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t id = min(__float_as_uint(pos[s].w), n - 1);
    gx[s] = make_float4(g[3 * id], g[3 * id + 1], g[3 * id + 2], 0.f);
    gv[s] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void k_unpack3(const float4* __restrict__ src, uint32_t n, float* __restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 v = src[i];
    dst[3 * i] = v.x;
    dst[3 * i + 1] = v.y;
    dst[3 * i + 2] = v.z;
}

__device__ __forceinline__ const SceneParams& scene_for(const SceneParams* __restrict__ scenes,
                                                        const float4& v) {
    return scenes[__float_as_uint(v.w)];
}

// Clamp and symplectic Euler.  The pre-clamp position is recomputed with the
// forward kernel's expressions from the stored sorted state and acceleration.
__global__ __launch_bounds__(kBlock) void k_adj_integrate(
    const float4* __restrict__ ipos, const float4* __restrict__ ivel,
    const float4* __restrict__ accel, const uint32_t* __restrict__ i2r,
    const float4* __restrict__ gx, const float4* __restrict__ gv,
    const SceneParams* __restrict__ scenes, const uint32_t* __restrict__ n_dev, float dt,
    float4* __restrict__ abar, float4* __restrict__ gxd, float4* __restrict__ gvd) {
    const uint32_t n = *n_dev;
    const uint32_t s = xcd_local_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (s >= n) return;
    // s: internal slot; the step's outputs (and their adjoints) sit at its
    // sorted slot i2r[s].
    const uint32_t o = min(i2r[s], n - 1);
    const float4 xi = ipos[s], vi = ivel[s], ai = accel[o], gxi = gx[o], gvi = gv[o];
    const StepParams& p = scene_for(scenes, vi).p;
    const bool clamp = (p.flags & SPH_FLAG_WALL_CLAMP) != 0u;
    const float e = p.restitution;
    auto axis = [&](float x, float v, float a, float gxo, float gvo, float lo, float hi, float& ga,
                    float& gxr, float& gvr) {
        const float vp = v + dt * a;
        const float xp = x + dt * vp;
        float gxp = gxo, gvp = gvo;
        if (clamp && (xp < lo || xp > hi)) {
            gxp = 0.0f;
            gvp = -e * gvo;
        }
        gvp += dt * gxp;
        gxr = gxp;
        gvr = gvp;
        ga = dt * gvp;
    };
    float4 A, X, V;
    axis(xi.x, vi.x, ai.x, gxi.x, gvi.x, p.lox, p.hix, A.x, X.x, V.x);
    axis(xi.y, vi.y, ai.y, gxi.y, gvi.y, p.loy, p.hiy, A.y, X.y, V.y);
    axis(xi.z, vi.z, ai.z, gxi.z, gvi.z, p.loz, p.hiz, A.z, X.z, V.z);
    A.w = X.w = V.w = 0.0f;
    abar[s] = A;
    gxd[s] = X;
    gvd[s] = V;
}

#ifndef SPH_ADJ_FORCE_TILE
#define SPH_ADJ_FORCE_TILE 512
#endif
#ifndef SPH_ADJ_DENSITY_TILE
#define SPH_ADJ_DENSITY_TILE 1024
#endif
constexpr int kAdjForceTile = SPH_ADJ_FORCE_TILE;
constexpr int kAdjDensityTile = SPH_ADJ_DENSITY_TILE;
constexpr int kWaves = kBlock / 64;
constexpr int kAdjGroup = 8;   // candidates loaded together before any test
constexpr int kAdjQueue = 16;  // per-lane queue of in-range candidates (force adjoint)

// 0 < r2 < h2 as one unsigned compare (r2 >= +0; non-negative floats order
// like their bit patterns).
__device__ __forceinline__ bool pair_in_range(float r2, uint32_t h2_bits_m1) {
    return __float_as_uint(r2) - 1u < h2_bits_m1;
}

// Force-pair adjoint; see the file comment for the terms.  Candidates are
// staged in LDS as (x, A), (v, B) and abar.  A lane first queues the tile
// indices of its in-range candidates (whole groups of candidates are loaded
// before any test) and evaluates the queued pair terms, in queue order, when
// some lane's queue is nearly full and at the end of each staged run.
__global__ __launch_bounds__(kBlock) void k_adj_force(
    const float4* __restrict__ tpos, const float4* __restrict__ tvel,
    const float2* __restrict__ aux, const float* __restrict__ rho,
    const uint32_t* __restrict__ i2r, const float4* __restrict__ abar,
    const uint32_t* __restrict__ cs, const uint32_t* __restrict__ ce,
    const uint4* __restrict__ bins, const SceneParams* __restrict__ scenes,
    const uint32_t* __restrict__ n_dev, float* __restrict__ grho, float4* __restrict__ gxd,
    float4* __restrict__ gvd, double* __restrict__ kacc, double* __restrict__ macc) {
    __shared__ float4 tx[kAdjForceTile + kAdjGroup];  // padded for whole-group loads
    __shared__ float4 tv[kAdjForceTile];
    __shared__ float4 ta[kAdjForceTile];
    __shared__ uint16_t queue[kAdjQueue][kBlock];
    __shared__ uint32_t red[2 * kWaves];
    const uint32_t n = *n_dev;
    const uint32_t block0 = xcd_local_block(blockIdx.x, gridDim.x) * kBlock;
    if (block0 >= n) return;  // block-uniform: the grid covers the capacity
    const uint32_t s = block0 + threadIdx.x;
    const bool active = s < n;
    const float4 vi = active ? tvel[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const SceneParams& sp = scenes[active ? __float_as_uint(vi.w) : 0u];
    const GridView g = sp.g;
    const float h = sp.p.h, h2 = sp.p.h2;
    const float4 xi = active ? tpos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 bi = active ? abar[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float2 ABi = active ? aux[s] : make_float2(0.f, 1.f);
    const Probe pr = make_probe(g, xi.x, xi.y, xi.z);
    const uint32_t h2_bits_m1 = __float_as_uint(h2) - 1u;
    float sumA = 0.f, S1 = 0.f, S2 = 0.f;
    float gPx = 0.f, gPy = 0.f, gPz = 0.f;   // sum (A_i + A_j)[F dA + F'/r (x.dA) x]
    float gVx = 0.f, gVy = 0.f, gVz = 0.f;   // sum B_j (dv.dA) x / r
    float gWx = 0.f, gWy = 0.f, gWz = 0.f;   // sum (h - r) B_j dA

    // xa = (x_j, A_j), vb = (v_j, B_j), bj = abar_j; requires 0 < r < h.
    auto pair = [&](float dx, float dy, float dz, float r2, const float4& xa, const float4& vb,
                    const float4& bj) {
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float r = r2 * inv_r;
        const float hr = h - r;
        const float F = hr * hr * inv_r;
        const float Fr = -hr * (h + r) * inv_r * inv_r * inv_r;
        const float ax = bi.x - bj.x, ay = bi.y - bj.y, az = bi.z - bj.z;
        const float ux = vb.x - vi.x, uy = vb.y - vi.y, uz = vb.z - vi.z;
        const float xdA = dx * ax + dy * ay + dz * az;
        const float udA = ux * ax + uy * ay + uz * az;
        sumA += F * xdA;
        const float wB = hr * vb.w;
        S1 += wB * (ux * bi.x + uy * bi.y + uz * bi.z);
        S2 += wB * (ux * bj.x + uy * bj.y + uz * bj.z);
        const float pa = ABi.x + xa.w;
        const float cP = pa * Fr * xdA, pF = pa * F;
        gPx += pF * ax + cP * dx;
        gPy += pF * ay + cP * dy;
        gPz += pF * az + cP * dz;
        const float cV = vb.w * udA * inv_r;
        gVx += cV * dx;
        gVy += cV * dy;
        gVz += cV * dz;
        gWx += wB * ax;
        gWy += wB * ay;
        gWz += wB * az;
    };
    // The scan loads whole groups of kAdjGroup slots: tx must hold them past
    // a full tile (the loads past a run are masked out by the range test).
    static_assert(sizeof(tx) / sizeof(tx[0]) >= kAdjForceTile + kAdjGroup,
                  "the force-adjoint tile must be padded by one group");
    walk_stencil_staged<kBlock, kAdjForceTile, true>(
        g, pr, cs, ce, n, active, red,
        [&](uint32_t k, uint32_t j) {
            const float4 x = tpos[j], v = tvel[j];
            const float2 ab = aux[j];
            tx[k] = make_float4(x.x, x.y, x.z, ab.x);
            tv[k] = make_float4(v.x, v.y, v.z, ab.y);
            ta[k] = abar[j];
        },
        [&](uint32_t t0, uint32_t t1, uint32_t) {
            uint32_t cnt = 0;
            auto flush = [&]() {
                for (uint32_t k = 0; k < cnt; ++k) {
                    const uint32_t t = queue[k][threadIdx.x];
                    const float4 xa = tx[t], va = tv[t], ba = ta[t];
                    keep_lds_quad(va);
                    keep_lds_quad(ba);
                    const float dx = xi.x - xa.x, dy = xi.y - xa.y, dz = xi.z - xa.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    pair(dx, dy, dz, r2, xa, va, ba);
                }
                cnt = 0;
            };
            for (uint32_t t = t0; t < t1; t += kAdjGroup) {
                float4 q[kAdjGroup];
#pragma unroll
                for (int u = 0; u < kAdjGroup; ++u) q[u] = tx[t + u];
#pragma unroll
                for (int u = 0; u < kAdjGroup; ++u) {
                    keep_lds_quad(q[u]);
                    const float dx = xi.x - q[u].x, dy = xi.y - q[u].y, dz = xi.z - q[u].z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    const bool hit = (t + u < t1) & pair_in_range(r2, h2_bits_m1);
                    queue[cnt][threadIdx.x] = static_cast<uint16_t>(t + u);
                    cnt += hit ? 1u : 0u;
                }
                if (__ballot(cnt > kAdjQueue - kAdjGroup) != 0ull) flush();
            }
            flush();
        },
        [&](uint32_t b, uint32_t e) {
            for (uint32_t j = b; j < e; ++j) {
                const float4 xj = tpos[j];
                const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (!pair_in_range(r2, h2_bits_m1)) continue;
                const float4 vj = tvel[j];
                const float2 ab = aux[j];
                pair(dx, dy, dz, r2, make_float4(xj.x, xj.y, xj.z, ab.x),
                     make_float4(vj.x, vj.y, vj.z, ab.y), abar[j]);
            }
        },
        bins);
    if (!active) return;
    const StepParams& p = sp.p;
    const float Bi = ABi.y, ri = rho[min(i2r[s], n - 1)];
    const float gA = p.c45m * sumA, gB = p.visc * (S1 - S2);
    grho[s] = gA * p.stiffness * (2.0f * p.rest_density - ri) * Bi * Bi * Bi - gB * Bi * Bi;
    kacc[s] += static_cast<double>(gA * (ri - p.rest_density) * Bi * Bi);
    macc[s] += static_cast<double>(p.c45m * Bi * S1);
    const float vb = p.visc * Bi;
    float4 X = gxd[s], V = gvd[s];
    X.x += p.c45m * gPx - vb * gVx;
    X.y += p.c45m * gPy - vb * gVy;
    X.z += p.c45m * gPz - vb * gVz;
    V.x -= vb * gWx;
    V.y -= vb * gWy;
    V.z -= vb * gWz;
    gxd[s] = X;
    gvd[s] = V;
}

// Density-pair adjoint, then the scatter to the step's input order.
// Candidates are staged in LDS as (x, dL/drho).  Five waves per SIMD fit in
// 93 VGPRs (the compiler's own choice, 106, allowed four).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5))) void k_adj_density(
    const float4* __restrict__ tpos, const float4* __restrict__ tvel,
    const float* __restrict__ grho, const uint32_t* __restrict__ cs,
    const uint32_t* __restrict__ ce, const uint4* __restrict__ bins,
    const uint32_t* __restrict__ i2r, const uint32_t* __restrict__ perm,
    const SceneParams* __restrict__ scenes, const uint32_t* __restrict__ n_dev,
    const float4* __restrict__ gxd, const float4* __restrict__ gvd, float4* __restrict__ gx,
    float4* __restrict__ gv) {
    __shared__ float4 tile[kAdjDensityTile + kAdjGroup];  // padded for whole-group loads
    __shared__ uint32_t red[2 * kWaves];
    const uint32_t n = *n_dev;
    const uint32_t block0 = xcd_local_block(blockIdx.x, gridDim.x) * kBlock;
    if (block0 >= n) return;  // block-uniform: the grid covers the capacity
    const uint32_t s = block0 + threadIdx.x;
    const bool active = s < n;
    const SceneParams& sp = scenes[active ? __float_as_uint(tvel[s].w) : 0u];
    const GridView g = sp.g;
    const float h2 = sp.p.h2, inv_h2 = 1.0f / h2;
    const float4 xi = active ? tpos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float gi = active ? grho[s] : 0.0f;
    const Probe pr = make_probe(g, xi.x, xi.y, xi.z);
    // Sums of a^2 (g_i + g_j) x_ij with a = 1 - r^2 / h^2 clamped to [0, 1]
    // (one clamped fused multiply-add): the pair's (h^2 - r^2)^2 is h^4 a^2,
    // and a = 0 out of range.  A coincident pair (x_ij = 0) adds zero.
    float sx = 0.f, sy = 0.f, sz = 0.f;
    auto pair = [&](float dx, float dy, float dz, float gj, bool live) {
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float a = live ? __builtin_amdgcn_fmed3f(__builtin_fmaf(-r2, inv_h2, 1.0f), 0.0f, 1.0f)
                             : 0.0f;
        const float c = a * a * (gi + gj);
        sx += c * dx;
        sy += c * dy;
        sz += c * dz;
    };
    // Padded: the scan below loads whole groups and multiplies the loads past
    // its run by a zero weight, which needs them finite.
    static_assert(sizeof(tile) / sizeof(tile[0]) >= kAdjDensityTile + kAdjGroup,
                  "the staged walk fills kTile + kPad slots");
    walk_stencil_staged<kBlock, kAdjDensityTile, true, kAdjGroup>(
        g, pr, cs, ce, n, active, red,
        [&](uint32_t k, uint32_t j) {
            const float4 x = tpos[j];
            tile[k] = make_float4(x.x, x.y, x.z, grho[j]);
        },
        [&](uint32_t t0, uint32_t t1, uint32_t) {
            for (uint32_t t = t0; t < t1; t += kAdjGroup) {
                float4 q[kAdjGroup];
#pragma unroll
                for (int u = 0; u < kAdjGroup; ++u) q[u] = tile[t + u];
#pragma unroll
                for (int u = 0; u < kAdjGroup; ++u) {
                    keep_lds_quad(q[u]);
                    // Branch-free: a candidate past the run (a finite pad
                    // or staged copy) adds zero.
                    pair(xi.x - q[u].x, xi.y - q[u].y, xi.z - q[u].z, q[u].w, t + u < t1);
                }
            }
        },
        [&](uint32_t b, uint32_t e) {
            for (uint32_t j = b; j < e; ++j) {
                const float4 xj = tpos[j];
                pair(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z, grho[j], true);
            }
        },
        bins);
    if (!active) return;
    const float k6 = -6.0f * sp.p.c6m * (h2 * h2);
    float4 X = gxd[s];
    X.x += k6 * sx;
    X.y += k6 * sy;
    X.z += k6 * sz;
    const uint32_t q = min(perm[min(i2r[s], n - 1)], n - 1);
    gx[q] = X;
    gv[q] = gvd[s];
}

// Per-scene sums of the per-slot FP64 accumulators in a fixed order: one
// block per scene, strided partial sums, then a tree in LDS.
__global__ __launch_bounds__(kBlock) void k_scene_reduce(const double* __restrict__ kacc,
                                                         const double* __restrict__ macc,
                                                         const uint32_t* __restrict__ off,
                                                         double* __restrict__ out) {
    __shared__ double sk[kBlock], sm[kBlock];
    const uint32_t sc = blockIdx.x;
    const uint32_t b = off[sc], e = off[sc + 1];
    double a = 0.0, m = 0.0;
    for (uint32_t i = b + threadIdx.x; i < e; i += kBlock) {
        a += kacc[i];
        m += macc[i];
    }
    sk[threadIdx.x] = a;
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            sk[threadIdx.x] += sk[threadIdx.x + w];
            sm[threadIdx.x] += sm[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * sc] = sk[0];
        out[2 * sc + 1] = sm[0];
    }
}

size_t al(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

}  // namespace

// One recomputed step's record inside a segment.
struct RolloutSlot {
    float4 *ipos, *ivel, *accel;  // internal x-order state; accelerations at sorted slots
    float* rho;                   // sorted slots
    float2* iaux;                 // internal order
    uint32_t *perm, *i2r, *cell_start, *cell_end;
    uint4* xbins;
    hipGraphExec_t graph = nullptr;
};

struct sph_rollout {
    sph_batch* b = nullptr;
    float dt = 0.0f;
    uint32_t n_steps = 0, every = 1, n_ckpt = 0;
    void* arena = nullptr;
    size_t arena_bytes = 0;
    float4 *ckpt_pos = nullptr, *ckpt_vel = nullptr;
    std::vector<RolloutSlot> slot;
    float4 *gx, *gv, *abar, *gxd, *gvd;
    float* grho;
    double *kacc, *macc, *red;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // With more than one segment the backward keeps two sets of K records:
    // the recomputation of one segment (on stream rs) overlaps the adjoint of
    // the next later one (on the batch's stream).
    uint32_t n_sets = 1;
    hipStream_t rs = nullptr;
    hipEvent_t ev_rec[2] = {nullptr, nullptr}, ev_adj[2] = {nullptr, nullptr}, ev_seed = nullptr;
    bool have_forward = false;
};

namespace {

// The rollout runs on the batch's own stream, ordered after the caller's
// stream (NULL: the null stream) on entry and before it on exit.
int enter(sph_rollout* r, void* stream) {
    SPH_HIP(hipEventRecord(r->ev_in, static_cast<hipStream_t>(stream)));
    SPH_HIP(hipStreamWaitEvent(r->b->stream, r->ev_in, 0));
    return SPH_OK;
}

int leave(sph_rollout* r, void* stream) {
    SPH_HIP(hipEventRecord(r->ev_out, r->b->stream));
    SPH_HIP(hipStreamWaitEvent(static_cast<hipStream_t>(stream), r->ev_out, 0));
    return SPH_OK;
}

StepBuffers slot_buffers(const sph_rollout* r, const RolloutSlot& s) {
    StepBuffers sb = r->b->buf;
    sb.ipos = s.ipos;
    sb.ivel = s.ivel;
    sb.iaux = s.iaux;
    sb.i2r = s.i2r;
    sb.xbins = s.xbins;
    sb.accel = s.accel;
    sb.rho = s.rho;
    sb.perm = s.perm;
    sb.cell_start = s.cell_start;
    sb.cell_end = s.cell_end;
    return sb;
}

int slot_graph(sph_rollout* r, RolloutSlot& s) {
    if (s.graph) return SPH_OK;
    hipStream_t st = r->b->stream;
    hipGraph_t g = nullptr;
    SPH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = batch_step(r->b, slot_buffers(r, s), r->dt, true, st);
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != SPH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
    const hipError_t ei = hipGraphInstantiate(&s.graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        s.graph = nullptr;
        return hip_error(ei, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    return SPH_OK;
}

int adjoint_step(sph_rollout* r, const RolloutSlot& s) {
    sph_batch* b = r->b;
    const uint32_t n = b->n;
    hipStream_t st = b->stream;
    hipLaunchKernelGGL(k_adj_integrate, blocks_for(n), dim3(kBlock), 0, st, s.ipos, s.ivel,
                       s.accel, s.i2r, r->gx, r->gv, b->scenes, b->buf.n_dev, r->dt, r->abar,
                       r->gxd, r->gvd);
    hipLaunchKernelGGL(k_adj_force, blocks_for(n), dim3(kBlock), 0, st, s.ipos, s.ivel, s.iaux,
                       s.rho, s.i2r, r->abar, s.cell_start, s.cell_end, s.xbins, b->scenes,
                       b->buf.n_dev, r->grho, r->gxd, r->gvd, r->kacc, r->macc);
    hipLaunchKernelGGL(k_adj_density, blocks_for(n), dim3(kBlock), 0, st, s.ipos, s.ivel, r->grho,
                       s.cell_start, s.cell_end, s.xbins, s.i2r, s.perm, b->scenes, b->buf.n_dev,
                       r->gxd, r->gvd, r->gx, r->gv);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

}  // namespace

extern "C" {

int sph_rollout_destroy(sph_rollout* r) {
    if (!r) return SPH_OK;
    if (r->b && r->b->stream) (void)hipStreamSynchronize(r->b->stream);
    if (r->rs) (void)hipStreamSynchronize(r->rs);
    for (auto& s : r->slot)
        if (s.graph) (void)hipGraphExecDestroy(s.graph);
    if (r->arena) (void)hipFree(r->arena);
    if (r->ev_in) (void)hipEventDestroy(r->ev_in);
    if (r->ev_out) (void)hipEventDestroy(r->ev_out);
    for (int k = 0; k < 2; ++k) {
        if (r->ev_rec[k]) (void)hipEventDestroy(r->ev_rec[k]);
        if (r->ev_adj[k]) (void)hipEventDestroy(r->ev_adj[k]);
    }
    if (r->ev_seed) (void)hipEventDestroy(r->ev_seed);
    if (r->rs) (void)hipStreamDestroy(r->rs);
    if (r->b) sph_batch_destroy(r->b);
    delete r;
    return SPH_OK;
}

int sph_rollout_create(const sph_config* cfgs, uint32_t n_scenes, const uint32_t* counts,
                       float dt, uint32_t n_steps, uint32_t checkpoint_every, sph_rollout** out) {
    if (!out) return set_error(SPH_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    if (!(dt > 0.0f) || !std::isfinite(dt)) return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (checkpoint_every == 0) return set_error(SPH_ERR_INVALID_ARGUMENT, "checkpoint_every must be >= 1");
    for (uint32_t s = 0; cfgs && s < n_scenes; ++s)
        if (cfgs[s].tension != 0.0f)
            return set_error(SPH_ERR_UNSUPPORTED, "rollout: tension is not differentiated");
    auto* r = new (std::nothrow) sph_rollout;
    if (!r) return set_error(SPH_ERR_INVALID_ARGUMENT, "out of host memory");
    int rc = batch_alloc(cfgs, n_scenes, counts, &r->b);
    if (rc != SPH_OK) {
        delete r;
        return rc;
    }
    r->dt = dt;
    r->n_steps = n_steps;
    r->every = std::min(checkpoint_every, std::max(n_steps, 1u));
    r->n_ckpt = (n_steps + r->every - 1) / r->every;
    const size_t N = r->b->n, C = r->b->total_cells, S = r->b->n_scenes;
    const uint32_t K = n_steps ? r->every : 0u;
    const size_t n_stored = r->n_ckpt ? r->n_ckpt - 1 : 0;  // the last segment keeps records
    r->n_sets = (r->n_ckpt > 1) ? 2u : 1u;
    std::vector<size_t> sizes = {al(16 * N * n_stored), al(16 * N * n_stored)};
    for (uint32_t k = 0; k < K * r->n_sets; ++k)
        for (size_t v : {16 * N, 16 * N, 16 * N, 4 * N, 8 * N, 4 * N, 4 * N, 4 * C, 4 * C,
                         kXBinBytes * C})
            sizes.push_back(al(v));
    for (size_t v : {16 * N, 16 * N, 16 * N, 16 * N, 16 * N, 4 * N, 8 * N, 8 * N, 16 * S})
        sizes.push_back(al(v));
    size_t total = 0;
    for (size_t v : sizes) total += v;
    hipError_t e = hipMalloc(&r->arena, total);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_in, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_out, hipEventDisableTiming);
    if (r->n_sets > 1) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&r->rs, hipStreamNonBlocking);
        for (int k = 0; k < 2; ++k) {
            if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_rec[k], hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_adj[k], hipEventDisableTiming);
        }
        if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_seed, hipEventDisableTiming);
    }
    if (e != hipSuccess) {
        sph_rollout_destroy(r);
        return hip_error(e, "rollout allocation", __FILE__, __LINE__);
    }
    r->arena_bytes = total;
    char* a = static_cast<char*>(r->arena);
    size_t i = 0;
    auto take = [&]() {
        void* p = a;
        a += sizes[i++];
        return p;
    };
    r->ckpt_pos = static_cast<float4*>(take());
    r->ckpt_vel = static_cast<float4*>(take());
    r->slot.resize(static_cast<size_t>(K) * r->n_sets);
    for (auto& s : r->slot) {
        s.ipos = static_cast<float4*>(take());
        s.ivel = static_cast<float4*>(take());
        s.accel = static_cast<float4*>(take());
        s.rho = static_cast<float*>(take());
        s.iaux = static_cast<float2*>(take());
        s.perm = static_cast<uint32_t*>(take());
        s.i2r = static_cast<uint32_t*>(take());
        s.cell_start = static_cast<uint32_t*>(take());
        s.cell_end = static_cast<uint32_t*>(take());
        s.xbins = static_cast<uint4*>(take());
    }
    r->gx = static_cast<float4*>(take());
    r->gv = static_cast<float4*>(take());
    r->abar = static_cast<float4*>(take());
    r->gxd = static_cast<float4*>(take());
    r->gvd = static_cast<float4*>(take());
    r->grho = static_cast<float*>(take());
    r->kacc = static_cast<double*>(take());
    r->macc = static_cast<double*>(take());
    r->red = static_cast<double*>(take());
    *out = r;
    return SPH_OK;
}

int sph_rollout_info(const sph_rollout* r, uint32_t* n_particles, uint32_t* n_scenes,
                     size_t* device_bytes) {
    if (!r) return set_error(SPH_ERR_INVALID_ARGUMENT, "rollout is null");
    if (n_particles) *n_particles = r->b->n;
    if (n_scenes) *n_scenes = r->b->n_scenes;
    if (device_bytes) *device_bytes = r->arena_bytes;
    return SPH_OK;
}

int sph_rollout_forward(sph_rollout* r, const float* pos0, const float* vel0,
                        const float* stiffness, const float* viscosity, float* final_pos,
                        void* stream) {
    if (!r || !pos0) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    sph_batch* b = r->b;
    const uint32_t n = b->n;
    int rc = enter(r, stream);
    if (rc != SPH_OK) return rc;
    rc = batch_set_parameters(b, stiffness, viscosity);
    if (rc != SPH_OK) return rc;
    hipLaunchKernelGGL(k_pack_batch, blocks_for(n), dim3(kBlock), 0, b->stream, pos0, vel0,
                       b->part_offset_dev, b->n_scenes, n, b->buf.pos, b->buf.vel);
    SPH_HIP(hipGetLastError());
    // The last segment runs through the segment-record graphs, so the backward
    // sweep starts from records the forward already wrote.
    const uint32_t last0 = r->n_ckpt ? (r->n_ckpt - 1) * r->every : 0u;
    for (uint32_t t = 0; t < r->n_steps; ++t) {
        if (t % r->every == 0 && t < last0) {
            const size_t o = static_cast<size_t>(t / r->every) * n;
            SPH_HIP(hipMemcpyAsync(r->ckpt_pos + o, b->buf.pos, 16 * size_t(n),
                                   hipMemcpyDeviceToDevice, b->stream));
            SPH_HIP(hipMemcpyAsync(r->ckpt_vel + o, b->buf.vel, 16 * size_t(n),
                                   hipMemcpyDeviceToDevice, b->stream));
        }
        if (t >= last0) {
            RolloutSlot& sl = r->slot[t - last0];
            rc = slot_graph(r, sl);
            if (rc != SPH_OK) return rc;
            SPH_HIP(hipGraphLaunch(sl.graph, b->stream));
        } else {
            rc = sph_batch_advance(b, r->dt, 1);
            if (rc != SPH_OK) return rc;
        }
    }
    if (final_pos) {
        hipLaunchKernelGGL(k_export_by_id, blocks_for(n), dim3(kBlock), 0, b->stream, b->buf.pos,
                           n, final_pos);
        SPH_HIP(hipGetLastError());
    }
    r->have_forward = true;
    return leave(r, stream);
}

int sph_rollout_backward(sph_rollout* r, const float* grad_final, float* grad_pos0,
                         float* grad_vel0, double* grad_stiffness, double* grad_viscosity,
                         void* stream) {
    if (!r || !grad_final) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (!r->have_forward)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "rollout: backward needs a forward first");
    sph_batch* b = r->b;
    const uint32_t n = b->n;
    int rc = enter(r, stream);
    if (rc != SPH_OK) return rc;
    hipLaunchKernelGGL(k_seed, blocks_for(n), dim3(kBlock), 0, b->stream, b->buf.pos, grad_final,
                       n, r->gx, r->gv);
    SPH_HIP(hipGetLastError());
    SPH_HIP(hipMemsetAsync(r->kacc, 0, 8 * size_t(n), b->stream));
    SPH_HIP(hipMemsetAsync(r->macc, 0, 8 * size_t(n), b->stream));
    // Segment c's records live in set (last - c) % 2 (the forward wrote the
    // last segment's into set 0).  Segment c - 1 is recomputed on rs while
    // the batch's stream runs the adjoint of segment c; a set is refilled only
    // after the adjoint that read it, and read only after its recomputation.
    const uint32_t last = r->n_ckpt ? r->n_ckpt - 1 : 0u;
    const uint32_t K = r->every;
    auto set_of = [&](uint32_t c) { return (last - c) & 1u; };
    auto slot_of = [&](uint32_t c, uint32_t t) -> RolloutSlot& {
        return r->slot[static_cast<size_t>(set_of(c)) * K + (t - c * K)];
    };
    // Every graph exists before any of them is queued (captures run on the
    // batch's stream).
    for (auto& sl : r->slot) {
        rc = slot_graph(r, sl);
        if (rc != SPH_OK) return rc;
    }
    hipStream_t rs = r->n_sets > 1 ? r->rs : b->stream;
    if (r->n_sets > 1) {  // the seed reads the final state the recomputation overwrites
        SPH_HIP(hipEventRecord(r->ev_seed, b->stream));
        SPH_HIP(hipStreamWaitEvent(rs, r->ev_seed, 0));
    }
    bool adj_done[2] = {false, false};
    for (uint32_t c = r->n_ckpt; c-- > 0;) {
        const uint32_t t0 = c * r->every, t1 = std::min(r->n_steps, t0 + r->every);
        if (c >= 1) {  // recompute segment c - 1 on rs, overlapping the adjoint of c
            const uint32_t p = c - 1, ps = set_of(p);
            if (r->n_sets > 1 && adj_done[ps]) SPH_HIP(hipStreamWaitEvent(rs, r->ev_adj[ps], 0));
            const size_t o = static_cast<size_t>(p) * n;
            SPH_HIP(hipMemcpyAsync(b->buf.pos, r->ckpt_pos + o, 16 * size_t(n),
                                   hipMemcpyDeviceToDevice, rs));
            SPH_HIP(hipMemcpyAsync(b->buf.vel, r->ckpt_vel + o, 16 * size_t(n),
                                   hipMemcpyDeviceToDevice, rs));
            const uint32_t p1 = std::min(r->n_steps, p * K + K);
            for (uint32_t t = p * K; t < p1; ++t) SPH_HIP(hipGraphLaunch(slot_of(p, t).graph, rs));
            if (r->n_sets > 1) SPH_HIP(hipEventRecord(r->ev_rec[ps], rs));
        }
        if (c < last && r->n_sets > 1) SPH_HIP(hipStreamWaitEvent(b->stream, r->ev_rec[set_of(c)], 0));
        for (uint32_t t = t1; t-- > t0;) {
            rc = adjoint_step(r, slot_of(c, t));
            if (rc != SPH_OK) return rc;
        }
        if (r->n_sets > 1) {
            SPH_HIP(hipEventRecord(r->ev_adj[set_of(c)], b->stream));
            adj_done[set_of(c)] = true;
        }
    }
    // The state now holds a recomputed segment's end (or, with one segment,
    // the forward's final state, whose records the adjoint has consumed).
    r->have_forward = false;
    if (grad_pos0)
        hipLaunchKernelGGL(k_unpack3, blocks_for(n), dim3(kBlock), 0, b->stream, r->gx, n,
                           grad_pos0);
    if (grad_vel0)
        hipLaunchKernelGGL(k_unpack3, blocks_for(n), dim3(kBlock), 0, b->stream, r->gv, n,
                           grad_vel0);
    hipLaunchKernelGGL(k_scene_reduce, dim3(b->n_scenes), dim3(kBlock), 0, b->stream, r->kacc,
                       r->macc, b->part_offset_dev, r->red);
    SPH_HIP(hipGetLastError());
    if (grad_stiffness || grad_viscosity) {
        std::vector<double> h(2 * size_t(b->n_scenes));
        SPH_HIP(hipMemcpyAsync(h.data(), r->red, 16 * size_t(b->n_scenes), hipMemcpyDeviceToHost,
                               b->stream));
        SPH_HIP(hipStreamSynchronize(b->stream));
        for (uint32_t s = 0; s < b->n_scenes; ++s) {
            if (grad_stiffness) grad_stiffness[s] = h[2 * s];
            if (grad_viscosity) grad_viscosity[s] = h[2 * s + 1];
        }
    }
    return leave(r, stream);
}

int sph_rollout_synchronize(sph_rollout* r) {
    if (!r) return set_error(SPH_ERR_INVALID_ARGUMENT, "rollout is null");
    SPH_HIP(hipStreamSynchronize(r->b->stream));
    return SPH_OK;
}

int sph_rollout_device_errors(sph_rollout* r, uint32_t* bits) {
    if (!r) return set_error(SPH_ERR_INVALID_ARGUMENT, "rollout is null");
    return sph_batch_device_errors(r->b, bits);
}

}  // extern "C"
