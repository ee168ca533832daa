// step.hip -- the single-scale SPH step on gfx950.
//
// Pass order of Solver::Advance (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1082-1198):
//   hash+hist      cell hash per particle fused with the radix-sort digit
//                  histograms, plus the empty-cell reset of the ranges
//                  (computeHashs :401-424 and the two cudaMemsets :1118-1127)
//   radix passes   stable (hash, slot) sort          (thrust, :1110-1115)
//   k_reorder      cell start/end, and the gather of x, v into sorted slots
//                  when a pass needs the sorted copies
//                  (reorderAndComputeCellStartEndD :478-534)
//   k_xorder       internal x-order of each cell (reads the state through
//                  the sort's permutation) and its x-bin table
//   k_density      Poly6 density                    (:536-599, :198-229)
//   k_force        spiky pressure + viscosity Laplacian (+ tension) fused with
//                  symplectic Euler and the wall clamp (:601-711, :713-755)
//
// State layout: float4 pos = (x, y, z, persistent id bits), float4 vel =
// (vx, vy, vz, 0), one 16-byte load per particle.  Within one row (cy, cz)
// of the grid the cells cx are consecutive hashes, so a stencil row of up to
// three cells is one contiguous run of sorted particles.
#include "neighbors.hpp"
#include "diagnostics.hpp"
#include "pairwalk.hpp"
#include "walls.hpp"

namespace sph {
namespace {

constexpr int kBlock = 256;

// Boundary detection compares each sorted key with its predecessor read
// straight from global memory (no block barrier, so no divergent
// __syncthreads as in the reference at Solver.cu:494-518).  With kCopy the
// state is also gathered into the sorted copies tpos / tvel; without it only
// the cell ranges are written (the internal x-order then reads the state
// through perm itself).
template <bool kCopy>
__global__ __launch_bounds__(kBlock) void k_reorder(const uint32_t* __restrict__ skeys,
                                                    const uint32_t* __restrict__ perm,
                                                    const float4* __restrict__ pos,
                                                    const float4* __restrict__ vel,
                                                    float4* __restrict__ tpos,
                                                    float4* __restrict__ tvel,
                                                    uint32_t* __restrict__ cell_start,
                                                    uint32_t* __restrict__ cell_end,
                                                    const uint32_t* __restrict__ n_dev,
                                                    uint32_t num_cells) {
    const uint32_t n = *n_dev;
    const uint32_t s = xcd_local_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
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
    if (kCopy) {
        const uint32_t j = min(perm[s], n - 1);
        tpos[s] = pos[j];
        tvel[s] = vel[j];
    }
}

// Internal x-order: slot s of the sorted order (cell c, run [b, e)) moves to
// b + rank, its rank by (x, sorted slot) among the cell's particles, so each
// cell's particles are sorted by x (deterministically: ties keep the sorted
// order).  Cells are short (tens to a few hundred particles), so the rank is
// a direct count over the cell.  The block stages
// the x of every particle in its cells' runs in LDS (lanes of one cell then
// read the same LDS word each step); runs beyond the staged range are read
// from global memory.  The state of sorted slot j is tpos[j] / tvel[j], or,
// with kPerm, pos[perm[j]] / vel[perm[j]] of the previous order (tpos / tvel
// then name pos / vel).
constexpr int kXorderTile = 2048;

template <bool kBatched, bool kPerm = false>
__global__ __launch_bounds__(kBlock) void k_xorder(const uint32_t* __restrict__ skeys,
                                                   const float4* __restrict__ tpos,
                                                   const float4* __restrict__ tvel,
                                                   const uint32_t* __restrict__ cell_start,
                                                   const uint32_t* __restrict__ cell_end,
                                                   const uint32_t* __restrict__ n_dev,
                                                   StepParams p,
                                                   const SceneParams* __restrict__ scenes,
                                                   float4* __restrict__ ipos,
                                                   float4* __restrict__ ivel,
                                                   uint32_t* __restrict__ i2r,
                                                   uint4* __restrict__ bins,
                                                   const uint32_t* __restrict__ perm = nullptr) {
    __shared__ float xs_tile[kXorderTile];
    __shared__ uint32_t red[2 * (kBlock / 64)];
    const uint32_t n = *n_dev;
    auto slot = [&](uint32_t j) { return kPerm ? min(perm[j], n - 1) : j; };
    const uint32_t block0 = blockIdx.x * kBlock;
    if (block0 >= n) return;  // block-uniform
    const uint32_t s = block0 + threadIdx.x;
    const bool active = s < n;
    uint32_t key = 0, b = kEmptyCell, e = 0u;
    if (active) {
        key = min(skeys[s], p.num_cells - 1);
        // s lies in its cell's run, so the run is non-empty and b <= s < e <= n.
        b = min(cell_start[key], s);
        e = max(min(cell_end[key], n), s + 1);
    }
    uint32_t ub, ue;
    block_run_union<kBlock>(b, e, red, ub, ue);
    const uint32_t seg_n = (ue > ub) ? min(ue - ub, static_cast<uint32_t>(kXorderTile)) : 0u;
    for (uint32_t k = threadIdx.x; k < seg_n; k += kBlock) xs_tile[k] = tpos[slot(ub + k)].x;
    __syncthreads();
    if (!active) return;
    const uint32_t src = slot(s);
    const float4 xs = tpos[src];
    const float4 vs = tvel[src];
    const bool staged = e <= ub + seg_n;
    // rank: particles before s in (x, sorted slot) order; xg: the smallest x
    // after it, which is the x of the next particle in the cell's x-order.
    uint32_t rank = 0;
    float xg = __builtin_inff();
    auto count = [&](uint32_t j, float xj) {
        const bool before = (xj < xs.x) | ((xj == xs.x) & (j < s));
        rank += before ? 1u : 0u;
        xg = (!before & (j != s)) ? fminf(xg, xj) : xg;
    };
    if (staged) {
        for (uint32_t j = b; j < e; ++j) count(j, xs_tile[j - ub]);
    } else {
        for (uint32_t j = b; j < e; ++j) count(j, tpos[slot(j)].x);
    }
    const uint32_t q = b + rank;
    ipos[q] = xs;
    ivel[q] = vs;
    i2r[q] = s;

    // The cell's x-bin table: cumulative counts of its kXBins x-bins (one
    // byte each).  Bins are non-decreasing along the x-order, so the particle
    // of rank r holds the count r + 1 for every bin from its own up to (not
    // including) the next particle's bin, or up to the last bin if it is the
    // cell's last particle; the first particle also writes 0 for the bins
    // below its own.  Every entry is written by exactly one particle.  Cells
    // of more than kXBinMax particles get the all-ones marker (no trimming).
    uint8_t* t = reinterpret_cast<uint8_t*>(bins + static_cast<size_t>(key) * kXBinQuads);
    const uint32_t cnt = e - b;
    if (cnt > kXBinMax) {
        if (rank == 0)
            for (int k = 0; k < kXBins; ++k) t[k] = static_cast<uint8_t>(kXBinUnbinned);
        return;
    }
    float ox = p.ox, inv = 1.0f / p.spacing;
    uint32_t local = key, dx = static_cast<uint32_t>(p.dx);
    if (kBatched) {
        const GridView& g = scenes[__float_as_uint(vs.w)].g;
        ox = g.ox;
        inv = g.inv_spacing;
        dx = static_cast<uint32_t>(g.dx);
        local = key - g.cell_offset;
    }
    const float cx = static_cast<float>(local % dx);
    const int kq = xbin_of((xs.x - ox) * inv - cx);
    const int kn = (rank + 1 < cnt) ? max(kq, xbin_of((xg - ox) * inv - cx)) : kXBins;
    if (rank == 0)
        for (int k = 0; k < kq; ++k) t[k] = 0u;
    const uint8_t c = static_cast<uint8_t>(rank + 1);
    for (int k = kq; k < kn; ++k) t[k] = c;
}


#ifndef SPH_DENSITY_TILE
#define SPH_DENSITY_TILE 1280
#endif
#ifndef SPH_FORCE_TILE
#define SPH_FORCE_TILE 640
#endif
#ifndef SPH_LIST_CAP
#define SPH_LIST_CAP 16
#endif
constexpr int kDensityTile = SPH_DENSITY_TILE;
constexpr int kForceTile = SPH_FORCE_TILE;
constexpr int kListCap = SPH_LIST_CAP;
constexpr int kWavesPerBlock = kBlock / 64;

// Launch variants of the neighbour passes, selected at run time by the tune
// word (StepParams::tune; 0 = the defaults above).  Every variant computes
// bit-identical results: the tile size only decides which runs are read from
// LDS instead of global memory, the list capacity only when queued pairs are
// evaluated, and the grid shape only which block processes which particles.
//   bits 0-3   density LDS tile: 1280 (default), 512, 1536, 2048, 768, 1024
//   bits 4-7   force tile / list capacity / list entry:  640/16/u16 (default),
//              768/16/u16, 768/24/u16, 768/8/u32, 512/16/u16, 1024/16/u16,
//              512/8/u32, 640/24/u16, 1024/8/u32, 512/24/u16
//   bits 8-11  density grid:       one block per block-size particles, or a
//   bits 12-15 force grid:         persistent grid of 1, 2, 4, 8 blocks per CU
//                                  looping over those blocks
//   bits 16-17 density block size, bits 18-19 force block size: 256 (0), 128
//              (1), 512 (2) threads; a non-default block size runs the pass's
//              default tile and queue
//   bits 20-21 density, bits 22-23 force: particles per lane, 1 (0) or 2 (1),
//              with the default tile, queue and block size (single scenes)
// The block size times the particles per lane decides how many targets share
// one staged union.
constexpr uint32_t kPersistPerCu[] = {0, 1, 2, 4, 8};
const SceneParams* const kNoScenes = nullptr;

// Internal x-order of a neighbour pass (kX): the pass reads the internal-order
// state, trims rows with the cells' x-bins and writes results at i2r[slot].
struct XOrder {
    const uint32_t* i2r;
    const uint4* bins;
    float2* iaux;  // density: internal-order copy of aux for the force pass
    const float4* bacc;  // force: boundary acceleration per sorted slot (or NULL)
    // force: device error word; the integration ORs kDevErrNonFinite into it
    // when it writes a NaN or infinite position or velocity (or NULL).
    uint32_t* err = nullptr;
};
constexpr uint32_t kNumCus = 256;

struct LaunchShape {
    dim3 grid;
    uint32_t nvb;  // blocks of `block` particles to cover
};

inline LaunchShape launch_shape(uint32_t cap_n, uint32_t persist_idx, uint32_t block = kBlock) {
    const uint32_t nvb = cap_n ? (cap_n + block - 1) / block : 1u;
    uint32_t g = nvb;
    if (persist_idx > 0 && persist_idx < 5) {
        g = std::min(nvb, kPersistPerCu[persist_idx] * kNumCus);
        // A multiple of the XCD count keeps each virtual block on the XCD the
        // chunked mapping assigns it to.
        if (g > kNumXcds) g -= g % kNumXcds;
    }
    return {dim3(g), nvb};
}

// The neighbour passes walk the nine grid rows (cy + dy, cz + dz),
// dy, dz in {-1, 0, 1}, as block-wide phases.  In each phase every lane finds
// its candidate run in that row (neighbors.hpp), the block takes the union of
// the runs of its 256 consecutive sorted targets -- one contiguous index range
// when the targets share a grid row -- and loads it into LDS with coalesced
// reads; each lane then scans its own run out of LDS.  A run that lies outside
// the loaded range (targets spanning distant rows, or a range larger than the
// tile) is read from global memory instead, and so are the rare extra rows a
// reference stencil reaches when rounding puts a bound two cells away.

// The pair arithmetic below is written with explicit fused multiply-adds and
// contraction off, so every kernel that evaluates a pair rounds it the same
// way whatever the surrounding code: the launch variants of a pass, which
// inline it into different code, give identical sums.
__device__ __forceinline__ float dist2(float dx, float dy, float dz) {
#pragma clang fp contract(off)
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Poly6 weight in units of h^2: a = 1 - r^2 / h^2 clamped to [0, 1] (one
// fused multiply-add with the clamp modifier), so (h^2 - r^2)^3 = h^6 a^3 for
// r < h and the term is 0 otherwise; a = 1 for r = 0.
__device__ __forceinline__ float poly6_unit(const float4& xi, const float4& xj, float inv_h2) {
#pragma clang fp contract(off)
    const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    const float r2 = dist2(dx, dy, dz);
    return __builtin_amdgcn_fmed3f(__builtin_fmaf(-r2, inv_h2, 1.0f), 0.0f, 1.0f);
}

// Sum of a^3 over a run of m candidates (poly6_unit), added one at a time in
// index order (a^2 * a + sum as one fused operation), whether the run is read
// from global memory or staged in LDS (kQuad: LDS, read in 16-byte elements).
constexpr int kScanGroup = 8;  // force scan: candidates loaded together before use

template <bool kQuad, class Load>
__device__ __forceinline__ float poly6_run(const float4& xi, float inv_h2, uint32_t m,
                                           Load&& load) {
#pragma clang fp contract(off)
    float rs = 0.0f;
    auto add = [&](const float4& q) {
        if (kQuad) keep_lds_quad(q);
        const float a = poly6_unit(xi, q, inv_h2);
        rs = __builtin_fmaf(a * a, a, rs);
    };
    uint32_t k = 0;
    for (; k + 8 <= m; k += 8) {
        float4 q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) q[u] = load(k + u);
#pragma unroll
        for (int u = 0; u < 8; ++u) add(q[u]);
    }
    for (; k + 4 <= m; k += 4) {
        const float4 q0 = load(k), q1 = load(k + 1), q2 = load(k + 2), q3 = load(k + 3);
        add(q0);
        add(q1);
        add(q2);
        add(q3);
    }
    for (; k < m; ++k) add(load(k));
    return rs;
}

// computeDensitiesPressuresD + computeDensityCell (Solver.cu:198-229,536-599):
// rho_i = m c6 sum_{r_ij < h} (h^2 - r_ij^2)^3, self included, evaluated as
// (m c6 h^6) sum a^3 with a = 1 - r_ij^2 / h^2 (poly6_run).  Each run is
// summed in plain FP32 and the run sums are combined with Kahan compensation
// (with kKahan).  Also emits the per-particle factors the force pass needs for
// every neighbour j: A_j = k (rho_j - rho0) / rho_j^2 and B_j = 1 / rho_j.
// Batched scenes (kBatched): each particle's scene id is in tvel[s].w and its
// constants come from scenes[id]; runs never leave a scene because each
// scene's cells occupy their own range of the global cell array.
template <bool kBatched>
__device__ __forceinline__ void scene_constants(const SceneParams* __restrict__ scenes,
                                                const float4* __restrict__ tvel, uint32_t s,
                                                bool active, StepParams& p, GridView& g) {
    if (kBatched) {
        const uint32_t sc = active ? __float_as_uint(tvel[s].w) : 0u;
        p = scenes[sc].p;
        g = scenes[sc].g;
    }
}

// kP targets per lane (particles per wavefront 64 kP): the block covers
// the kP kB consecutive targets block0 + q kB + lane, q < kP, stages one union
// of all their runs per row and scans each lane's targets in turn against it.
// Each target's sums are taken exactly as with one target per lane.
template <bool kKahan, bool kFinalize, bool kBatched, int kTile, bool kX, int kB, int kP>
__device__ __forceinline__ void density_block(uint32_t block0, uint32_t n, float4* tile,
                                              uint32_t* red, const float4* __restrict__ tpos,
                                              const uint32_t* __restrict__ cs,
                                              const uint32_t* __restrict__ ce,
                                              const StepParams& p_in, const GridView& g_in,
                                              float* __restrict__ rho, float2* __restrict__ aux,
                                              const float4* __restrict__ tvel,
                                              const SceneParams* __restrict__ scenes,
                                              const XOrder& xo) {
    static_assert(kP == 1 || !kBatched, "batched scenes run one target per lane");
    StepParams p = p_in;
    GridView g = g_in;
    uint32_t s[kP];
    bool active[kP];
    float4 xi[kP];
    Probe pr[kP];
    float sum[kP], comp[kP];
#pragma unroll
    for (int q = 0; q < kP; ++q) {
        s[q] = block0 + q * kB + threadIdx.x;
        active[q] = s[q] < n;
        if (kBatched) scene_constants<kBatched>(scenes, tvel, s[q], active[q], p, g);
        xi[q] = active[q] ? tpos[s[q]] : make_float4(0.f, 0.f, 0.f, 0.f);
        pr[q] = make_probe(g, xi[q].x, xi[q].y, xi[q].z);
        sum[q] = 0.0f;
        comp[q] = 0.0f;
    }
    const float inv_h2 = 1.0f / p.h2;
    auto accumulate = [&](int q, float rs) {
        if (kKahan) {
            const float y = rs - comp[q];
            const float t = sum[q] + y;
            comp[q] = (t - sum[q]) - y;
            sum[q] = t;
        } else {
            sum[q] += rs;
        }
    };
    for (int dz = -1; dz <= 1; ++dz) {
        for (int dy = -1; dy <= 1; ++dy) {
            uint32_t b[kP], e[kP];
            uint32_t lb = kEmptyCell, le = 0u;
#pragma unroll
            for (int q = 0; q < kP; ++q) {
                b[q] = kEmptyCell;
                e[q] = 0u;
                if (active[q])
                    probe_row<kX>(g, pr[q], cs, ce, xo.bins, n, pr[q].cy + dy, pr[q].cz + dz,
                                  b[q], e[q]);
                lb = min(lb, b[q]);
                le = max(le, e[q]);
            }
            uint32_t ub, ue;
            block_run_union<kB>(lb, le, red, ub, ue);
            // Loaded range [ub, seg_e); empty when every lane's run is empty.
            const uint32_t seg_n = (ue > ub) ? min(ue - ub, static_cast<uint32_t>(kTile)) : 0u;
            const uint32_t seg_e = ub + seg_n;
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < seg_n; k += kB) tile[k] = tpos[ub + k];
            __syncthreads();
#pragma unroll
            for (int q = 0; q < kP; ++q) {
                const bool staged = b[q] < e[q] && b[q] >= ub && e[q] <= seg_e;
                if (staged) {
                    const float4* t = tile + (b[q] - ub);
                    accumulate(q, poly6_run<true>(xi[q], inv_h2, e[q] - b[q],
                                                  [&](uint32_t k) { return t[k]; }));
                }
                if (b[q] < e[q] && !staged) {
                    const float4* t = tpos + b[q];
                    accumulate(q, poly6_run<false>(xi[q], inv_h2, e[q] - b[q],
                                                   [&](uint32_t k) { return t[k]; }));
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kP; ++q) {
        if (!active[q]) continue;
        for (int rz = pr[q].sz; rz <= pr[q].ez; ++rz)
            for (int ry = pr[q].sy; ry <= pr[q].ey; ++ry) {
                if (abs(ry - pr[q].cy) <= 1 && abs(rz - pr[q].cz) <= 1) continue;
                uint32_t b, e;
                probe_row<kX>(g, pr[q], cs, ce, xo.bins, n, ry, rz, b, e);
                if (b < e) {
                    const float4* t = tpos + b;
                    accumulate(q, poly6_run<false>(xi[q], inv_h2, e - b,
                                                   [&](uint32_t k) { return t[k]; }));
                }
            }
        const float r = p.c6m_h6 * sum[q];
        const uint32_t o = kX ? xo.i2r[s[q]] : s[q];
        rho[o] = r;
        if (kFinalize) {
            const float inv = __builtin_amdgcn_rcpf(r);
            const float2 ab = make_float2(p.stiffness * (r - p.rest_density) * inv * inv, inv);
            aux[o] = ab;
            if (kX) xo.iaux[s[q]] = ab;
        }
    }
}

// Occupancy targets of the neighbour passes (waves per SIMD): the density
// pass fits 6 waves in 84 VGPRs, the force pass 5 in 102 with the
// 640-candidate tile (LDS 28.8 KB per block); the hints make the compiler
// allocate within those budgets.  Batched scenes take the same targets (left
// to itself the compiler gave their force pass 116 VGPRs, 4 waves; with the
// hints the batch step went from 1.92-1.96 to 1.89 ms).  The tension and
// two-scale own-set forces keep the compiler's choice (94 VGPRs for the
// latter, 5 waves already).
constexpr int density_waves(bool /*batched*/, int ppl = 1) { return ppl == 1 ? 6 : 4; }
constexpr int force_waves(bool /*batched*/, bool tension, bool integrate, int ppl = 1) {
    return (!tension && integrate) ? (ppl == 1 ? 5 : 4) : 1;
}

template <bool kKahan, bool kFinalize, bool kBatched, int kTile, bool kX, int kB, int kP>
__global__ __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(density_waves(kBatched, kP))))
void k_density(const float4* __restrict__ tpos,
                                                    const uint32_t* __restrict__ cs,
                                                    const uint32_t* __restrict__ ce,
                                                    StepParams p_in, GridView g_in,
                                                    const uint32_t* __restrict__ n_dev,
                                                    float* __restrict__ rho,
                                                    float2* __restrict__ aux,
                                                    const float4* __restrict__ tvel,
                                                    const SceneParams* __restrict__ scenes,
                                                    XOrder xo, uint32_t nvb) {
    __shared__ float4 tile[kTile];
    __shared__ uint32_t red[2 * (kB / 64)];
    const uint32_t n = *n_dev;
    for (uint32_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
        const uint32_t block0 = xcd_local_block(vb, nvb) * (kB * kP);
        if (block0 < n)  // block-uniform
            density_block<kKahan, kFinalize, kBatched, kTile, kX, kB, kP>(
                block0, n, tile, red, tpos, cs, ce, p_in, g_in, rho, aux, tvel, scenes, xo);
    }
}

// Pair sums of the force pass for one target i (position xi, velocity vi,
// A_i): pressure P, viscosity V and tension T.
template <bool kTension>
struct PairSums {
    float Px = 0.f, Py = 0.f, Pz = 0.f;
    float Vx = 0.f, Vy = 0.f, Vz = 0.f;
    float Tx = 0.f, Ty = 0.f, Tz = 0.f;

    // xa = (x_j, A_j), vb = (v_j, B_j); requires 0 < r_ij < h.
    __device__ __forceinline__ void add(const float4& xi, const float4& vi, float ai, float h,
                                        float h2, const float4& xa, const float4& vb) {
#pragma clang fp contract(off)
        const float dx = xi.x - xa.x, dy = xi.y - xa.y, dz = xi.z - xa.z;
        const float r2 = dist2(dx, dy, dz);
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float hr = __builtin_fmaf(-r2, inv_r, h);
        const float tpr = (ai + xa.w) * hr * hr * inv_r;
        Px = __builtin_fmaf(tpr, dx, Px);
        Py = __builtin_fmaf(tpr, dy, Py);
        Pz = __builtin_fmaf(tpr, dz, Pz);
        const float tvs = hr * vb.w;
        Vx = __builtin_fmaf(tvs, vb.x - vi.x, Vx);
        Vy = __builtin_fmaf(tvs, vb.y - vi.y, Vy);
        Vz = __builtin_fmaf(tvs, vb.z - vi.z, Vz);
        if (kTension) {
            const float a = h2 - r2;
            const float w = a * a * a;
            Tx = __builtin_fmaf(w, dx, Tx);
            Ty = __builtin_fmaf(w, dy, Ty);
            Tz = __builtin_fmaf(w, dz, Tz);
        }
    }
};

// The force pass's result for target i at sorted slot o: the acceleration
// (plus the boundary acceleration bacc[o] when bacc is non-NULL), then, with
// kIntegrate, integrateD (Solver.cu:713-755): v += dt a; x += dt v; and the
// wall clamp.  Without kIntegrate only the acceleration is written.
template <bool kTension, bool kIntegrate>
__device__ __forceinline__ void force_finish(const PairSums<kTension>& S, const StepParams& p,
                                             const float4& xi, const float4& vi,
                                             const float2& ai, uint32_t o,
                                             const float4* __restrict__ bacc, float dt,
                                             float4* __restrict__ pos, float4* __restrict__ vel,
                                             float4* __restrict__ accel,
                                             uint32_t* __restrict__ err) {
#pragma clang fp contract(off)
    const float vc = p.visc * ai.y;
    float ax = __builtin_fmaf(vc, S.Vx, p.c45m * S.Px) + p.gx;
    float ay = __builtin_fmaf(vc, S.Vy, p.c45m * S.Py) + p.gy;
    float az = __builtin_fmaf(vc, S.Vz, p.c45m * S.Pz) + p.gz;
    if (kTension) {
        const float tc = p.tens * ai.y;
        ax = __builtin_fmaf(-tc, S.Tx, ax);
        ay = __builtin_fmaf(-tc, S.Ty, ay);
        az = __builtin_fmaf(-tc, S.Tz, az);
    }
    if (bacc) {  // a_i = f_i / rho_i + b_i + g (Solver.cu:708-710)
        const float4 b = bacc[o];
        ax += b.x;
        ay += b.y;
        az += b.z;
    }
    if (!kIntegrate) {
        accel[o] = make_float4(ax, ay, az, 0.0f);
        return;
    }
    float vx = __builtin_fmaf(dt, ax, vi.x), vy = __builtin_fmaf(dt, ay, vi.y),
          vz = __builtin_fmaf(dt, az, vi.z);
    float x = __builtin_fmaf(dt, vx, xi.x), y = __builtin_fmaf(dt, vy, xi.y),
          z = __builtin_fmaf(dt, vz, xi.z);
    if (p.flags & SPH_FLAG_WALL_CLAMP) {
        const float e = p.restitution;
        if (x < p.lox) { x = p.lox; vx = -e * vx; }
        if (x > p.hix) { x = p.hix; vx = -e * vx; }
        if (y < p.loy) { y = p.loy; vy = -e * vy; }
        if (y > p.hiy) { y = p.hiy; vy = -e * vy; }
        if (z < p.loz) { z = p.loz; vz = -e * vz; }
        if (z > p.hiz) { z = p.hiz; vz = -e * vz; }
    }
    pos[o] = make_float4(x, y, z, xi.w);
    vel[o] = make_float4(vx, vy, vz, vi.w);
    if (accel) accel[o] = make_float4(ax, ay, az, 0.0f);
    // A NaN or infinity in any component makes the sum non-finite.
    if (err && !__builtin_isfinite(((x + y) + (z + vx)) + (vy + vz))) atomicOr(err, kDevErrNonFinite);
}

// a_i = m c45 sum_j (A_i + A_j)(h-r)^2/r x_ij            (pressure)
//     + (mu m c45 / rho_i) sum_j B_j (h-r)(v_j - v_i)     (viscosity)
//     - (sigma m c6 / rho_i) sum_j (h^2-r^2)^3 x_ij       (tension)
//     + g,                     0 < r_ij < h
// which is f_i / rho_i of Solver.cu:231-309,601-711 with rho_i factored out,
// fused with integrateD (Solver.cu:713-755) and the wall clamp.
//
// Candidates are tested as they are scanned; in-range ones are appended to
// the lane's column of an LDS list (list[k][lane], conflict-free for any k)
// and their pair terms are evaluated when some lane's column is full and at
// the end of every phase, so the pair body runs for neighbours only.  The
// tile holds (x, y, z, A_j) and (vx, vy, vz, B_j) per candidate.
// With kIntegrate = false (two-scale own-set pass) the kernel only writes the
// acceleration, gravity included, to accel.
template <bool kTension, bool kIntegrate, bool kBatched, int kTile, int kCap, class TList,
          bool kX, int kB, int kP>
__device__ __forceinline__ void force_block(uint32_t block0, uint32_t n, float4* tp, float4* tv,
                                            TList (*list)[kB], uint32_t* red,
                                            const float4* __restrict__ tpos,
                                            const float4* __restrict__ tvel,
                                            const float2* __restrict__ aux,
                                            const uint32_t* __restrict__ cs,
                                            const uint32_t* __restrict__ ce,
                                            const StepParams& p_in, const GridView& g_in,
                                            float dt, float4* __restrict__ pos,
                                            float4* __restrict__ vel,
                                            float4* __restrict__ accel,
                                            const SceneParams* __restrict__ scenes,
                                            const XOrder& xo) {
    static_assert(kP == 1 || !kBatched, "batched scenes run one target per lane");
    StepParams p = p_in;
    GridView g = g_in;
    uint32_t s[kP];
    bool active[kP];
    float4 xi[kP], vi[kP];
    float2 ai[kP];
    Probe pr[kP];
    PairSums<kTension> S[kP];
#pragma unroll
    for (int q = 0; q < kP; ++q) {
        s[q] = block0 + q * kB + threadIdx.x;
        active[q] = s[q] < n;
        if (kBatched) scene_constants<kBatched>(scenes, tvel, s[q], active[q], p, g);
        xi[q] = active[q] ? tpos[s[q]] : make_float4(0.f, 0.f, 0.f, 0.f);
        vi[q] = active[q] ? tvel[s[q]] : make_float4(0.f, 0.f, 0.f, 0.f);
        ai[q] = active[q] ? aux[s[q]] : make_float2(0.f, 1.f);
        pr[q] = make_probe(g, xi[q].x, xi[q].y, xi[q].z);
    }
    uint32_t cnt = 0;

    auto pair = [&](int q, const float4& xa, const float4& vb) {
        S[q].add(xi[q], vi[q], ai[q].x, p.h, p.h2, xa, vb);
    };
    // 0 < r2 < h2 as one unsigned compare: r2 >= +0 is a sum of squares, and
    // for non-negative floats the bit patterns order like the values.
    const uint32_t h2_bits_m1 = __float_as_uint(p.h2) - 1u;
    auto in_range = [&](int q, const float4& xj) {
        const float dx = xi[q].x - xj.x, dy = xi[q].y - xj.y, dz = xi[q].z - xj.z;
        const float r2 = dist2(dx, dy, dz);
        return __float_as_uint(r2) - 1u < h2_bits_m1;
    };
    // Queued pairs in queue order; two list entries are fetched per step so
    // their tile reads overlap.
    auto flush = [&](int q) {
        uint32_t k = 0;
        for (; k + 2 <= cnt; k += 2) {
            const uint32_t t0 = list[k][threadIdx.x], t1 = list[k + 1][threadIdx.x];
            const float4 a0 = tp[t0], b0 = tv[t0], a1 = tp[t1], b1 = tv[t1];
            pair(q, a0, b0);
            pair(q, a1, b1);
        }
        if (k < cnt) {
            const uint32_t t0 = list[k][threadIdx.x];
            pair(q, tp[t0], tv[t0]);
        }
        cnt = 0;
    };
    auto global_run = [&](int q, uint32_t b, uint32_t e) {
        for (uint32_t j = b; j < e; ++j) {
            const float4 xj = tpos[j];
            if (in_range(q, xj)) {
                const float4 vj = tvel[j];
                const float2 aj = aux[j];
                pair(q, make_float4(xj.x, xj.y, xj.z, aj.x), make_float4(vj.x, vj.y, vj.z, aj.y));
            }
        }
    };

    for (int dz = -1; dz <= 1; ++dz) {
        for (int dy = -1; dy <= 1; ++dy) {
            uint32_t b[kP], e[kP];
            uint32_t lb = kEmptyCell, le = 0u;
#pragma unroll
            for (int q = 0; q < kP; ++q) {
                b[q] = kEmptyCell;
                e[q] = 0u;
                if (active[q])
                    probe_row<kX>(g, pr[q], cs, ce, xo.bins, n, pr[q].cy + dy, pr[q].cz + dz,
                                  b[q], e[q]);
                lb = min(lb, b[q]);
                le = max(le, e[q]);
            }
            uint32_t ub, ue;
            block_run_union<kB>(lb, le, red, ub, ue);
            const uint32_t seg_n = (ue > ub) ? min(ue - ub, static_cast<uint32_t>(kTile)) : 0u;
            const uint32_t seg_e = ub + seg_n;
            __syncthreads();
            for (uint32_t k = threadIdx.x; k < seg_n; k += kB) {
                const float4 x = tpos[ub + k];
                const float4 v = tvel[ub + k];
                const float2 a = aux[ub + k];
                tp[k] = make_float4(x.x, x.y, x.z, a.x);
                tv[k] = make_float4(v.x, v.y, v.z, a.y);
            }
            __syncthreads();
            constexpr int kG = kCap >= 16 ? kScanGroup : kScanGroup / 2;
#pragma unroll
            for (int q = 0; q < kP; ++q) {
                const bool staged = b[q] < e[q] && b[q] >= ub && e[q] <= seg_e;
                if (staged) {
                    // Whole groups of candidates are loaded before any test (the
                    // tile is padded); every candidate's index is written at the
                    // queue tail and the tail advances only for in-range ones.
                    const uint32_t t0 = b[q] - ub, t1 = e[q] - ub;
                    for (uint32_t t = t0; t < t1; t += kG) {
                        float4 c[kG];
#pragma unroll
                        for (int u = 0; u < kG; ++u) c[u] = tp[t + u];
#pragma unroll
                        for (int u = 0; u < kG; ++u) {
                            keep_lds_quad(c[u]);
                            const bool hit = (t + u < t1) & in_range(q, c[u]);  // no short circuit
                            list[cnt][threadIdx.x] = static_cast<TList>(t + u);
                            cnt += hit ? 1u : 0u;
                        }
                        if (__ballot(cnt > kCap - kG) != 0ull) flush(q);
                    }
                }
                if (b[q] < e[q] && !staged) global_run(q, b[q], e[q]);
                flush(q);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kP; ++q) {
        if (!active[q]) continue;
        for (int rz = pr[q].sz; rz <= pr[q].ez; ++rz)
            for (int ry = pr[q].sy; ry <= pr[q].ey; ++ry) {
                if (abs(ry - pr[q].cy) <= 1 && abs(rz - pr[q].cz) <= 1) continue;
                uint32_t b, e;
                probe_row<kX>(g, pr[q], cs, ce, xo.bins, n, ry, rz, b, e);
                global_run(q, b, e);
            }
        const uint32_t o = kX ? xo.i2r[s[q]] : s[q];
        force_finish<kTension, kIntegrate>(S[q], p, xi[q], vi[q], ai[q], o, xo.bacc, dt, pos,
                                           vel, accel, xo.err);
    }
}

template <bool kTension, bool kIntegrate, bool kBatched, int kTile, int kCap, class TList,
          bool kX, int kB, int kP>
__global__ __launch_bounds__(kB)
__attribute__((amdgpu_waves_per_eu(force_waves(kBatched, kTension, kIntegrate, kP))))
void k_force(const float4* __restrict__ tpos,
                                                  const float4* __restrict__ tvel,
                                                  const float2* __restrict__ aux,
                                                  const uint32_t* __restrict__ cs,
                                                  const uint32_t* __restrict__ ce,
                                                  StepParams p_in, GridView g_in,
                                                  const uint32_t* __restrict__ n_dev,
                                                  float dt, float4* __restrict__ pos,
                                                  float4* __restrict__ vel,
                                                  float4* __restrict__ accel,
                                                  const SceneParams* __restrict__ scenes,
                                                  XOrder xo, uint32_t nvb) {
    static_assert(kCap >= kScanGroup, "a scan group queues up to kCap / 2 candidates");
    static_assert(kTile + kScanGroup <= 65536 || sizeof(TList) == 4, "list entries index the tile");
    __shared__ float4 tp[kTile + kScanGroup];  // padded for whole-group loads
    __shared__ float4 tv[kTile];
    __shared__ TList list[kCap][kB];
    __shared__ uint32_t red[2 * (kB / 64)];
    const uint32_t n = *n_dev;
    for (uint32_t vb = blockIdx.x; vb < nvb; vb += gridDim.x) {
        const uint32_t block0 = xcd_local_block(vb, nvb) * (kB * kP);
        if (block0 < n)  // block-uniform
            force_block<kTension, kIntegrate, kBatched, kTile, kCap, TList, kX, kB, kP>(
                block0, n, tp, tv, list, red, tpos, tvel, aux, cs, ce, p_in, g_in, dt, pos, vel,
                accel, scenes, xo);
    }
}

// XSPH integrator (integrateXSPHD, SINGLE_SCALE/Solver.cu:346-395, 757-853,
// with each axis taking its own correction component): in sorted order, after
// the force pass has written every particle's acceleration,
//   v'_i = v_i + dt a_i,  v''_i = v'_i + eps m c6 sum_{r < h} B_j (h^2 - r^2)^3 (v'_j - v'_i),
//   x'_i = x_i + dt v''_i, then the wall clamp.
// The block-staged stencil walk holds (x_j, B_j) and v'_j per candidate.
constexpr int kXsphTile = 512;

__global__ __launch_bounds__(kBlock) void k_xsph_integrate(
    const float4* __restrict__ tpos, const float4* __restrict__ tvel,
    const float2* __restrict__ aux, const float4* __restrict__ accel,
    const uint32_t* __restrict__ cs, const uint32_t* __restrict__ ce, StepParams p, GridView g,
    const uint32_t* __restrict__ n_dev, float dt, float4* __restrict__ pos,
    float4* __restrict__ vel) {
    __shared__ float4 tx[kXsphTile];
    __shared__ float4 tv[kXsphTile];
    __shared__ uint32_t red[2 * kWavesPerBlock];
    const uint32_t n = *n_dev;
    const uint32_t s = xcd_local_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const bool active = s < n;
    const float4 xi = active ? tpos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 vi = active ? tvel[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (active) {
        const float4 a = accel[s];
        vi.x += dt * a.x;
        vi.y += dt * a.y;
        vi.z += dt * a.z;
    }
    const Probe pr = make_probe(g, xi.x, xi.y, xi.z);
    float Sx = 0.f, Sy = 0.f, Sz = 0.f;
    auto pair = [&](const float4& xb, const float4& v) {
        const float dx = xi.x - xb.x, dy = xi.y - xb.y, dz = xi.z - xb.z;
        const float r2 = dist2(dx, dy, dz);
        if (r2 < p.h2) {
            const float a = p.h2 - r2;
            const float w = a * a * a * xb.w;
            Sx += w * (v.x - vi.x);
            Sy += w * (v.y - vi.y);
            Sz += w * (v.z - vi.z);
        }
    };
    auto vprime = [&](uint32_t j) {
        const float4 v = tvel[j], a = accel[j];
        return make_float4(v.x + dt * a.x, v.y + dt * a.y, v.z + dt * a.z, 0.0f);
    };
    walk_stencil_staged<kBlock, kXsphTile>(
        g, pr, cs, ce, n, active, red,
        [&](uint32_t k, uint32_t j) {
            const float4 x = tpos[j];
            tx[k] = make_float4(x.x, x.y, x.z, aux[j].y);
            tv[k] = vprime(j);
        },
        [&](uint32_t t0, uint32_t t1, uint32_t) {
            for (uint32_t t = t0; t < t1; ++t) {
                const float4 q = tx[t], v = tv[t];
                keep_lds_quad(q);
                keep_lds_quad(v);
                pair(q, v);
            }
        },
        [&](uint32_t b, uint32_t e) {
            for (uint32_t j = b; j < e; ++j) {
                const float4 x = tpos[j];
                pair(make_float4(x.x, x.y, x.z, aux[j].y), vprime(j));
            }
        });
    if (!active) return;
    float vx = vi.x + p.xsph * Sx, vy = vi.y + p.xsph * Sy, vz = vi.z + p.xsph * Sz;
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
    vel[s] = make_float4(vx, vy, vz, vi.w);
}

// Variant dispatch.  Only the passes of the single-scale and batched steps
// (density with aux, force with integration and no tension) are tunable; the
// other instantiations use the defaults.  The single-scale and batched passes
// (density with aux, force with integration) run in the internal x-order; the
// two-scale own-set passes (neither) in the sorted order.
// Block size of a tunable pass from its tune-word bits: 0 = 256, 1 = 128,
// 2 = 512 threads; 3 is invalid, refused by tune_word_valid.

template <bool kKahan, bool kFinalize, bool kBatched, bool kX = kFinalize, class... A>
void launch_density(uint32_t tune, uint32_t cap_n, hipStream_t st, A... args) {
    // The step's density pass and the two-scale own-set pass (x-order, no
    // finalisation) take the tune word.
    constexpr bool kTunable = kFinalize || kX;
    const uint32_t v = kTunable ? (tune & 0xFu) : 0u;
    const uint32_t bv = kTunable ? ((tune >> 16) & 0x3u) : 0u;
    const uint32_t pv = (kTunable && !kBatched) ? ((tune >> 20) & 0x3u) : 0u;
    const uint32_t persist = kTunable ? ((tune >> 8) & 0xFu) : 0u;
#define SPH_DENSITY_VARIANT(TILE, B, P)                                                         \
    do {                                                                                        \
        const LaunchShape sh = launch_shape(cap_n, persist, (B) * (P));                         \
        hipLaunchKernelGGL((k_density<kKahan, kFinalize, kBatched, TILE, kX, B, P>), sh.grid,  \
                           dim3(B), 0, st, args..., sh.nvb);                                    \
    } while (0)
    if constexpr (kTunable) {
        if constexpr (!kBatched)
            if (pv == 1) { SPH_DENSITY_VARIANT(kDensityTile, kBlock, 2); return; }
        if (bv == 1) { SPH_DENSITY_VARIANT(kDensityTile, 128, 1); return; }
        if (bv == 2) { SPH_DENSITY_VARIANT(kDensityTile, 512, 1); return; }
        switch (v) {
            case 1: SPH_DENSITY_VARIANT(512, kBlock, 1); return;
            case 2: SPH_DENSITY_VARIANT(1536, kBlock, 1); return;
            case 3: SPH_DENSITY_VARIANT(2048, kBlock, 1); return;
            case 4: SPH_DENSITY_VARIANT(768, kBlock, 1); return;
            case 5: SPH_DENSITY_VARIANT(1024, kBlock, 1); return;
            default: break;
        }
    }
    SPH_DENSITY_VARIANT(kDensityTile, kBlock, 1);
#undef SPH_DENSITY_VARIANT
}

template <bool kTension, bool kIntegrate, bool kBatched, bool kX = kIntegrate, class... A>
void launch_force(uint32_t tune, uint32_t cap_n, hipStream_t st, A... args) {
    // The step's force pass and the two-scale own-set pass (x-order, no
    // integration) take the tune word; the tension forms run the default.
    constexpr bool kTunable = !kTension && (kIntegrate || kX);
    const uint32_t v = kTunable ? ((tune >> 4) & 0xFu) : 0u;
    const uint32_t bv = kTunable ? ((tune >> 18) & 0x3u) : 0u;
    const uint32_t pv = (kTunable && !kBatched) ? ((tune >> 22) & 0x3u) : 0u;
    const uint32_t persist = kTunable ? ((tune >> 12) & 0xFu) : 0u;
#define SPH_FORCE_VARIANT(TILE, CAP, TLIST, B, P)                                               \
    do {                                                                                        \
        const LaunchShape sh = launch_shape(cap_n, persist, (B) * (P));                         \
        hipLaunchKernelGGL((k_force<kTension, kIntegrate, kBatched, TILE, CAP, TLIST, kX, B, P>), \
                           sh.grid, dim3(B), 0, st, args..., sh.nvb);                           \
    } while (0)
    if constexpr (kTunable) {
        if constexpr (!kBatched)
            if (pv == 1) { SPH_FORCE_VARIANT(kForceTile, kListCap, uint16_t, kBlock, 2); return; }
        if (bv == 1) { SPH_FORCE_VARIANT(kForceTile, kListCap, uint16_t, 128, 1); return; }
        if (bv == 2) { SPH_FORCE_VARIANT(kForceTile, kListCap, uint16_t, 512, 1); return; }
        switch (v) {
            case 1: SPH_FORCE_VARIANT(768, 16, uint16_t, kBlock, 1); return;
            case 2: SPH_FORCE_VARIANT(768, 24, uint16_t, kBlock, 1); return;
            case 3: SPH_FORCE_VARIANT(768, 8, uint32_t, kBlock, 1); return;
            case 4: SPH_FORCE_VARIANT(512, 16, uint16_t, kBlock, 1); return;
            case 5: SPH_FORCE_VARIANT(1024, 16, uint16_t, kBlock, 1); return;
            case 6: SPH_FORCE_VARIANT(512, 8, uint32_t, kBlock, 1); return;
            case 7: SPH_FORCE_VARIANT(640, 24, uint16_t, kBlock, 1); return;
            case 8: SPH_FORCE_VARIANT(1024, 8, uint32_t, kBlock, 1); return;
            case 9: SPH_FORCE_VARIANT(512, 24, uint16_t, kBlock, 1); return;
            default: break;
        }
    }
    SPH_FORCE_VARIANT(kForceTile, kListCap, uint16_t, kBlock, 1);
#undef SPH_FORCE_VARIANT
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
    const GridView g = make_grid_view(p.ox, p.oy, p.oz, p.spacing, p.dx, p.dy, p.dz, p.h);
    switch (pass) {
        case kPassHash: {
            const int rc = hash_hist_cells(p, b.n_dev, b.pos, b.keys, key_bits, b.sort_temp,
                                           b.sort_temp_bytes, b.cell_start, b.cell_end, stream);
            if (rc != SPH_OK) return rc;
            break;
        }
        case kPassSort: {
            const int rc = sort_cells(p, b.keys, b.skeys, b.perm, key_bits, b.sort_temp, stream);
            if (rc != SPH_OK) return rc;
            break;
        }
        case kPassReorder: {
            // The sorted copies tpos / tvel are read by the colour, wall and
            // XSPH passes and by a step without the x-order; otherwise the
            // x-order reads the state through perm and no copy is made.
            const bool copy = !b.ipos || b.color || (b.walls && b.bacc) ||
                              (p.flags & SPH_FLAG_XSPH) || b.keep_sorted;
            if (copy)
                hipLaunchKernelGGL(k_reorder<true>, blocks_for(p.n), dim3(kBlock), 0, stream,
                                   b.skeys, b.perm, b.pos, b.vel, b.tpos, b.tvel, b.cell_start,
                                   b.cell_end, b.n_dev, p.num_cells);
            else
                hipLaunchKernelGGL(k_reorder<false>, blocks_for(p.n), dim3(kBlock), 0, stream,
                                   b.skeys, b.perm, b.pos, b.vel, b.tpos, b.tvel, b.cell_start,
                                   b.cell_end, b.n_dev, p.num_cells);
            if (b.ipos) {
                if (copy)
                    hipLaunchKernelGGL(k_xorder<false>, blocks_for(p.n), dim3(kBlock), 0, stream,
                                       b.skeys, b.tpos, b.tvel, b.cell_start, b.cell_end, b.n_dev,
                                       p, kNoScenes, b.ipos, b.ivel, b.i2r, b.xbins, nullptr);
                else
                    hipLaunchKernelGGL((k_xorder<false, true>), blocks_for(p.n), dim3(kBlock), 0,
                                       stream, b.skeys, b.pos, b.vel, b.cell_start, b.cell_end,
                                       b.n_dev, p, kNoScenes, b.ipos, b.ivel, b.i2r, b.xbins,
                                       b.perm);
            }
            break;
        }
        case kPassDensity:
            if (!b.ipos) return set_error(SPH_ERR_INVALID_ARGUMENT, "step buffers lack the x-order");
            if (p.flags & SPH_FLAG_KAHAN)
                launch_density<true, true, false>(p.tune, p.n, stream, b.ipos, b.cell_start,
                                                  b.cell_end, p, g, b.n_dev, b.rho, b.aux, b.ivel,
                                                  kNoScenes, XOrder{b.i2r, b.xbins, b.iaux});
            else
                launch_density<false, true, false>(p.tune, p.n, stream, b.ipos, b.cell_start,
                                                   b.cell_end, p, g, b.n_dev, b.rho, b.aux, b.ivel,
                                                   kNoScenes, XOrder{b.i2r, b.xbins, b.iaux});
            break;
        case kPassForce: {
            float4* acc = write_accel ? b.accel : nullptr;
            if (!b.ipos) return set_error(SPH_ERR_INVALID_ARGUMENT, "step buffers lack the x-order");
            if (b.color) {  // colour values from the force pass's inputs
                const int rc = launch_energy_color(b.tpos, b.tvel, b.cell_start, b.cell_end,
                                                   b.n_dev, p.n, g, p.mass, b.color, stream);
                if (rc != SPH_OK) return rc;
            }
            const bool walls = b.walls && b.bacc;
            if (walls) {
                const int rc = launch_wall_accel(*b.walls, b.tpos, b.n_dev, p.n, b.bacc, stream);
                if (rc != SPH_OK) return rc;
            }
            if (p.flags & SPH_FLAG_XSPH) {
                // Accelerations of every particle first (sorted order), then
                // the XSPH integration, which reads its neighbours' ones.
                const XOrder xa{nullptr, nullptr, nullptr, walls ? b.bacc : nullptr};
                float4* const no_state = nullptr;
                if (p.tens != 0.0f)
                    launch_force<true, false, false>(0u, p.n, stream, b.tpos, b.tvel, b.aux,
                                                     b.cell_start, b.cell_end, p, g, b.n_dev, 0.0f,
                                                     no_state, no_state, b.accel, kNoScenes, xa);
                else
                    launch_force<false, false, false>(0u, p.n, stream, b.tpos, b.tvel, b.aux,
                                                      b.cell_start, b.cell_end, p, g, b.n_dev,
                                                      0.0f, no_state, no_state, b.accel, kNoScenes,
                                                      xa);
                hipLaunchKernelGGL(k_xsph_integrate, blocks_for(p.n), dim3(kBlock), 0, stream,
                                   b.tpos, b.tvel, b.aux, b.accel, b.cell_start, b.cell_end, p, g,
                                   b.n_dev, dt, b.pos, b.vel);
                break;
            }
            const XOrder xo{b.i2r, b.xbins, nullptr, walls ? b.bacc : nullptr,
                            radix_sort_error_word(b.sort_temp)};
            if (p.tens != 0.0f)
                launch_force<true, true, false>(p.tune, p.n, stream, b.ipos, b.ivel, b.iaux,
                                                b.cell_start, b.cell_end, p, g, b.n_dev, dt, b.pos,
                                                b.vel, acc, kNoScenes, xo);
            else
                launch_force<false, true, false>(p.tune, p.n, stream, b.ipos, b.ivel, b.iaux,
                                                 b.cell_start, b.cell_end, p, g, b.n_dev, dt,
                                                 b.pos, b.vel, acc, kNoScenes, xo);
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

// The two-scale own-set passes in the internal x-order: density without the
// A/B finalisation (the complementary sums complete it), the internal-order
// copy of the finalised A/B, then the force without integration.  Results
// land at the sorted slots, as in the sorted-order passes above.
int launch_density_own_x(const StepParams& p, const StepBuffers& b, hipStream_t stream) {
    if (p.n == 0) return SPH_OK;
    if (!b.ipos) return set_error(SPH_ERR_INVALID_ARGUMENT, "step buffers lack the x-order");
    const GridView g = make_grid_view(p.ox, p.oy, p.oz, p.spacing, p.dx, p.dy, p.dz, p.h);
    const XOrder xo{b.i2r, b.xbins, nullptr};
    if (p.flags & SPH_FLAG_KAHAN)
        launch_density<true, false, false, true>(p.tune, p.n, stream, b.ipos, b.cell_start,
                                                 b.cell_end, p, g, b.n_dev, b.rho, b.aux, b.ivel,
                                                 kNoScenes, xo);
    else
        launch_density<false, false, false, true>(p.tune, p.n, stream, b.ipos, b.cell_start,
                                                  b.cell_end, p, g, b.n_dev, b.rho, b.aux,
                                                  b.ivel, kNoScenes, xo);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

__global__ __launch_bounds__(kBlock) void k_gather_aux(const float2* __restrict__ aux,
                                                       const uint32_t* __restrict__ i2r,
                                                       const uint32_t* __restrict__ n_dev,
                                                       float2* __restrict__ iaux) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s < *n_dev) iaux[s] = aux[i2r[s]];
}

int launch_gather_aux(const StepParams& p, const StepBuffers& b, hipStream_t stream) {
    return launch_gather_x(b.aux, b.i2r, b.n_dev, p.n, b.iaux, stream);
}

int launch_gather_x(const float2* src, const uint32_t* i2r, const uint32_t* n_dev, uint32_t cap,
                    float2* dst, hipStream_t stream) {
    if (cap == 0) return SPH_OK;
    hipLaunchKernelGGL(k_gather_aux, blocks_for(cap), dim3(kBlock), 0, stream, src, i2r, n_dev,
                       dst);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

int launch_force_own_x(const StepParams& p, const GridView& g, const StepBuffers& b,
                       float4* accel_own, hipStream_t stream) {
    if (p.n == 0) return SPH_OK;
    if (!b.ipos) return set_error(SPH_ERR_INVALID_ARGUMENT, "step buffers lack the x-order");
    float4* const no_state = nullptr;
    const XOrder xo{b.i2r, b.xbins, nullptr, nullptr};
    if (p.tens != 0.0f)
        launch_force<true, false, false, true>(p.tune, p.n, stream, b.ipos, b.ivel, b.iaux,
                                               b.cell_start, b.cell_end, p, g, b.n_dev, 0.0f,
                                               no_state, no_state, accel_own, kNoScenes, xo);
    else
        launch_force<false, false, false, true>(p.tune, p.n, stream, b.ipos, b.ivel, b.iaux,
                                                b.cell_start, b.cell_end, p, g, b.n_dev, 0.0f,
                                                no_state, no_state, accel_own, kNoScenes, xo);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

// Pair work of one neighbour pass of the most recent step, over its internal
// x-order: the candidates its runs walk (the nine rows and any rows beyond)
// and the pairs with 0 < r < h among them.  Wave sums, then one 64-bit atomic
// per wave per counter.
__global__ __launch_bounds__(kBlock) void k_count_pairs(const float4* __restrict__ ipos,
                                                        const uint32_t* __restrict__ cs,
                                                        const uint32_t* __restrict__ ce,
                                                        const uint4* __restrict__ bins,
                                                        const uint32_t* __restrict__ n_dev,
                                                        GridView g, float h2,
                                                        unsigned long long* __restrict__ out) {
    const uint32_t n = *n_dev;
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    uint32_t cand = 0, nbr = 0;
    if (s < n) {
        const float4 xi = ipos[s];
        const Probe pr = make_probe(g, xi.x, xi.y, xi.z);
        const uint32_t h2_bits_m1 = __float_as_uint(h2) - 1u;
        auto run = [&](uint32_t b, uint32_t e) {
            if (b >= e) return;
            cand += e - b;
            for (uint32_t j = b; j < e; ++j) {
                const float4 xj = ipos[j];
                const float r2 = dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z);
                nbr += (__float_as_uint(r2) - 1u < h2_bits_m1) ? 1u : 0u;
            }
        };
        for (int rz = pr.sz; rz <= pr.ez; ++rz)
            for (int ry = pr.sy; ry <= pr.ey; ++ry) {
                uint32_t b, e;
                probe_row<true>(g, pr, cs, ce, bins, n, ry, rz, b, e);
                run(b, e);
            }
    }
    cand = __ockl_wfred_add_u32(cand);
    nbr = __ockl_wfred_add_u32(nbr);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], static_cast<unsigned long long>(cand));
        atomicAdd(&out[1], static_cast<unsigned long long>(nbr));
    }
}

int launch_count_pairs(const StepParams& p, const StepBuffers& b, unsigned long long* out,
                       hipStream_t stream) {
    if (p.n == 0 || !b.ipos) return SPH_OK;
    const GridView g = make_grid_view(p.ox, p.oy, p.oz, p.spacing, p.dx, p.dy, p.dz, p.h);
    hipLaunchKernelGGL(k_count_pairs, blocks_for(p.n), dim3(kBlock), 0, stream, b.ipos,
                       b.cell_start, b.cell_end, b.xbins, b.n_dev, g, p.h2, out);
    SPH_HIP(hipGetLastError());
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

int launch_batch_pass(int pass, const StepParams& cap, const StepBuffers& b,
                      const SceneParams* scenes, bool kahan, bool tension, float dt,
                      uint32_t key_bits, bool write_accel, hipStream_t stream) {
    if (cap.n == 0) return SPH_OK;
    const GridView g0 = make_grid_view(cap.ox, cap.oy, cap.oz, cap.spacing, cap.dx, cap.dy, cap.dz,
                                       cap.h);
    switch (pass) {
        case kPassHash: {
            const int rc = hash_hist_cells_batched(cap, b.n_dev, b.pos, b.vel, scenes, b.keys,
                                                   key_bits, b.sort_temp, b.sort_temp_bytes,
                                                   b.cell_start, b.cell_end, stream);
            if (rc != SPH_OK) return rc;
            break;
        }
        case kPassSort:
            return launch_pass(pass, cap, b, dt, key_bits, write_accel, stream);
        case kPassReorder: {
            // Cell ranges only, then the x-order reads the state through perm
            // (no pass of a batched step reads the sorted copies).
            if (!b.ipos) return set_error(SPH_ERR_INVALID_ARGUMENT, "step buffers lack the x-order");
            hipLaunchKernelGGL(k_reorder<false>, blocks_for(cap.n), dim3(kBlock), 0, stream, b.skeys,
                               b.perm, b.pos, b.vel, b.tpos, b.tvel, b.cell_start, b.cell_end,
                               b.n_dev, cap.num_cells);
            hipLaunchKernelGGL((k_xorder<true, true>), blocks_for(cap.n), dim3(kBlock), 0, stream,
                               b.skeys, b.pos, b.vel, b.cell_start, b.cell_end, b.n_dev, cap,
                               scenes, b.ipos, b.ivel, b.i2r, b.xbins, b.perm);
            break;
        }
        case kPassDensity:
            if (!b.ipos) return set_error(SPH_ERR_INVALID_ARGUMENT, "step buffers lack the x-order");
            if (kahan)
                launch_density<true, true, true>(cap.tune, cap.n, stream, b.ipos, b.cell_start,
                                                 b.cell_end, cap, g0, b.n_dev, b.rho, b.aux,
                                                 b.ivel, scenes, XOrder{b.i2r, b.xbins, b.iaux});
            else
                launch_density<false, true, true>(cap.tune, cap.n, stream, b.ipos, b.cell_start,
                                                  b.cell_end, cap, g0, b.n_dev, b.rho, b.aux,
                                                  b.ivel, scenes, XOrder{b.i2r, b.xbins, b.iaux});
            break;
        case kPassForce: {
            if (!b.ipos) return set_error(SPH_ERR_INVALID_ARGUMENT, "step buffers lack the x-order");
            float4* acc = write_accel ? b.accel : nullptr;
            const XOrder xo{b.i2r, b.xbins, nullptr, nullptr, radix_sort_error_word(b.sort_temp)};
            if (tension)
                launch_force<true, true, true>(cap.tune, cap.n, stream, b.ipos, b.ivel, b.iaux,
                                               b.cell_start, b.cell_end, cap, g0, b.n_dev, dt,
                                               b.pos, b.vel, acc, scenes, xo);
            else
                launch_force<false, true, true>(cap.tune, cap.n, stream, b.ipos, b.ivel, b.iaux,
                                                b.cell_start, b.cell_end, cap, g0, b.n_dev, dt,
                                                b.pos, b.vel, acc, scenes, xo);
            break;
        }
        default:
            return set_error(SPH_ERR_INVALID_ARGUMENT, "unknown pass");
    }
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

}  // namespace sph

namespace sph {

bool tune_word_valid(uint32_t tune, bool batched) {
    const uint32_t dtile = tune & 0xFu, fvar = (tune >> 4) & 0xFu;
    const uint32_t dblock = (tune >> 16) & 0x3u, fblock = (tune >> 18) & 0x3u;
    const uint32_t dppl = (tune >> 20) & 0x3u, fppl = (tune >> 22) & 0x3u;
    return (tune >> 24) == 0u && dtile <= 5u && fvar <= 9u && ((tune >> 8) & 0xFu) <= 4u &&
           ((tune >> 12) & 0xFu) <= 4u && dblock <= 2u && fblock <= 2u && dppl <= 1u &&
           fppl <= 1u && (dblock == 0u || dtile == 0u) && (fblock == 0u || fvar == 0u) &&
           (dppl == 0u || (dtile == 0u && dblock == 0u)) &&
           (fppl == 0u || (fvar == 0u && fblock == 0u)) && (!batched || (dppl | fppl) == 0u);
}

int tune_neighbour_passes(StepParams p, const StepBuffers& b, const SceneParams* scenes,
                          bool kahan, bool tension, float dt, uint32_t key_bits, uint32_t reps,
                          hipStream_t stream, uint32_t* tune_out, float* ms_out) {
    if (p.n == 0) return set_error(SPH_ERR_INVALID_ARGUMENT, "autotune: no particles");
    reps = std::max(reps, 1u);
    auto run = [&](int pass, uint32_t tune) {
        p.tune = tune;
        return scenes ? launch_batch_pass(pass, p, b, scenes, kahan, tension, dt, key_bits, false,
                                          stream)
                      : launch_pass(pass, p, b, dt, key_bits, false, stream);
    };
    // The force pass integrates into pos/vel: keep the state to restore it.
    void* saved = nullptr;
    const size_t bytes = 16 * static_cast<size_t>(p.n);
    SPH_HIP(hipMalloc(&saved, 2 * bytes));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SPH_OK;
    auto fail = [&](int code) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipStreamSynchronize(stream);
        (void)hipFree(saved);
        return code;
    };
    hipError_t e = hipMemcpyAsync(saved, b.pos, bytes, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(static_cast<char*>(saved) + bytes, b.vel, bytes,
                           hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) return fail(hip_error(e, "autotune setup", __FILE__, __LINE__));
    for (int pass = kPassHash; pass <= kPassReorder && rc == SPH_OK; ++pass) rc = run(pass, 0u);
    if (rc != SPH_OK) return fail(rc);
    // Mean milliseconds of one launch of `pass` with tune word t (after one
    // untimed launch).
    auto time_pass = [&](int pass, uint32_t t, float& ms) -> int {
        int r = run(pass, t);
        if (r != SPH_OK) return r;
        SPH_HIP(hipEventRecord(e0, stream));
        for (uint32_t i = 0; i < reps && r == SPH_OK; ++i) r = run(pass, t);
        if (r != SPH_OK) return r;
        SPH_HIP(hipEventRecord(e1, stream));
        SPH_HIP(hipEventSynchronize(e1));
        SPH_HIP(hipEventElapsedTime(&ms, e0, e1));
        ms /= static_cast<float>(reps);
        return SPH_OK;
    };
    // Density: tile first (256-thread blocks), then the 128- and 512-thread
    // blocks and two particles per lane with the default tile, then the grid
    // shape for the best of them.
    float d_def = 0.f, d_best = 0.f, ms = 0.f;
    uint32_t d_shape = 0, d_grid = 0;  // d_shape: tile bits, or block bits << 16
    for (uint32_t v = 0; v <= 5 && rc == SPH_OK; ++v) {
        rc = time_pass(kPassDensity, v, ms);
        if (v == 0) d_def = d_best = ms;
        else if (ms < d_best) { d_best = ms; d_shape = v; }
    }
    for (uint32_t bv = 1; bv <= 2 && rc == SPH_OK; ++bv) {
        rc = time_pass(kPassDensity, bv << 16, ms);
        if (ms < d_best) { d_best = ms; d_shape = bv << 16; }
    }
    if (!scenes && rc == SPH_OK) {  // two particles per lane (single scene)
        rc = time_pass(kPassDensity, 1u << 20, ms);
        if (ms < d_best) { d_best = ms; d_shape = 1u << 20; }
    }
    for (uint32_t gsh = 1; gsh <= 4 && rc == SPH_OK; ++gsh) {
        rc = time_pass(kPassDensity, d_shape | (gsh << 8), ms);
        if (ms < d_best) { d_best = ms; d_grid = gsh; }
    }
    if (rc != SPH_OK) return fail(rc);
    const uint32_t dword = d_shape | (d_grid << 8);
    rc = run(kPassDensity, dword);  // aux for the force pass
    // Force: variant first, then the 128- and 512-thread blocks and two
    // particles per lane with the default variant, then the grid shape
    // (tension runs the default).
    float f_def = 0.f, f_best = 0.f;
    uint32_t f_shape = 0, f_grid = 0;  // variant bits << 4, or block bits << 18
    for (uint32_t v = 0; v <= (tension ? 0u : 9u) && rc == SPH_OK; ++v) {
        rc = time_pass(kPassForce, dword | (v << 4), ms);
        if (v == 0) f_def = f_best = ms;
        else if (ms < f_best) { f_best = ms; f_shape = v << 4; }
    }
    for (uint32_t bv = 1; bv <= (tension ? 0u : 2u) && rc == SPH_OK; ++bv) {
        rc = time_pass(kPassForce, dword | (bv << 18), ms);
        if (ms < f_best) { f_best = ms; f_shape = bv << 18; }
    }
    if (!scenes && !tension && rc == SPH_OK) {  // two particles per lane (single scene)
        rc = time_pass(kPassForce, dword | (1u << 22), ms);
        if (ms < f_best) { f_best = ms; f_shape = 1u << 22; }
    }
    for (uint32_t gsh = 1; gsh <= (tension ? 0u : 4u) && rc == SPH_OK; ++gsh) {
        rc = time_pass(kPassForce, dword | f_shape | (gsh << 12), ms);
        if (ms < f_best) { f_best = ms; f_grid = gsh; }
    }
    if (rc != SPH_OK) return fail(rc);
    e = hipMemcpyAsync(b.pos, saved, bytes, hipMemcpyDeviceToDevice, stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b.vel, static_cast<char*>(saved) + bytes, bytes,
                           hipMemcpyDeviceToDevice, stream);
    if (e != hipSuccess) return fail(hip_error(e, "autotune restore", __FILE__, __LINE__));
    *tune_out = dword | f_shape | (f_grid << 12);
    if (ms_out) {
        ms_out[0] = d_def;
        ms_out[1] = d_best;
        ms_out[2] = f_def;
        ms_out[3] = f_best;
    }
    return fail(SPH_OK);
}

}  // namespace sph
