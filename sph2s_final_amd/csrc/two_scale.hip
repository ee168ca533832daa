// two_scale.hip -- the two-scale step: a low-resolution set (h0, m0) and a
// high-resolution set (h1 = h0/2, m1 = m0/8) advanced together, with the 1->8
// split of low-resolution particles inside an axis-aligned region and the
// 8->1 merge of sibling groups that have all left it.
//
// Reference: SPH2S_FINAL_TWO_SCALE/Solver.cu (Advance :1568-1614,
// computeNeighborhoods :1616-1680, the complementary density/force sums
// :252-290, :394-499, :732-845, :847-1048, integrateD + insertHighResParticles
// :1163-1289, :541-586) and Solver.h:16-50 (h1 = h0/2, m1 = m0/8, one grid per
// resolution).  Two couplings:
//   * fixed (the default): each set sees the other through the reference's
//     averaged kernels 1/2 (W_h0 + W_h1) (and their gradients / Laplacians)
//     with cutoff h0 and the other set's mass, with blend weight lambda = 1
//     throughout (no fade-in/fade-out timeline) and the split/merge below;
//   * blended: the reference's own coupling -- every neighbour term weighted
//     by the neighbour's lambda, integrateD's fade-in/fade-out state machine
//     and velocity cap, splits that fade the parent out and the children in,
//     and a faded merge (own-set sums in k_blend_density and
//     k_blend_force_integrate, complementary sums in the fixed coupling's
//     walks with lambda weights).
// Both couplings run the two sets' passes on two streams (graph branches),
// joined where one set's passes read the other's results of the step.
// Departures of the fixed coupling, all deliberate:
//   * split: children at x +/- h1/4 on each axis (a uniform child lattice of
//     spacing h1/2, the low-resolution lattice's h-relative spacing), mass m1,
//     velocity copied; the reference places them at r/sqrt(3) with
//     r = 1/2 (3 m0 / (4 pi rho))^(1/3) and fades them in;
//   * merge (absent in the reference): a group of 8 siblings merges when all 8
//     are outside the region expanded by h0/2, into one low-resolution particle
//     at their centroid with their mean velocity (mass and momentum conserved);
//   * bookkeeping: split/merge flags and counts come from prefix sums in
//     particle order (deterministic), counts stay on the device, and
//     emissions beyond capacity are dropped and flagged instead of written out
//     of bounds (the reference's atomicAdd has no bound, :566, and its
//     atomic compaction order is nondeterministic, :1287).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "diagnostics.hpp"
#include "neighbors.hpp"
#include "pairwalk.hpp"
#include "walls.hpp"

namespace sph {
namespace {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;

// Cross-resolution constants (the complementary sums of TWO_SCALE/Solver.cu).
struct CrossParams {
    float h0, h1, h02, h12;
    float c6_0, c6_1;      // 315 / (64 pi h^9)
    float c45_0, c45_1;    // 45 / (pi h^6)
    float m_other;         // mass of the other set's particles
    float stiffness, rest_density;
    float visc;            // mu
    float tension;         // sigma
    float gx, gy, gz, restitution;
    float lox, loy, loz, hix, hiy, hiz;
    uint32_t flags;
    float vmax;            // velocity cap c h_A / dt of the target set; 0 = off
    uint32_t* err;         // target set's device error word (non-finite state), or NULL
};

// The velocity cap of integrateD (TWO_SCALE/Solver.cu:1212-1223):
// |v| <= vmax.  The reference scales by min(|v|, vmax) / |v|, which is exactly
// 1 when |v| <= vmax (and 0/0 at |v| = 0, which is left alone here).
__device__ __forceinline__ void cap_velocity(float vmax, float& vx, float& vy, float& vz) {
    if (vmax > 0.0f) {
        const float nv = __fsqrt_rn(vx * vx + vy * vy + vz * vz);
        if (nv > vmax) {
            const float q = __fdiv_rn(vmax, nv);
            vx *= q;
            vy *= q;
            vz *= q;
        }
    }
}

// x + offset of child k: bit 0 -> x, bit 1 -> y, bit 2 -> z; set = +d.
__device__ __forceinline__ float child_offset(int k, int axis, float d) {
    return ((k >> axis) & 1) ? d : -d;
}

// ---------------------------------------------------------------------------
// The walk list is kept in 64 segments, each with its own counter 64 bytes
// from the next (appends from one block go to segment blockIdx % 64), so the
// appends do not all contend for one address.
constexpr uint32_t kWalkSegments = 64;
constexpr uint32_t kWalkCounterStride = 16;  // words

// Occupancy of the low-resolution grid by each set, so that a particle with no
// particle of the other set within reach skips the walk over that set's grid
// (most low particles lie away from the refinement region, and most high
// particles deep inside it).  Per low-grid cell: occ_h holds the step's mark
// when the cell holds a high particle, occ_l when it holds a low particle;
// near = both tests ORed over the cell's 3 x 3 x 3 block (bit 0 high, bit 1
// low).  The mark is *stamp + 1, and *stamp counts the steps (k_scan_counts
// advances it), so a new step's marks need no clearing: a cell marked in an
// earlier step holds an older value.  After 2^32 steps an old value can come
// round again; such a cell reads as occupied, which only lets a walk run that
// finds no candidates (the test only skips empty sums).
__device__ __forceinline__ uint32_t occupancy_mark(const uint32_t* __restrict__ stamp) {
    return *stamp + 1u;
}

// Marks the cells of set S's particles; block 0 also zeroes the set's
// walk-list counters for the step's k_cross_list (which runs later on the
// same stream).
__global__ void k_mark_occupancy(const float4* __restrict__ pos, const uint32_t* __restrict__ n_dev,
                                 GridView g, const uint32_t* __restrict__ stamp,
                                 uint32_t* __restrict__ occ, uint32_t* __restrict__ walk_counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (blockIdx.x == 0)
        for (uint32_t k = threadIdx.x; k < kWalkSegments; k += blockDim.x)
            walk_counts[k * kWalkCounterStride] = 0u;
    if (i >= *n_dev) return;
    const float4 x = pos[i];
    occ[cell_hash(x.x, x.y, x.z, g.ox, g.oy, g.oz, g.spacing, g.dx, g.dy, g.dz)] =
        occupancy_mark(stamp);
}

__global__ void k_dilate_occupancy(const uint32_t* __restrict__ occ_h,
                                   const uint32_t* __restrict__ occ_l, int dx, int dy, int dz,
                                   const uint32_t* __restrict__ stamp,
                                   uint32_t* __restrict__ near) {
    const uint32_t mark = occupancy_mark(stamp);
    const uint32_t cells = static_cast<uint32_t>(dx) * static_cast<uint32_t>(dy) *
                           static_cast<uint32_t>(dz);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < cells;
         c += gridDim.x * blockDim.x) {
        const int cx = static_cast<int>(c % static_cast<uint32_t>(dx));
        const int cy = static_cast<int>((c / static_cast<uint32_t>(dx)) % static_cast<uint32_t>(dy));
        const int cz = static_cast<int>(c / (static_cast<uint32_t>(dx) * static_cast<uint32_t>(dy)));
        bool h = false, l = false;
        for (int z = max(cz - 1, 0); z <= min(cz + 1, dz - 1); ++z)
            for (int y = max(cy - 1, 0); y <= min(cy + 1, dy - 1); ++y) {
                const uint32_t row = static_cast<uint32_t>(dx) * (y + dy * z);
                for (int x = max(cx - 1, 0); x <= min(cx + 1, dx - 1); ++x) {
                    h |= occ_h[row + x] == mark;
                    l |= occ_l[row + x] == mark;
                }
            }
        near[c] = (h ? 1u : 0u) | (l ? 2u : 0u);
    }
}

// The skip test of one cross walk: occ marks the low-grid cells holding a
// particle of the other set (NULL: no test, always walk) with the step's mark
// (occupancy_mark(stamp)), near the dilated marks (bit `bit`).
struct Nearby {
    const uint32_t* occ;
    const uint32_t* near;
    const uint32_t* stamp;
    uint32_t bit;
};

// True when a low-grid cell within reach of (x, y, z) -- the cells overlapping
// the box x +/- 1.001 h0 (a margin far above rounding) -- holds a particle of
// the other set.  When those cells lie within the 3 x 3 x 3 block around the
// particle's own cell (all but particles within 0.001 h0 of a cell face), the
// block's dilated mark answers with one load; otherwise the cells are read.
__device__ __forceinline__ bool other_nearby(const Nearby& nb, const GridView& ga, float x,
                                             float y, float z) {
    if (!nb.occ) return true;
    // Fast path: the particle's position in cell units (one multiply; off by
    // far less than the 0.001 margins from the exact division).  The box
    // x +/- 1.001 h0 stays within cells c - 1 .. c + 1 when the fraction f of
    // the particle's cell coordinate satisfies q - 1 < f < 2 - q with
    // q = 1.001 h0 / spacing; the margins cover the rounding of q and f.
    {
        const float q = ga.h * 1.001f * ga.inv_spacing;
        const float lo = q - 1.0f + 0.001f, hi = 2.0f - q - 0.001f;
        const float tx = (x - ga.ox) * ga.inv_spacing, ty = (y - ga.oy) * ga.inv_spacing,
                    tz = (z - ga.oz) * ga.inv_spacing;
        const float fx = floorf(tx), fy = floorf(ty), fz = floorf(tz);
        if (tx - fx > lo && tx - fx < hi && ty - fy > lo && ty - fy < hi && tz - fz > lo &&
            tz - fz < hi && fx >= 0.0f && fy >= 0.0f && fz >= 0.0f &&
            fx < static_cast<float>(ga.dx) && fy < static_cast<float>(ga.dy) &&
            fz < static_cast<float>(ga.dz)) {
            const uint32_t c = static_cast<uint32_t>(fx) +
                               static_cast<uint32_t>(ga.dx) *
                                   (static_cast<uint32_t>(fy) +
                                    static_cast<uint32_t>(ga.dy) * static_cast<uint32_t>(fz));
            return (nb.near[c] & nb.bit) != 0u;
        }
    }
    const float r = ga.h * 1.001f;
    const int sx = cell_coord_off(x, -r, ga.ox, ga.spacing, ga.dx);
    const int ex = cell_coord_off(x, r, ga.ox, ga.spacing, ga.dx);
    const int sy = cell_coord_off(y, -r, ga.oy, ga.spacing, ga.dy);
    const int ey = cell_coord_off(y, r, ga.oy, ga.spacing, ga.dy);
    const int sz = cell_coord_off(z, -r, ga.oz, ga.spacing, ga.dz);
    const int ez = cell_coord_off(z, r, ga.oz, ga.spacing, ga.dz);
    const int cx = cell_coord(x, ga.ox, ga.spacing, ga.dx);
    const int cy = cell_coord(y, ga.oy, ga.spacing, ga.dy);
    const int cz = cell_coord(z, ga.oz, ga.spacing, ga.dz);
    if (sx >= cx - 1 && ex <= cx + 1 && sy >= cy - 1 && ey <= cy + 1 && sz >= cz - 1 &&
        ez <= cz + 1)
        return (nb.near[static_cast<uint32_t>(cx) +
                        static_cast<uint32_t>(ga.dx) * (cy + ga.dy * cz)] & nb.bit) != 0u;
    const uint32_t mark = occupancy_mark(nb.stamp);
    bool any = false;
    for (int cz2 = sz; cz2 <= ez; ++cz2)
        for (int cy2 = sy; cy2 <= ey; ++cy2) {
            const uint32_t row = static_cast<uint32_t>(ga.dx) * (cy2 + ga.dy * cz2);
            for (int cx2 = sx; cx2 <= ex; ++cx2) any |= nb.occ[row + cx2] == mark;
        }
    return any;
}

// ---------------------------------------------------------------------------
// The complementary sums by lane group.  Few particles have a particle of the other
// set within h0 (those near the interface), but a low particle there sees
// hundreds of high ones (8 per low-resolution volume), and one lane walking
// them alone holds its whole wave -- and the kernel -- for the length of that
// serial walk.  So the walkers are listed first (k_cross_list, which also
// finalises the others), and each listed particle gets a group of 16, 32 or
// 64 lanes (as many as its stencil has rows, rounded up): the lanes probe its
// stencil rows in parallel, the rows' runs are laid end to end, lane l takes
// candidates l, l + kG, ... of that sequence, and the lanes' partial sums are
// reduced in a fixed order (so the result does not depend on which wave or
// group takes which particle or on the list's order).  A wave holds several
// groups: a walker's row probes, prefix sum and reductions then share their
// instructions with the other groups' (a high walker has 9 rows and a few
// hundred candidates, little for 64 lanes), and the wave has several walkers'
// chains of dependent loads in flight.
constexpr uint32_t kWalkBlocks = 4096;  // 16384 waves share a step's walk list

// Null arguments of the fixed coupling's walks (no blend weights, no separate
// complementary acceleration).
float4* const kNoXacc = nullptr;
const float2* const kNoBlend = nullptr;

inline uint32_t walk_segment_cap(uint32_t cap) {
    const uint32_t blocks = (cap + kBlock - 1) / kBlock;
    return ((blocks + kWalkSegments - 1) / kWalkSegments) * kBlock;
}

// Entry w of the segmented walk list: seg_incl holds the wave's inclusive
// prefix of the 64 segment counts (seg_incl[63] = total > w).
__device__ __forceinline__ uint32_t walk_entry(const uint32_t* __restrict__ list,
                                               uint32_t seg_cap, const uint32_t* seg_incl,
                                               uint32_t w) {
    uint32_t k = 0u;  // the number of segments whose inclusive count is <= w
    for (uint32_t step = 32u; step >= 1u; step >>= 1)
        if (seg_incl[k + step - 1u] <= w) k += step;
    const uint32_t begin = (k > 0u) ? seg_incl[k - 1u] : 0u;
    return list[k * seg_cap + (w - begin)];
}

// The wave's inclusive prefix of the segment counts, into its LDS row;
// returns the list's length.
__device__ __forceinline__ uint32_t walk_segments(const uint32_t* __restrict__ counts,
                                                  uint32_t* seg_incl) {
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t incl = counts[lane * kWalkCounterStride];
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t v = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), off));
        if (lane >= static_cast<uint32_t>(off)) incl += v;
    }
    seg_incl[lane] = incl;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return static_cast<uint32_t>(__shfl(static_cast<int>(incl), 63));
}

__device__ __forceinline__ uint64_t lanes_below(uint32_t lane) {
    return (lane == 0u) ? 0ull : (~0ull >> (64u - lane));
}

// Appends `value` of every lane with `flag` to list[*count ...] (one atomic
// per wave).  Every lane of the wave must call it.
__device__ __forceinline__ void append_flagged(bool flag, uint32_t value,
                                               uint32_t* __restrict__ count,
                                               uint32_t* __restrict__ list) {
    const uint64_t m = __ballot(flag);
    if (m == 0ull) return;  // wave-uniform
    const uint32_t lane = threadIdx.x & 63u;
    const int leader = __builtin_ctzll(m);
    uint32_t base = 0u;
    if (static_cast<int>(lane) == leader) base = atomicAdd(count, __popcll(m));
    base = static_cast<uint32_t>(__shfl(static_cast<int>(base), leader));
    if (flag) list[base + __popcll(m & lanes_below(lane))] = value;
}

// Sum over each group of kG lanes (lanes q kG .. q kG + kG - 1) in a fixed
// butterfly order (every lane of the wave active).
template <int kG>
__device__ __forceinline__ float group_sum(float v) {
    for (int off = kG / 2; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Calls body(j, pos_b[j]) for the candidates j of one particle's stencil rows
// over grid g, spread over a group of kG lanes.  The wave holds 64 / kG groups, each
// with its own particle (a group with `active` false has no rows).  Row r of
// a chunk of kG rows is probed by the group's lane r, its run
// [rb[r], rb[r] + len_r) takes positions [incl[r] - len_r, incl[r]) of the
// chunk's rows laid end to end, and the group's lane l visits positions
// l, l + kG, ...  rb_s and incl_s are the wave's own 64-entry LDS rows
// (group q uses entries q kG .. q kG + kG - 1).  A lane takes its positions
// kWalkUnroll at a time: it finds their candidates, loads their positions
// together (several loads in flight), then calls body in position order, so
// its terms are added in the same order as one at a time.  Every lane of the
// wave calls it.
constexpr int kWalkUnroll = 4;
template <int kG, class Body>
__device__ __forceinline__ void group_walk(const GridView& g, const uint32_t* __restrict__ cs,
                                           const uint32_t* __restrict__ ce,
                                           const float4* __restrict__ pos_b, uint32_t n,
                                           bool active, float x, float y, float z,
                                           uint32_t* rb_s, uint32_t* incl_s, Body&& body) {
    static_assert(kG == 16 || kG == 32 || kG == 64, "walker group width");
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gl = lane & (kG - 1u), g0 = lane - gl;
    const int sx = cell_coord_off(x, -g.h, g.ox, g.spacing, g.dx);
    const int sy = cell_coord_off(y, -g.h, g.oy, g.spacing, g.dy);
    const int sz = cell_coord_off(z, -g.h, g.oz, g.spacing, g.dz);
    const int ex = cell_coord_off(x, g.h, g.ox, g.spacing, g.dx);
    const int ey = cell_coord_off(y, g.h, g.oy, g.spacing, g.dy);
    const int ez = cell_coord_off(z, g.h, g.oz, g.spacing, g.dz);
    const float fx = grid_offset(x, g.ox, g.spacing, g.dx);
    const float fy = grid_offset(y, g.oy, g.spacing, g.dy);
    const float fz = grid_offset(z, g.oz, g.spacing, g.dz);
    const int ny = ey - sy + 1;
    const int nrows = active ? ny * (ez - sz + 1) : 0;
    // Rows in chunks of kG (the launch picks kG so that one chunk holds a
    // walker's rows); the wave runs as many chunks as its largest group needs.
    const uint32_t chunks =
        __ockl_wfred_max_u32(static_cast<uint32_t>((nrows + kG - 1) / kG));
    for (uint32_t ch = 0; ch < chunks; ++ch) {
        const int r = static_cast<int>(ch * kG + gl);
        uint32_t b = kEmptyCell, e = 0u;
        if (r < nrows) {
            const int cy = sy + r % ny, cz = sz + r / ny;
            const float gy = slab_gap(fy, cy, g.spacing);
            const float gz = slab_gap(fz, cz, g.spacing);
            const float d2 = gy * gy + gz * gz;
            if (d2 < g.prune_h2) {
                const float w = __builtin_sqrtf(g.prune_h2 - d2);
                const int lo = max(sx, static_cast<int>(floorf((fx - w) * g.inv_spacing)));
                const int hi = min(ex, static_cast<int>(floorf((fx + w) * g.inv_spacing)));
                if (lo <= hi)
                    row_run(cs, ce, g.cell_offset + static_cast<uint32_t>(g.dx) * (cy + g.dy * cz),
                            lo, hi, n, b, e);
            }
        }
        const uint32_t len = (b < e) ? e - b : 0u;
        // Inclusive prefix of the run lengths over the group's lanes.
        uint32_t incl = len;
        for (int off = 1; off < kG; off <<= 1) {
            const uint32_t v = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), off, kG));
            if (gl >= static_cast<uint32_t>(off)) incl += v;
        }
        const uint32_t total =
            static_cast<uint32_t>(__shfl(static_cast<int>(incl), static_cast<int>(g0 + kG - 1u)));
        rb_s[lane] = b;
        incl_s[lane] = incl;
        // (the wave's own LDS rows: no barrier beyond the wave is needed)
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // incl_s[g0 + kG - 1] = total > f, so the row search stays in the group.
        uint32_t row = g0;
        for (uint32_t f = gl; f < total; f += kWalkUnroll * kG) {
            uint32_t j[kWalkUnroll];
            float4 xj[kWalkUnroll];
#pragma unroll
            for (int u = 0; u < kWalkUnroll; ++u) {
                const uint32_t fu = f + static_cast<uint32_t>(u) * kG;
                j[u] = 0u;
                if (fu < total) {
                    while (incl_s[row] <= fu) ++row;
                    const uint32_t begin = (row > g0) ? incl_s[row - 1u] : 0u;
                    j[u] = rb_s[row] + (fu - begin);
                }
            }
#pragma unroll
            for (int u = 0; u < kWalkUnroll; ++u)
                if (f + static_cast<uint32_t>(u) * kG < total) xj[u] = pos_b[j[u]];
#pragma unroll
            for (int u = 0; u < kWalkUnroll; ++u)
                if (f + static_cast<uint32_t>(u) * kG < total) body(j[u], xj[u]);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Lanes per walker for walks over grid g: the smallest group of 16, 32 or 64
// lanes that probes a walker's stencil rows in one chunk.  A search radius h
// over cells of pitch `spacing` spans at most ceil(2 h / spacing) + 1 cells per
// axis (3 for the low grid at h0, 5 for the high grid at h0 = 2 h1), so 9 or 25
// rows.  The width fixes the order in which a walker's terms are added (lane
// partial sums, then the group's butterfly), so a set's walks always use the
// same one.
inline int walk_lanes(const GridView& g) {
    const int cells = static_cast<int>(std::ceil(2.0f * g.h * g.inv_spacing - 1e-3f)) + 1;
    const int rows = cells * cells;
    return rows <= 16 ? 16 : rows <= 32 ? 32 : 64;
}

// Lists set A's particles whose complementary walk is not empty and
// finalises the others: density unchanged, A and B from it, and (with xacc_a,
// the blended coupling) a complementary acceleration of -0, which the
// integration adds without changing the own-set sum (x + -0 = x for every x).
__global__ __launch_bounds__(kBlock) void k_cross_list(
    const float4* __restrict__ tpos_a, const uint32_t* __restrict__ n_a,
    const uint32_t* __restrict__ n_b, Nearby occ, GridView ga, CrossParams c,
    float* __restrict__ rho_a, float2* __restrict__ aux_a, float4* __restrict__ xacc_a,
    uint32_t* __restrict__ counts, uint32_t* __restrict__ list, uint32_t seg_cap) {
    const uint32_t na = *n_a, nb = *n_b;
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t seg = blockIdx.x % kWalkSegments;
    const bool live = s < na;
    bool walk = false;
    if (live) {
        const float4 xi = tpos_a[s];
        walk = nb > 0 && other_nearby(occ, ga, xi.x, xi.y, xi.z);
        if (!walk) {
            const float r = rho_a[s];
            const float inv = __builtin_amdgcn_rcpf(r);
            aux_a[s] = make_float2(c.stiffness * (r - c.rest_density) * inv * inv, inv);
            if (xacc_a) xacc_a[s] = make_float4(-0.0f, -0.0f, -0.0f, 0.0f);
        }
    }
    append_flagged(walk, s, counts + seg * kWalkCounterStride, list + seg * seg_cap);
}

// complementary density of the listed particles: rho_i += m_B sum_{j in B,
// r < h0} 1/2 (W_h0 + W_h1), then A_i, B_i.  One group of kG lanes per listed
// particle, 64 / kG particles per wave.
// kBlend: every term weighted by lambda_j (tbl_b[j].x; Solver.cu:252-290).
template <bool kBlend, int kG>
__global__ __launch_bounds__(kBlock) void k_cross_density_walk(
    const float4* __restrict__ tpos_a, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ list, uint32_t seg_cap, const float4* __restrict__ tpos_b,
    const float2* __restrict__ tbl_b, const uint32_t* __restrict__ cs_b,
    const uint32_t* __restrict__ ce_b, const uint32_t* __restrict__ n_b, GridView gb,
    CrossParams c, float* __restrict__ rho_a, float2* __restrict__ aux_a) {
    __shared__ uint32_t rb_s[kBlock / 64][64];
    __shared__ uint32_t incl_s[kBlock / 64][64];
    __shared__ uint32_t seg_s[kBlock / 64][64];
    constexpr uint32_t kPer = 64 / kG;  // walkers per wave
    const uint32_t nb = *n_b;
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t nw = walk_segments(counts, seg_s[wv]);
    const uint32_t stride = gridDim.x * (kBlock / 64) * kPer;
    for (uint32_t w0 = (blockIdx.x * (kBlock / 64) + wv) * kPer; w0 < nw; w0 += stride) {
        const uint32_t w = w0 + lane / kG;
        const bool act = w < nw;
        const uint32_t s = act ? walk_entry(list, seg_cap, seg_s[wv], w) : 0u;
        const float4 xi = act ? tpos_a[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float sum = 0.0f;
        group_walk<kG>(gb, cs_b, ce_b, tpos_b, nb, act, xi.x, xi.y, xi.z, rb_s[wv], incl_s[wv],
                       [&](uint32_t j, const float4& xj) {
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < c.h02) {
                const float a0 = c.h02 - r2;
                const float a1 = c.h12 - r2;
                const float w0 = c.c6_0 * a0 * a0 * a0;
                const float w1 = (r2 < c.h12) ? c.c6_1 * a1 * a1 * a1 : 0.0f;
                if (kBlend)
                    sum += tbl_b[j].x * (0.5f * (w0 + w1));
                else
                    sum += 0.5f * (w0 + w1);
            }
        });
        sum = group_sum<kG>(sum);
        if (act && (lane & (kG - 1u)) == 0u) {
            const float r = rho_a[s] + c.m_other * sum;
            rho_a[s] = r;
            const float inv = __builtin_amdgcn_rcpf(r);
            aux_a[s] = make_float2(c.stiffness * (r - c.rest_density) * inv * inv, inv);
        }
    }
}

// complementary force (TWO_SCALE/Solver.cu:394-499) of the listed particles.
// One group of kG lanes each.  Fixed coupling (lambda = 1): added to their own-set
// acceleration in acc_a.  kBlend: every term weighted by lambda_j (tbl_b[j].x)
// and the sum written to acc_a, which the blended force kernel adds.
template <bool kTension, bool kBlend, int kG>
__global__ __launch_bounds__(kBlock) void k_cross_force_walk(
    const float4* __restrict__ tpos_a, const float4* __restrict__ tvel_a,
    const float2* __restrict__ aux_a, const uint32_t* __restrict__ counts,
    const uint32_t* __restrict__ list, uint32_t seg_cap, const float4* __restrict__ tpos_b,
    const float4* __restrict__ tvel_b, const float2* __restrict__ aux_b,
    const float2* __restrict__ tbl_b, const uint32_t* __restrict__ cs_b,
    const uint32_t* __restrict__ ce_b, const uint32_t* __restrict__ n_b, GridView gb,
    CrossParams c, float4* __restrict__ acc_a) {
    __shared__ uint32_t rb_s[kBlock / 64][64];
    __shared__ uint32_t incl_s[kBlock / 64][64];
    __shared__ uint32_t seg_s[kBlock / 64][64];
    constexpr uint32_t kPer = 64 / kG;  // walkers per wave
    const uint32_t nb = *n_b;
    const uint32_t wv = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint32_t nw = walk_segments(counts, seg_s[wv]);
    const uint32_t stride = gridDim.x * (kBlock / 64) * kPer;
    for (uint32_t w0 = (blockIdx.x * (kBlock / 64) + wv) * kPer; w0 < nw; w0 += stride) {
        const uint32_t w = w0 + lane / kG;
        const bool act = w < nw;
        const uint32_t s = act ? walk_entry(list, seg_cap, seg_s[wv], w) : 0u;
        const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 xi = act ? tpos_a[s] : zero;
        const float4 vi = act ? tvel_a[s] : zero;
        const float2 ai = act ? aux_a[s] : make_float2(0.0f, 0.0f);
        float Px = 0.f, Py = 0.f, Pz = 0.f, Vx = 0.f, Vy = 0.f, Vz = 0.f, Tx = 0.f, Ty = 0.f,
              Tz = 0.f;
        group_walk<kG>(gb, cs_b, ce_b, tpos_b, nb, act, xi.x, xi.y, xi.z, rb_s[wv], incl_s[wv],
                       [&](uint32_t j, const float4& xj) {
            const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < c.h02 && r2 > 0.0f) {
                const float4 vj = tvel_b[j];
                const float2 aj = aux_b[j];
                const float inv_r = __builtin_amdgcn_rsqf(r2);
                const float r = r2 * inv_r;
                const float hr0 = c.h0 - r;
                const float hr1 = fmaxf(c.h1 - r, 0.0f);
                // 1/2 (|grad W_h0| + |grad W_h1|) / r and 1/2 (lap W_h0 + lap W_h1)
                const float gg = 0.5f * (c.c45_0 * hr0 * hr0 + c.c45_1 * hr1 * hr1) * inv_r;
                const float l = 0.5f * (c.c45_0 * hr0 + c.c45_1 * hr1);
                const float lj = kBlend ? tbl_b[j].x : 1.0f;
                const float tp = kBlend ? lj * (ai.x + aj.x) * gg : (ai.x + aj.x) * gg;
                Px += tp * dx;
                Py += tp * dy;
                Pz += tp * dz;
                const float tv = kBlend ? lj * l * aj.y : l * aj.y;
                Vx += tv * (vj.x - vi.x);
                Vy += tv * (vj.y - vi.y);
                Vz += tv * (vj.z - vi.z);
                if (kTension) {
                    const float a0 = c.h02 - r2, a1 = c.h12 - r2;
                    const float wt = (kBlend ? lj * 0.5f : 0.5f) *
                                     (c.c6_0 * a0 * a0 * a0 +
                                      ((r2 < c.h12) ? c.c6_1 * a1 * a1 * a1 : 0.0f));
                    Tx += wt * dx;
                    Ty += wt * dy;
                    Tz += wt * dz;
                }
            }
        });
        Px = group_sum<kG>(Px);
        Py = group_sum<kG>(Py);
        Pz = group_sum<kG>(Pz);
        Vx = group_sum<kG>(Vx);
        Vy = group_sum<kG>(Vy);
        Vz = group_sum<kG>(Vz);
        if (kTension) {
            Tx = group_sum<kG>(Tx);
            Ty = group_sum<kG>(Ty);
            Tz = group_sum<kG>(Tz);
        }
        if (act && (lane & (kG - 1u)) == 0u) {
            const float mv = c.m_other * c.visc * ai.y;
            float ax = c.m_other * Px + mv * Vx, ay = c.m_other * Py + mv * Vy,
                  az = c.m_other * Pz + mv * Vz;
            if (kTension) {
                const float tc = c.tension * c.m_other * ai.y;
                ax -= tc * Tx;
                ay -= tc * Ty;
                az -= tc * Tz;
            }
            if (kBlend) {
                acc_a[s] = make_float4(ax, ay, az, 0.0f);
            } else {
                const float4 a = acc_a[s];
                acc_a[s] = make_float4(a.x + ax, a.y + ay, a.z + az, a.w);
            }
        }
    }
}

// The walks' launches at the group width that walk_lanes picks for the grid
// walked.
template <bool kBlend, class... Args>
void launch_density_walk(int lanes, hipStream_t s, Args... args) {
    if (lanes == 16)
        hipLaunchKernelGGL((k_cross_density_walk<kBlend, 16>), dim3(kWalkBlocks), dim3(kBlock), 0,
                           s, args...);
    else if (lanes == 32)
        hipLaunchKernelGGL((k_cross_density_walk<kBlend, 32>), dim3(kWalkBlocks), dim3(kBlock), 0,
                           s, args...);
    else
        hipLaunchKernelGGL((k_cross_density_walk<kBlend, 64>), dim3(kWalkBlocks), dim3(kBlock), 0,
                           s, args...);
}

template <bool kTension, bool kBlend, class... Args>
void launch_force_walk(int lanes, hipStream_t s, Args... args) {
    if (lanes == 16)
        hipLaunchKernelGGL((k_cross_force_walk<kTension, kBlend, 16>), dim3(kWalkBlocks),
                           dim3(kBlock), 0, s, args...);
    else if (lanes == 32)
        hipLaunchKernelGGL((k_cross_force_walk<kTension, kBlend, 32>), dim3(kWalkBlocks),
                           dim3(kBlock), 0, s, args...);
    else
        hipLaunchKernelGGL((k_cross_force_walk<kTension, kBlend, 64>), dim3(kWalkBlocks),
                           dim3(kBlock), 0, s, args...);
}

// integrateD of one set (TWO_SCALE/Solver.cu:1163-1223 with lambda = 1):
// a = own-set + complementary acceleration (+ boundary), v += dt a, the
// velocity cap, x += dt v and the wall clamp.
__global__ __launch_bounds__(kBlock) void k_cross_integrate(
    const float4* __restrict__ tpos_a, const float4* __restrict__ tvel_a,
    const float4* __restrict__ acc_a, const uint32_t* __restrict__ n_a, CrossParams c,
    const float4* __restrict__ bacc, float dt, float4* __restrict__ pos_a,
    float4* __restrict__ vel_a, float4* __restrict__ acc_out) {
    const uint32_t s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= *n_a) return;
    const float4 xi = tpos_a[s];
    const float4 vi = tvel_a[s];
    const float4 a = acc_a[s];
    float ax = a.x, ay = a.y, az = a.z;
    if (bacc) {  // boundary particles (Solver.cu:1046-1048)
        const float4 b = bacc[s];
        ax += b.x;
        ay += b.y;
        az += b.z;
    }
    float vx = vi.x + dt * ax, vy = vi.y + dt * ay, vz = vi.z + dt * az;
    cap_velocity(c.vmax, vx, vy, vz);
    float x = xi.x + dt * vx, y = xi.y + dt * vy, z = xi.z + dt * vz;
    if (c.flags & SPH_FLAG_WALL_CLAMP) {
        const float e = c.restitution;
        if (x < c.lox) { x = c.lox; vx = -e * vx; }
        if (x > c.hix) { x = c.hix; vx = -e * vx; }
        if (y < c.loy) { y = c.loy; vy = -e * vy; }
        if (y > c.hiy) { y = c.hiy; vy = -e * vy; }
        if (z < c.loz) { z = c.loz; vz = -e * vz; }
        if (z > c.hiz) { z = c.hiz; vz = -e * vz; }
    }
    pos_a[s] = make_float4(x, y, z, xi.w);
    vel_a[s] = make_float4(vx, vy, vz, vi.w);
    if (c.err && !__builtin_isfinite(((x + y) + (z + vx)) + (vy + vz)))
        atomicOr(c.err, kDevErrNonFinite);
    if (acc_out) acc_out[s] = make_float4(ax, ay, az, 0.0f);
}

// ---------------------------------------------------------------------------
// Blended coupling (TWO_SCALE/Solver.cu:215-499, 732-1083, 1163-1289): every
// neighbour term carries the neighbour's blend weight lambda_j.  Each set's
// own sums walk its own grid in the internal x-order with the block-staged
// stencil walk (the kernels below); the complementary sums are the fixed
// coupling's walks (one lane group per listed particle) with lambda weights.  Blend
// state words: float2 (lambda, state bits).

constexpr int kBlendDensityTile = 1024;
constexpr int kBlendForceTile = 512;

// The own-set part of rho_i = m_A sum_{A, r < h_A} lambda_j W_hA + m_B
// sum_{B, r < h0} lambda_j 1/2 (W_h0 + W_h1), and A_i, B_i from it (the
// complementary walk adds the second sum and recomputes A_i, B_i for the
// particles it lists).  Targets in the internal order; results at the sorted
// slot i2r[s].
__global__ __launch_bounds__(kBlock) void k_blend_density(
    const float4* __restrict__ ipos_a, const float2* __restrict__ ibl_a,
    const uint32_t* __restrict__ cs_a, const uint32_t* __restrict__ ce_a,
    const uint4* __restrict__ bins_a, const uint32_t* __restrict__ i2r_a,
    const uint32_t* __restrict__ n_a, GridView ga, StepParams pa, CrossParams c,
    float* __restrict__ rho_a, float2* __restrict__ aux_a) {
    __shared__ float4 tile[kBlendDensityTile];
    __shared__ uint32_t red[2 * (kBlock / 64)];
    const uint32_t na = *n_a;
    const uint32_t s = xcd_local_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const bool active = s < na;
    const float4 xi = active ? ipos_a[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const Probe pr = make_probe(ga, xi.x, xi.y, xi.z);
    float own = 0.0f;
    auto pair = [&](const float4& xj, float lj) {
        const float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < pa.h2) {
            const float a = pa.h2 - r2;
            own += lj * (a * a * a);
        }
    };
    walk_stencil_staged<kBlock, kBlendDensityTile, true>(
        ga, pr, cs_a, ce_a, na, active, red,
        [&](uint32_t k, uint32_t j) {
            const float4 x = ipos_a[j];
            tile[k] = make_float4(x.x, x.y, x.z, ibl_a[j].x);
        },
        [&](uint32_t t0, uint32_t t1, uint32_t) {
            for (uint32_t t = t0; t < t1; ++t) {
                const float4 q = tile[t];
                keep_lds_quad(q);
                pair(q, q.w);
            }
        },
        [&](uint32_t b, uint32_t e) {
            for (uint32_t j = b; j < e; ++j) pair(ipos_a[j], ibl_a[j].x);
        },
        bins_a);
    if (!active) return;
    const float r = pa.c6m * own;
    const uint32_t o = i2r_a[s];
    rho_a[o] = r;
    const float inv = __builtin_amdgcn_rcpf(r);
    aux_a[o] = make_float2(c.stiffness * (r - c.rest_density) * inv * inv, inv);
}

// Blended acceleration: the lambda-weighted own-set sums (targets in the
// internal order) plus gravity, plus the complementary acceleration of the
// step's walk (xacc_a, -0 for the particles it does not list) and the
// boundary term; then the velocity update with the cap, the position update
// and the wall clamp, and integrateD's blend update; results at the sorted
// slot i2r[s].
template <bool kTension>
__global__ __launch_bounds__(kBlock) void k_blend_force_integrate(
    const float4* __restrict__ ipos_a, const float4* __restrict__ ivel_a,
    const float2* __restrict__ iaux_a, const float2* __restrict__ ibl_a,
    const uint32_t* __restrict__ cs_a, const uint32_t* __restrict__ ce_a,
    const uint4* __restrict__ bins_a, const uint32_t* __restrict__ i2r_a,
    const uint32_t* __restrict__ n_a, GridView ga, StepParams pa, CrossParams c,
    const float4* __restrict__ xacc_a, const float4* __restrict__ bacc, float dt, float inc,
    float4* __restrict__ pos_a, float4* __restrict__ vel_a, float2* __restrict__ nbl_a,
    float4* __restrict__ acc_out) {
    __shared__ float4 tx[kBlendForceTile];
    __shared__ float4 tv[kBlendForceTile];
    __shared__ float tl[kBlendForceTile];
    __shared__ uint32_t red[2 * (kBlock / 64)];
    const uint32_t na = *n_a;
    const uint32_t s = xcd_local_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const bool active = s < na;
    const float4 xi = active ? ipos_a[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 vi = active ? ivel_a[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float2 ai = active ? iaux_a[s] : make_float2(0.f, 1.f);
    const Probe pr = make_probe(ga, xi.x, xi.y, xi.z);
    float Px = 0.f, Py = 0.f, Pz = 0.f, Vx = 0.f, Vy = 0.f, Vz = 0.f, Tx = 0.f, Ty = 0.f,
          Tz = 0.f;
    const uint32_t h2_bits_m1 = __float_as_uint(pa.h2) - 1u;
    // xa = (x_j, A_j), vb = (v_j, B_j); 0 < r < h_A.
    auto pair = [&](const float4& xa, const float4& vb, float lj) {
        const float dx = xi.x - xa.x, dy = xi.y - xa.y, dz = xi.z - xa.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        if (__float_as_uint(r2) - 1u >= h2_bits_m1) return;
        const float inv_r = __builtin_amdgcn_rsqf(r2);
        const float hr = pa.h - r2 * inv_r;
        const float tpr = lj * (ai.x + xa.w) * hr * hr * inv_r;
        Px += tpr * dx;
        Py += tpr * dy;
        Pz += tpr * dz;
        const float tvs = lj * hr * vb.w;
        Vx += tvs * (vb.x - vi.x);
        Vy += tvs * (vb.y - vi.y);
        Vz += tvs * (vb.z - vi.z);
        if (kTension) {
            const float a = pa.h2 - r2;
            const float w = lj * (a * a * a);
            Tx += w * dx;
            Ty += w * dy;
            Tz += w * dz;
        }
    };
    walk_stencil_staged<kBlock, kBlendForceTile, true>(
        ga, pr, cs_a, ce_a, na, active, red,
        [&](uint32_t k, uint32_t j) {
            const float4 x = ipos_a[j], v = ivel_a[j];
            const float2 a = iaux_a[j];
            tx[k] = make_float4(x.x, x.y, x.z, a.x);
            tv[k] = make_float4(v.x, v.y, v.z, a.y);
            tl[k] = ibl_a[j].x;
        },
        [&](uint32_t t0, uint32_t t1, uint32_t) {
            for (uint32_t t = t0; t < t1; ++t) {
                const float4 q = tx[t];
                keep_lds_quad(q);
                pair(q, tv[t], tl[t]);
            }
        },
        [&](uint32_t b, uint32_t e) {
            for (uint32_t j = b; j < e; ++j) {
                const float4 x = ipos_a[j], v = ivel_a[j];
                const float2 a = iaux_a[j];
                pair(make_float4(x.x, x.y, x.z, a.x), make_float4(v.x, v.y, v.z, a.y),
                     ibl_a[j].x);
            }
        },
        bins_a);
    if (!active) return;
    const uint32_t o = i2r_a[s];  // the target's sorted slot
    const float vc = pa.visc * ai.y;
    float ax = pa.c45m * Px + vc * Vx + pa.gx;
    float ay = pa.c45m * Py + vc * Vy + pa.gy;
    float az = pa.c45m * Pz + vc * Vz + pa.gz;
    if (kTension) {
        const float tc = pa.tens * ai.y;
        ax -= tc * Tx;
        ay -= tc * Ty;
        az -= tc * Tz;
    }
    // The complementary sums (Solver.cu:394-499, weighted by lambda_j) of
    // the step's walk; -0 for the particles it does not list.
    const float4 xa = xacc_a[o];
    ax += xa.x;
    ay += xa.y;
    az += xa.z;
    if (bacc) {  // boundary particles (Solver.cu:1046-1048)
        const float4 b = bacc[o];
        ax += b.x;
        ay += b.y;
        az += b.z;
    }
    float vx = vi.x + dt * ax, vy = vi.y + dt * ay, vz = vi.z + dt * az;
    cap_velocity(c.vmax, vx, vy, vz);
    float x = xi.x + dt * vx, y = xi.y + dt * vy, z = xi.z + dt * vz;
    if (c.flags & SPH_FLAG_WALL_CLAMP) {
        const float e = c.restitution;
        if (x < c.lox) { x = c.lox; vx = -e * vx; }
        if (x > c.hix) { x = c.hix; vx = -e * vx; }
        if (y < c.loy) { y = c.loy; vy = -e * vy; }
        if (y > c.hiy) { y = c.hiy; vy = -e * vy; }
        if (z < c.loz) { z = c.loz; vz = -e * vz; }
        if (z > c.hiz) { z = c.hiz; vz = -e * vz; }
    }
    pos_a[o] = make_float4(x, y, z, xi.w);
    vel_a[o] = make_float4(vx, vy, vz, vi.w);
    if (c.err && !__builtin_isfinite(((x + y) + (z + vx)) + (vy + vz)))
        atomicOr(c.err, kDevErrNonFinite);
    if (acc_out) acc_out[o] = make_float4(ax, ay, az, 0.0f);
    // Blend update (Solver.cu:1242-1257): lambda += c[mode] inc with
    // c = {0, -1, +1, 0}; a fading-in particle reaching 1 settles.  Removal
    // (fading out, lambda <= 0) happens in the compaction.
    const float2 bl = ibl_a[s];
    uint32_t st = __float_as_uint(bl.y);
    const uint32_t mode = st & 3u;
    float lam = bl.x;
    if (mode == 1u) lam = __fsub_rn(lam, inc);
    if (mode == 2u) lam = __fadd_rn(lam, inc);
    if (mode == 2u && lam >= 1.0f) {
        st &= 4u;
        lam = 1.0f;
    }
    nbl_a[o] = make_float2(lam, __uint_as_float(st));
}

__global__ void k_gather_blend(const uint32_t* __restrict__ perm, const uint32_t* __restrict__ n_dev,
                               const float2* __restrict__ src, float2* __restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    dst[i] = src[perm[i]];
}

__global__ void k_fill_blend(const uint32_t* __restrict__ n_dev, float2 value,
                             float2* __restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    dst[i] = value;
}

__global__ void k_fill_float(const uint32_t* __restrict__ n_dev, float value,
                             float* __restrict__ dst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    dst[i] = value;
}

// Low particles whose id is in the sorted list `ids` and that are settled
// (mode 0; the surface bit 8 may be set) get mode 3 (split requested).
__global__ void k_mark_ids(const float4* __restrict__ pos, const uint32_t* __restrict__ n_dev,
                           const uint32_t* __restrict__ ids, uint32_t n_ids,
                           float2* __restrict__ bl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    const uint32_t id = __float_as_uint(pos[i].w);
    uint32_t lo = 0, hi = n_ids;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (ids[mid] < id)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < n_ids && ids[lo] == id) {
        const float2 b = bl[i];
        const uint32_t st = __float_as_uint(b.y);
        if ((st & 7u) == 0u) bl[i] = make_float2(b.x, __uint_as_float(st | 3u));
    }
}

// ---------------------------------------------------------------------------
// split / merge bookkeeping

struct Region {
    float lox, loy, loz, hix, hiy, hiz;
    float margin;      // merge hysteresis
    float child_d;     // child offset per axis (lattice offsets)
};

// How a split places its children and whether groups merge.
struct SplitParams {
    uint32_t merge;        // 0: never merge
    uint32_t ref_offsets;  // 1: r/sqrt(3) diagonals with r from the parent's density
    float m0;              // low-resolution mass
};

__device__ __forceinline__ bool inside(const Region& r, float x, float y, float z) {
    return x >= r.lox && x <= r.hix && y >= r.loy && y <= r.hiy && z >= r.loz && z <= r.hiz;
}

__device__ __forceinline__ bool outside_margin(const Region& r, float x, float y, float z) {
    return x < r.lox - r.margin || x > r.hix + r.margin || y < r.loy - r.margin ||
           y > r.hiy + r.margin || z < r.loz - r.margin || z > r.hiz + r.margin;
}

// High-resolution particles carry (slot << 29) | group in vel.w.
constexpr uint32_t kGroupMask = (1u << 29) - 1u;

// Blended coupling: a fading-out particle whose weight reached 0 is removed.
__device__ __forceinline__ bool blend_removed(const float2& b) {
    return (__float_as_uint(b.y) & 3u) == 1u && b.x <= 0.0f;
}

// Flags, per particle of either set: bit 0 = not kept, bit 1 = emits (a low
// particle its 8 children, a high merge leader the merged particle), bit 2
// (blended high set) = starts fading out.
//
// High set: children outside the expanded region record their state in their
// group's slots and count themselves (blended: only settled ones, and every
// live child marks its group in grp_live bit 0).
__global__ __launch_bounds__(kBlock) void k_merge_mark(
    const float4* __restrict__ pos_h, const float4* __restrict__ vel_h,
    const float2* __restrict__ bl_h, const uint32_t* __restrict__ n_h, Region rg, uint32_t merge,
    float4* __restrict__ grp_pos, float4* __restrict__ grp_vel, uint32_t* __restrict__ grp_cnt,
    uint32_t* __restrict__ grp_live, uint32_t num_groups) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= *n_h) return;
    const float4 x = pos_h[i];
    const float4 v = vel_h[i];
    const uint32_t tag = __float_as_uint(v.w);
    const uint32_t g = tag & kGroupMask, k = tag >> 29;
    if (g >= num_groups) return;
    if (bl_h) {
        const float2 b = bl_h[i];
        if (blend_removed(b)) return;
        atomicOr(&grp_live[g], 1u);
        if ((__float_as_uint(b.y) & 3u) != 0u) return;
    }
    if (merge && outside_margin(rg, x.x, x.y, x.z)) {
        grp_pos[static_cast<size_t>(g) * 8 + k] = x;
        grp_vel[static_cast<size_t>(g) * 8 + k] = v;
        atomicAdd(&grp_cnt[g], 1u);
    }
}

// The block's counts over its particles' flags (every thread of the block
// calls it; `live` false for slots past the set's count): counts[block] =
// kept (bit 0 clear), counts[nblocks + block] = emitting (bit 1 set).
__device__ __forceinline__ void block_flag_counts(bool live, uint32_t f,
                                                  uint32_t* __restrict__ counts,
                                                  uint32_t nblocks) {
    __shared__ uint32_t wc[2][kBlock / 64];
    const uint64_t b0 = __ballot(live && !(f & 1u)), b1 = __ballot(live && (f & 2u) != 0u);
    const uint32_t wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 0u) {
        wc[0][wave] = static_cast<uint32_t>(__popcll(b0));
        wc[1][wave] = static_cast<uint32_t>(__popcll(b1));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c0 = 0u, c1 = 0u;
        for (int w = 0; w < kBlock / 64; ++w) {
            c0 += wc[0][w];
            c1 += wc[1][w];
        }
        counts[blockIdx.x] = c0;
        counts[nblocks + blockIdx.x] = c1;
    }
}

// Low set.  Fixed: inside the region -> split and removed (3).  Blended:
// removed (1); or split (2) when settled inside the region or requested
// (state 3) and no child of its group is alive; every live particle marks
// its id in grp_live bit 1.  Then the block's counts (block_flag_counts).
__global__ __launch_bounds__(kBlock) void k_split_mark(const float4* __restrict__ pos_l,
                                                       const float2* __restrict__ bl_l,
                                                       const uint32_t* __restrict__ n_l, Region rg,
                                                       uint32_t* __restrict__ grp_live,
                                                       uint32_t num_groups,
                                                       uint32_t* __restrict__ flag_l,
                                                       uint32_t* __restrict__ counts_l,
                                                       uint32_t nbl) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < *n_l;
    uint32_t f = 0u;
    if (live) {
        const float4 x = pos_l[i];
        const bool in = inside(rg, x.x, x.y, x.z);
        if (!bl_l) {
            f = in ? 3u : 0u;
        } else {
            const float2 b = bl_l[i];
            if (blend_removed(b)) {
                f = 1u;
            } else {
                const uint32_t id = __float_as_uint(x.w);
                uint32_t children = 0u;
                if (id < num_groups) children = atomicOr(&grp_live[id], 2u) & 1u;
                const uint32_t mode = __float_as_uint(b.y) & 3u;
                const bool want = mode == 3u || (mode == 0u && in);
                f = (want && !children) ? 2u : 0u;
            }
        }
        flag_l[i] = f;
    }
    block_flag_counts(live, f, counts_l, nbl);
}

// High set.  Fixed: a group with 8 children outside the expanded region
// merges (1, leader 1 | 2).  Blended: removed (1); a group of 8 settled
// children outside, with no live low particle of its id, starts fading out
// (4, leader 4 | 2).  Then the block's counts (block_flag_counts).
__global__ __launch_bounds__(kBlock) void k_merge_flags(const float4* __restrict__ vel_h,
                                                        const float2* __restrict__ bl_h,
                                                        const uint32_t* __restrict__ n_h,
                                                        const uint32_t* __restrict__ grp_cnt,
                                                        const uint32_t* __restrict__ grp_live,
                                                        uint32_t num_groups,
                                                        uint32_t* __restrict__ flag_h,
                                                        uint32_t* __restrict__ counts_h,
                                                        uint32_t nbh) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < *n_h;
    uint32_t f = 0u;
    if (live) {
        const uint32_t tag = __float_as_uint(vel_h[i].w);
        const uint32_t g = tag & kGroupMask, k = tag >> 29;
        const uint32_t lead = (k == 0u) ? 2u : 0u;
        if (!bl_h) {
            const bool merge = g < num_groups && grp_cnt[g] == 8u;
            f = merge ? (1u | lead) : 0u;
        } else if (blend_removed(bl_h[i])) {
            f = 1u;
        } else {
            const bool merge = g < num_groups && grp_cnt[g] == 8u && !(grp_live[g] & 2u);
            f = merge ? (4u | lead) : 0u;
        }
        flag_h[i] = f;
    }
    block_flag_counts(live, f, counts_h, nbh);
}

// Words of the split/merge totals block `tot`: 0-3 the four scan totals
// (kept low, split low, kept high, merge leaders), 4-5 the new counts, 6-7 the
// counts the step ran with, then the error word, the step stamp and the
// scan's ticket.
constexpr uint32_t kTotWords = 16;
constexpr uint32_t kErrorWordIndex = 8;
constexpr uint32_t kStampWord = 9;
constexpr uint32_t kTicketWord = 10;

// Exclusive scans of the four block-count sequences (one workgroup each),
// over the blocks that hold live particles (the scatters read no other
// block's offsets).  tot[] = {keepL, splitL, keepH, leadH, newL, newH}.
__global__ __launch_bounds__(1024) void k_scan_counts(uint32_t* __restrict__ cl, uint32_t nbl,
                                                      uint32_t* __restrict__ ch, uint32_t nbh,
                                                      uint32_t* __restrict__ n_l,
                                                      uint32_t* __restrict__ n_h,
                                                      uint32_t cap_l, uint32_t cap_h,
                                                      uint32_t* __restrict__ tot,
                                                      uint32_t* __restrict__ error_word,
                                                      uint32_t* __restrict__ stamp) {
    // Block q scans count sequence q (kept low, split low, kept high, merge
    // leaders): a thread sums a contiguous chunk of at most a few counts, the
    // chunk sums are scanned with wave prefix sums and the 16 wave totals,
    // the chunk is rewritten as exclusive offsets, and tot[q] gets the total.
    __shared__ uint32_t wtot[16];
    const uint32_t q = blockIdx.x, t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t* d = (q == 0) ? cl : (q == 1) ? cl + nbl : (q == 2) ? ch : ch + nbh;
    const uint32_t live_blocks = ((q < 2 ? *n_l : *n_h) + kBlock - 1u) / kBlock;
    const uint32_t len = min((q < 2) ? nbl : nbh, live_blocks);
    const uint32_t chunk = (len + 1023) / 1024;
    const uint32_t b = min(t * chunk, len), e = min(b + chunk, len);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += d[i];
    uint32_t x = sum;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(x), off));
        if (lane >= static_cast<uint32_t>(off)) x += y;
    }
    if (lane == 63) wtot[wave] = x;
    __syncthreads();
    uint32_t before = 0;
    for (uint32_t w = 0; w < wave; ++w) before += wtot[w];
    uint32_t run = before + x - sum;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t v = d[i];
        d[i] = run;
        run += v;
    }
    if (t != 1023) return;
    // Publish the total; the last of the four blocks to do so commits the
    // step's counts (it reads the totals after the others' releases).
    __hip_atomic_store(&tot[q], before + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t* ticket = tot + kTicketWord;
    const uint32_t done =
        __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (done != gridDim.x - 1u) return;
    __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t tq[4];
    for (int k = 0; k < 4; ++k)
        tq[k] = __hip_atomic_load(&tot[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // New counts from the totals (kept + merged low; kept + 8 x split high),
    // clamped to the capacities; a split or merge past them was dropped.
    // tot[6], tot[7] keep the counts the step ran with (the scatters' bound,
    // and for inspection).  Every block read *n_l, *n_h before its ticket.
    const uint64_t new_l = static_cast<uint64_t>(tq[0]) + tq[3];
    const uint64_t new_h = static_cast<uint64_t>(tq[2]) + 8ull * tq[1];
    if (new_l > cap_l || new_h > cap_h) atomicOr(error_word, 2u);
    tot[4] = static_cast<uint32_t>(new_l < cap_l ? new_l : cap_l);
    tot[5] = static_cast<uint32_t>(new_h < cap_h ? new_h : cap_h);
    tot[6] = *n_l;
    tot[7] = *n_h;
    *n_l = tot[4];
    *n_h = tot[5];
    // The step stamp advances, so the next step's occupancy marks differ
    // from this step's.
    *stamp += 1u;
}

// In-block exclusive rank of a predicate (block of 256, in index order).
__device__ __forceinline__ uint32_t block_rank(bool pred, uint32_t* wsum) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(pred);
    const uint32_t below = static_cast<uint32_t>(
        __popcll(m & (lane == 0 ? 0ull : (~0ull >> (64 - lane)))));
    if (lane == 0) wsum[wave] = static_cast<uint32_t>(__popcll(m));
    __syncthreads();
    uint32_t off = 0;
    for (int w = 0; w < wave; ++w) off += wsum[w];
    __syncthreads();
    return off + below;
}

// Low set: kept particles to new_l[keep rank] (blended: a splitting parent
// starts fading out); emitting ones write 8 children to
// new_h[keepH + 8 * split rank + k].  ids: child id = id_base + 8 * pid + k,
// group = pid, slot = k.  Child offsets: +/- child_d per axis, or (reference
// offsets) r/sqrt(3) per axis with r = 1/2 (3 m0 / (4 pi rho))^(1/3)
// (insertHighResParticles, TWO_SCALE/Solver.cu:541-586).  Blend state words
// (bl_l != NULL): children start at lambda 0, state 6.
__device__ __forceinline__ void scatter_low(
    uint32_t blk, const float4* __restrict__ pos_l, const float4* __restrict__ vel_l,
    const float* __restrict__ rho_l, const float2* __restrict__ bl_l,
    const uint32_t* __restrict__ flag_l, const uint32_t* __restrict__ cl, uint32_t nbl,
    const uint32_t* __restrict__ tot,
    Region rg, SplitParams sp, uint32_t id_base, uint32_t cap_l, uint32_t cap_h,
    uint32_t* __restrict__ grp_live, uint32_t num_groups, const float* __restrict__ col_l,
    float4* __restrict__ new_pos_l, float4* __restrict__ new_vel_l, float2* __restrict__ new_bl_l,
    float* __restrict__ new_col_l, float4* __restrict__ new_pos_h, float4* __restrict__ new_vel_h,
    float2* __restrict__ new_bl_h, float* __restrict__ new_col_h) {
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t i = blk * kBlock + threadIdx.x;
    const uint32_t n = tot[6];  // the count the step ran with
    const bool live = i < n;
    const uint32_t f = live ? flag_l[i] : 1u;
    const uint32_t rk = block_rank(live && !(f & 1u), wsum);
    const uint32_t rs = block_rank(live && (f & 2u), wsum);
    if (!live) return;
    const float4 x = pos_l[i];
    const float4 v = vel_l[i];
    const uint32_t pid = __float_as_uint(x.w);
    if (grp_live && pid < num_groups) grp_live[pid] = 0u;
    if (!(f & 1u)) {
        const uint32_t dst = cl[blk] + rk;
        if (dst < cap_l) {
            new_pos_l[dst] = x;
            new_vel_l[dst] = make_float4(v.x, v.y, v.z, 0.0f);
            if (bl_l) {
                const float2 b = bl_l[i];
                new_bl_l[dst] = (f & 2u) ? make_float2(b.x, __uint_as_float(1u)) : b;
            }
            if (col_l) new_col_l[dst] = col_l[i];
        }
    }
    if (f & 2u) {
        const uint32_t base = tot[2] + 8u * (cl[nbl + blk] + rs);
        float d = rg.child_d;
        if (sp.ref_offsets) {
            // 3 / (4 pi) and 1/sqrt(3) as the reference's float constants
            const float r = 0.5f * cbrtf(0.238732414637843f * sp.m0 / rho_l[i]);
            d = r * 0.57735026919f;
        }
        for (int k = 0; k < 8; ++k) {
            const uint32_t dst = base + k;
            if (dst >= cap_h) break;
            new_pos_h[dst] = make_float4(x.x + child_offset(k, 0, d), x.y + child_offset(k, 1, d),
                                         x.z + child_offset(k, 2, d),
                                         __uint_as_float(id_base + 8u * pid + k));
            new_vel_h[dst] = make_float4(v.x, v.y, v.z,
                                         __uint_as_float((static_cast<uint32_t>(k) << 29) |
                                                         (pid & kGroupMask)));
            if (bl_l) new_bl_h[dst] = make_float2(0.0f, __uint_as_float(6u));
            if (col_l) new_col_h[dst] = col_l[i];
        }
    }
}

// High set: kept particles to new_h[keep rank] (blended: merging siblings
// start fading out); merge leaders emit the merged particle (centroid, mean
// velocity, id = group; blended: lambda 0, state 2) to
// new_l[keepL + lead rank].  Every high particle clears its group's counter
// and live bits for the next step.
__device__ __forceinline__ void scatter_high(
    uint32_t blk, const float4* __restrict__ pos_h, const float4* __restrict__ vel_h,
    const float2* __restrict__ bl_h, const uint32_t* __restrict__ flag_h,
    const uint32_t* __restrict__ ch, uint32_t nbh,
    const uint32_t* __restrict__ tot, const float4* __restrict__ grp_pos,
    const float4* __restrict__ grp_vel, uint32_t* __restrict__ grp_cnt,
    uint32_t* __restrict__ grp_live, uint32_t num_groups, uint32_t cap_l, uint32_t cap_h,
    const float* __restrict__ col_h, float4* __restrict__ new_pos_l,
    float4* __restrict__ new_vel_l, float2* __restrict__ new_bl_l, float* __restrict__ new_col_l,
    float4* __restrict__ new_pos_h, float4* __restrict__ new_vel_h,
    float2* __restrict__ new_bl_h, float* __restrict__ new_col_h) {
    __shared__ uint32_t wsum[kBlock / 64];
    const uint32_t i = blk * kBlock + threadIdx.x;
    const uint32_t n = tot[7];
    const bool live = i < n;
    const uint32_t f = live ? flag_h[i] : 1u;
    const uint32_t rk = block_rank(live && !(f & 1u), wsum);
    const uint32_t rl = block_rank(live && (f & 2u), wsum);
    if (!live) return;
    const float4 v = vel_h[i];
    const uint32_t g = __float_as_uint(v.w) & kGroupMask;
    if (g < num_groups) {
        grp_cnt[g] = 0u;
        if (grp_live) grp_live[g] = 0u;
    }
    if (!(f & 1u)) {
        const uint32_t dst = ch[blk] + rk;
        if (dst < cap_h) {
            new_pos_h[dst] = pos_h[i];
            new_vel_h[dst] = v;
            if (bl_h) {
                const float2 b = bl_h[i];
                new_bl_h[dst] =
                    (f & 4u) ? make_float2(b.x, __uint_as_float((__float_as_uint(b.y) & 4u) | 1u))
                             : b;
            }
            if (col_h) new_col_h[dst] = col_h[i];
        }
    }
    if (f & 2u) {
        const uint32_t dst = tot[0] + ch[nbh + blk] + rl;
        if (dst < cap_l) {
            float sx = 0.f, sy = 0.f, sz = 0.f, ux = 0.f, uy = 0.f, uz = 0.f;
            for (int k = 0; k < 8; ++k) {
                const float4 p = grp_pos[static_cast<size_t>(g) * 8 + k];
                const float4 q = grp_vel[static_cast<size_t>(g) * 8 + k];
                sx += p.x;
                sy += p.y;
                sz += p.z;
                ux += q.x;
                uy += q.y;
                uz += q.z;
            }
            new_pos_l[dst] = make_float4(0.125f * sx, 0.125f * sy, 0.125f * sz, __uint_as_float(g));
            new_vel_l[dst] = make_float4(0.125f * ux, 0.125f * uy, 0.125f * uz, 0.0f);
            if (bl_h) new_bl_l[dst] = make_float2(0.0f, __uint_as_float(2u));
            if (col_h) new_col_l[dst] = col_h[i];  // the leader's (slot 0) value
        }
    }
}

// Both sets' scatters in one launch: blocks 0 .. nbl - 1 take the low set's
// slots, the next nbh blocks the high set's.  Their writes are disjoint: the
// low set's kept particles go to new_l[0, keepL) and its children to
// new_h[keepH, ...), the high set's kept particles to new_h[0, keepH) and its
// merged particles to new_l[keepL, ...); both only zero the group words.
__global__ __launch_bounds__(kBlock) void k_scatter(
    const float4* __restrict__ pos_l, const float4* __restrict__ vel_l,
    const float* __restrict__ rho_l, const float2* __restrict__ bl_l,
    const uint32_t* __restrict__ flag_l, const uint32_t* __restrict__ cl, uint32_t nbl,
    const float* __restrict__ col_l, const float4* __restrict__ pos_h,
    const float4* __restrict__ vel_h, const float2* __restrict__ bl_h,
    const uint32_t* __restrict__ flag_h, const uint32_t* __restrict__ ch, uint32_t nbh,
    const float* __restrict__ col_h, const uint32_t* __restrict__ tot, Region rg, SplitParams sp,
    uint32_t id_base, uint32_t cap_l, uint32_t cap_h, const float4* __restrict__ grp_pos,
    const float4* __restrict__ grp_vel, uint32_t* __restrict__ grp_cnt,
    uint32_t* __restrict__ grp_live, uint32_t num_groups, float4* __restrict__ new_pos_l,
    float4* __restrict__ new_vel_l, float2* __restrict__ new_bl_l, float* __restrict__ new_col_l,
    float4* __restrict__ new_pos_h, float4* __restrict__ new_vel_h,
    float2* __restrict__ new_bl_h, float* __restrict__ new_col_h) {
    if (blockIdx.x < nbl)
        scatter_low(blockIdx.x, pos_l, vel_l, rho_l, bl_l, flag_l, cl, nbl, tot, rg, sp, id_base,
                    cap_l, cap_h, grp_live, num_groups, col_l, new_pos_l, new_vel_l, new_bl_l,
                    new_col_l, new_pos_h, new_vel_h, new_bl_h, new_col_h);
    else
        scatter_high(blockIdx.x - nbl, pos_h, vel_h, bl_h, flag_h, ch, nbh, tot, grp_pos, grp_vel,
                     grp_cnt, grp_live, num_groups, cap_l, cap_h, col_h, new_pos_l, new_vel_l,
                     new_bl_l, new_col_l, new_pos_h, new_vel_h, new_bl_h, new_col_h);
}

__global__ void k_pack_two_scale(const float* __restrict__ pxyz, const float* __restrict__ vxyz,
                                 const uint32_t* __restrict__ ids,
                                 const uint32_t* __restrict__ tags, uint32_t n,
                                 float4* __restrict__ pos, float4* __restrict__ vel) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos[i] = make_float4(pxyz[3 * i], pxyz[3 * i + 1], pxyz[3 * i + 2],
                         __uint_as_float(ids ? ids[i] : i));
    const float w = __uint_as_float(tags ? tags[i] : 0u);
    vel[i] = vxyz ? make_float4(vxyz[3 * i], vxyz[3 * i + 1], vxyz[3 * i + 2], w)
                  : make_float4(0.f, 0.f, 0.f, w);
}

__global__ void k_copy_state(const float4* __restrict__ sp, const float4* __restrict__ sv,
                             const float2* __restrict__ sb, const uint32_t* __restrict__ n_dev,
                             float4* __restrict__ dp, float4* __restrict__ dv,
                             float2* __restrict__ db) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *n_dev) return;
    dp[i] = sp[i];
    dv[i] = sv[i];
    if (sb) db[i] = sb[i];
}

inline dim3 blocks_for(uint32_t n) { return dim3((n + kBlock - 1) / kBlock > 0 ? (n + kBlock - 1) / kBlock : 1); }

}  // namespace
}  // namespace sph

using namespace sph;

// One resolution's buffers; `state_*` is the compacted state the next step
// reads, `pos/vel` (in StepBuffers) the integrated state before split/merge.
struct TwoScaleSet {
    sph_config cfg{};        // h, mass and grid of this resolution
    uint32_t cap = 0, num_cells = 0, key_bits = 1;
    StepBuffers buf{};
    float4* state_pos = nullptr;
    float4* state_vel = nullptr;
    float4* accel = nullptr;
    uint32_t* flag = nullptr;
    uint32_t* counts = nullptr;   // 2 * nblocks
    // Blend state words (lambda, state bits) of the blended coupling: the
    // compacted state, its sorted copy and the step's updated words (sorted).
    float2* state_blend = nullptr;
    float2* tblend = nullptr;
    float2* nblend = nullptr;
    // Colour values (SPH_DEBUG_COLOR): the step's sorted order and the
    // compacted order, carried through split/merge.
    float* tcolor = nullptr;
    float* color = nullptr;
    uint32_t* walk = nullptr;     // the step's particles with a complementary walk (sorted slots)
    float2* iblend = nullptr;     // blended coupling: the step's blend words in the internal order
};

struct sph_two_scale {
    sph_config cfg{};            // low-resolution configuration
    TwoScaleSet set[2];          // 0 = low (h0, m0), 1 = high (h1, m1)
    Region region{};
    uint32_t num_groups = 0;     // low-resolution ids ever issued (group ids)
    uint32_t id_base = 0;        // first high-resolution id
    float4* grp_pos = nullptr;
    float4* grp_vel = nullptr;
    uint32_t* grp_cnt = nullptr;
    uint32_t* tot = nullptr;     // 16 words: kTotWords
    uint32_t* error_word = nullptr;
    uint32_t* stamp = nullptr;   // steps run (the occupancy marks' base)
    uint32_t* occ = nullptr;     // per low-grid cell: 1 when it holds a high particle
    uint32_t* grp_live = nullptr;  // per group: bit 0 live child, bit 1 live low particle (blended)
    Walls walls{};               // boundary particles, sorted on the low grid
    float4* bacc[2] = {nullptr, nullptr};  // per set: boundary acceleration per sorted slot
    sph2_options opt{SPH2_COUPLING_FIXED, SPH2_SPLIT_LATTICE, 1u, 0.0125f, 0.0f};
    hipStream_t stream = nullptr;
    // The low set's passes run on a second stream forked from and joined back
    // to `stream` (branches of the step's graph when captured), with events
    // ev[] at the points where one set's passes read the other's results.
    hipStream_t stream2 = nullptr;
    hipEvent_t ev[6] = {};
    // Per set: the tune word of its own-set density and force passes (the
    // fields of the single-scale tune word; sph2_set_tune, sph2_autotune).
    uint32_t tune[2] = {0u, 0u};
    void* arena = nullptr;
    hipGraphExec_t graph = nullptr;
    float graph_dt = 0.0f;
    bool graph_accel = false;
    uint32_t graph_debug = 0;
    uint32_t debug = 0;
    uint32_t* surf = nullptr;    // per low id: at the free surface (surface detector)
    uint32_t* walk_count = nullptr;  // per set: the counters of its walk list's segments
    // Renderer buffers per set (sph2_enable_render_buffers): AoS xyz, colour, state.
    float* render_xyz[2] = {nullptr, nullptr};
    float* render_color[2] = {nullptr, nullptr};
    int32_t* render_state[2] = {nullptr, nullptr};
    void* render_arena = nullptr;
    uint32_t render_flags = 0;
};

namespace {

void destroy_streams2(sph_two_scale* t) {
    for (hipEvent_t& ev : t->ev)
        if (ev) (void)hipEventDestroy(ev);
    if (t->stream2) (void)hipStreamDestroy(t->stream2);
    if (t->stream) (void)hipStreamDestroy(t->stream);
}

// `to` waits for the work issued on `from` so far (event k).
int wait_for(sph_two_scale* t, int k, hipStream_t from, hipStream_t to) {
    SPH_HIP(hipEventRecord(t->ev[k], from));
    SPH_HIP(hipStreamWaitEvent(to, t->ev[k], 0));
    return SPH_OK;
}

void destroy_graph2(sph_two_scale* t) {
    if (t->graph) {
        (void)hipGraphExecDestroy(t->graph);
        t->graph = nullptr;
    }
}

uint32_t* walk_counts(const sph_two_scale* t, int target) {
    return t->walk_count + static_cast<uint32_t>(target) * kWalkSegments * kWalkCounterStride;
}

// The skip test of set `target`'s walk over the other set: the low-grid cells
// holding the other set's particles, and their dilated marks.
Nearby nearby(const sph_two_scale* t, int target) {
    const uint32_t C0 = t->set[0].num_cells;
    Nearby nb;
    nb.occ = (target == 0) ? t->occ : t->occ + C0;
    nb.near = t->occ + 2 * C0;
    nb.stamp = t->stamp;
    nb.bit = (target == 0) ? 1u : 2u;
    return nb;
}

CrossParams cross_params(const sph_two_scale* t, int target) {
    const sph_config& lo = t->set[0].cfg;
    const sph_config& hi = t->set[1].cfg;
    CrossParams c{};
    c.h0 = lo.h;
    c.h1 = hi.h;
    c.h02 = lo.h * lo.h;
    c.h12 = hi.h * hi.h;
    const double h0 = lo.h, h1 = hi.h;
    c.c6_0 = static_cast<float>(315.0 / (64.0 * kPi * std::pow(h0, 9)));
    c.c6_1 = static_cast<float>(315.0 / (64.0 * kPi * std::pow(h1, 9)));
    c.c45_0 = static_cast<float>(45.0 / (kPi * std::pow(h0, 6)));
    c.c45_1 = static_cast<float>(45.0 / (kPi * std::pow(h1, 6)));
    c.m_other = t->set[1 - target].cfg.mass;
    c.stiffness = lo.stiffness;
    c.rest_density = lo.rest_density;
    c.visc = lo.viscosity;
    c.tension = lo.tension;
    c.gx = lo.gravity[0];
    c.gy = lo.gravity[1];
    c.gz = lo.gravity[2];
    c.restitution = lo.restitution;
    c.lox = lo.domain_lo[0];
    c.loy = lo.domain_lo[1];
    c.loz = lo.domain_lo[2];
    c.hix = lo.domain_hi[0];
    c.hiy = lo.domain_hi[1];
    c.hiz = lo.domain_hi[2];
    c.flags = lo.flags;
    c.vmax = 0.0f;
    c.err = radix_sort_error_word(t->set[target].buf.sort_temp);
    return c;
}

bool blended(const sph_two_scale* t) { return t->opt.coupling == SPH2_COUPLING_BLENDED; }

const float4* wall_acc(const sph_two_scale* t, int r) { return t->walls.n ? t->bacc[r] : nullptr; }

// c h_A / dt in FP32, as integrateD computes 0.8f * h / dt (Solver.cu:1214-1217).
float velocity_limit(const sph_two_scale* t, int target, float dt) {
    return t->opt.velocity_cap > 0.0f ? t->opt.velocity_cap * t->set[target].cfg.h / dt : 0.0f;
}

// The integrated states -> split/merge -> compacted states and new counts.
// The split/merge stage in two parts.  The marks: the high set's merge marks
// and flags and per-block counts on hs, the low set's split flags and
// per-block counts on ls.  Under the fixed coupling the two sets' marks are
// independent (hs and ls may be the two branches of a step); the blended
// coupling's split marks read the merge marks' grp_live bits, so there both
// must be one stream.  The commit (on t->stream, after both): the scan of the
// counts, the scatters and the new counts.
int launch_split_merge_marks(sph_two_scale* t, hipStream_t hs, hipStream_t ls) {
    TwoScaleSet& L = t->set[0];
    TwoScaleSet& H = t->set[1];
    const uint32_t nbl = blocks_for(L.cap).x, nbh = blocks_for(H.cap).x;
    // Blended coupling: the step's updated blend words; fixed: none (NULL).
    const bool bl = blended(t);
    const float2* bl_l = bl ? L.nblend : nullptr;
    const float2* bl_h = bl ? H.nblend : nullptr;
    uint32_t* live = bl ? t->grp_live : nullptr;
    SplitParams sp{};
    sp.merge = t->opt.merge;
    sp.ref_offsets = t->opt.split_offsets == SPH2_SPLIT_REFERENCE ? 1u : 0u;
    sp.m0 = L.cfg.mass;
    if (bl && hs != ls) return set_error(SPH_ERR_INVALID_ARGUMENT, "blended marks need one stream");
    // High marks first: a split waits for its group's live children.
    hipLaunchKernelGGL(k_merge_mark, dim3(nbh), dim3(kBlock), 0, hs, H.buf.pos, H.buf.vel, bl_h,
                       H.buf.n_dev, t->region, sp.merge, t->grp_pos, t->grp_vel, t->grp_cnt,
                       live, t->num_groups);
    hipLaunchKernelGGL(k_split_mark, dim3(nbl), dim3(kBlock), 0, ls, L.buf.pos, bl_l, L.buf.n_dev,
                       t->region, live, t->num_groups, L.flag, L.counts, nbl);
    hipLaunchKernelGGL(k_merge_flags, dim3(nbh), dim3(kBlock), 0, hs, H.buf.vel, bl_h,
                       H.buf.n_dev, t->grp_cnt, live, t->num_groups, H.flag, H.counts, nbh);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

int launch_split_merge_commit(sph_two_scale* t) {
    TwoScaleSet& L = t->set[0];
    TwoScaleSet& H = t->set[1];
    hipStream_t st = t->stream;
    const uint32_t nbl = blocks_for(L.cap).x, nbh = blocks_for(H.cap).x;
    const bool bl = blended(t);
    const float2* bl_l = bl ? L.nblend : nullptr;
    const float2* bl_h = bl ? H.nblend : nullptr;
    uint32_t* live = bl ? t->grp_live : nullptr;
    SplitParams sp{};
    sp.merge = t->opt.merge;
    sp.ref_offsets = t->opt.split_offsets == SPH2_SPLIT_REFERENCE ? 1u : 0u;
    sp.m0 = L.cfg.mass;
    hipLaunchKernelGGL(k_scan_counts, dim3(4), dim3(1024), 0, st, L.counts, nbl, H.counts, nbh,
                       L.buf.n_dev, H.buf.n_dev, L.cap, H.cap, t->tot, t->error_word, t->stamp);
    const bool col = (t->debug & SPH_DEBUG_COLOR) != 0u;
    hipLaunchKernelGGL(k_scatter, dim3(nbl + nbh), dim3(kBlock), 0, st, L.buf.pos, L.buf.vel,
                       L.buf.rho, bl_l, L.flag, L.counts, nbl, col ? L.tcolor : nullptr, H.buf.pos,
                       H.buf.vel, bl_h, H.flag, H.counts, nbh, col ? H.tcolor : nullptr, t->tot,
                       t->region, sp, t->id_base, L.cap, H.cap, t->grp_pos, t->grp_vel,
                       t->grp_cnt, live, t->num_groups, L.state_pos, L.state_vel, L.state_blend,
                       L.color, H.state_pos, H.state_vel, H.state_blend, H.color);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

// The whole split/merge stage on t->stream.
int launch_split_merge(sph_two_scale* t) {
    const int rc = launch_split_merge_marks(t, t->stream, t->stream);
    return rc == SPH_OK ? launch_split_merge_commit(t) : rc;
}

// Blended coupling: lambda-weighted densities of both sets, then the
// lambda-weighted forces with integration and the blend update, then the
// split/merge stage.
// The blended coupling's passes (Solver.cu:215-499, 732-1083) in the fixed
// coupling's shape: each set's lambda-weighted own-set density; the
// complementary sums of the particles with a non-empty walk, one lane group each
// (k_cross_list lists them and finalises the others), weighted by lambda_j;
// then each set's own-set force with the integration and blend update, which
// adds the walk's complementary acceleration (kept in the set's own-set
// acceleration buffer, unused by this coupling).  The high set's passes run
// on t->stream, the low set's on t->stream2; the complementary forces wait
// for both sets' final A, B.
int launch_blended_passes(sph_two_scale* t, const StepParams* P, const GridView* G,
                          const GridView* X, float dt, bool accel_out) {
    const hipStream_t st = t->stream, sl = t->stream2;
    const hipStream_t sst[2] = {sl, st};
    int rc = wait_for(t, 2, st, sl);  // the dilation and the high set's sorted state
    if (rc != SPH_OK) return rc;
    for (int r = 1; r >= 0; --r) {
        TwoScaleSet& A = t->set[r];
        TwoScaleSet& B = t->set[1 - r];
        const hipStream_t s = sst[r];
        const CrossParams c = cross_params(t, r);
        hipLaunchKernelGGL(k_blend_density, blocks_for(A.cap), dim3(kBlock), 0, s, A.buf.ipos,
                           A.iblend, A.buf.cell_start, A.buf.cell_end, A.buf.xbins, A.buf.i2r,
                           A.buf.n_dev, G[r], P[r], c, A.buf.rho, A.buf.aux);
        hipLaunchKernelGGL(k_cross_list, blocks_for(A.cap), dim3(kBlock), 0, s, A.buf.tpos,
                           A.buf.n_dev, B.buf.n_dev, nearby(t, r), X[0], c, A.buf.rho, A.buf.aux,
                           A.accel, walk_counts(t, r), A.walk, walk_segment_cap(A.cap));
        launch_density_walk<true>(walk_lanes(X[1 - r]), s, A.buf.tpos, walk_counts(t, r), A.walk,
                                  walk_segment_cap(A.cap), B.buf.tpos, B.tblend,
                                  B.buf.cell_start, B.buf.cell_end, B.buf.n_dev, X[1 - r], c,
                                  A.buf.rho, A.buf.aux);
    }
    SPH_HIP(hipEventRecord(t->ev[3], sl));  // the low set's A, B
    SPH_HIP(hipEventRecord(t->ev[4], st));  // the high set's A, B
    SPH_HIP(hipStreamWaitEvent(st, t->ev[3], 0));
    SPH_HIP(hipStreamWaitEvent(sl, t->ev[4], 0));
    for (int r = 1; r >= 0; --r) {
        TwoScaleSet& A = t->set[r];
        TwoScaleSet& B = t->set[1 - r];
        const hipStream_t s = sst[r];
        CrossParams c = cross_params(t, r);
        if (c.tension != 0.0f)
            launch_force_walk<true, true>(walk_lanes(X[1 - r]), s, A.buf.tpos, A.buf.tvel,
                                          A.buf.aux, walk_counts(t, r), A.walk,
                                          walk_segment_cap(A.cap), B.buf.tpos, B.buf.tvel,
                                          B.buf.aux, B.tblend, B.buf.cell_start, B.buf.cell_end,
                                          B.buf.n_dev, X[1 - r], c, A.accel);
        else
            launch_force_walk<false, true>(walk_lanes(X[1 - r]), s, A.buf.tpos, A.buf.tvel,
                                          A.buf.aux, walk_counts(t, r), A.walk,
                                          walk_segment_cap(A.cap), B.buf.tpos, B.buf.tvel,
                                          B.buf.aux, B.tblend, B.buf.cell_start, B.buf.cell_end,
                                          B.buf.n_dev, X[1 - r], c, A.accel);
        c.vmax = velocity_limit(t, r, dt);
        float4* acc = accel_out ? A.buf.accel : nullptr;
        const float inc = t->opt.blend_increment;
        rc = launch_gather_aux(P[r], A.buf, s);  // the final A, B in the internal order
        if (rc != SPH_OK) return rc;
        if (P[r].tens != 0.0f)
            hipLaunchKernelGGL(k_blend_force_integrate<true>, blocks_for(A.cap), dim3(kBlock), 0,
                               s, A.buf.ipos, A.buf.ivel, A.buf.iaux, A.iblend, A.buf.cell_start,
                               A.buf.cell_end, A.buf.xbins, A.buf.i2r, A.buf.n_dev, G[r], P[r], c,
                               A.accel, wall_acc(t, r), dt, inc, A.buf.pos, A.buf.vel, A.nblend,
                               acc);
        else
            hipLaunchKernelGGL(k_blend_force_integrate<false>, blocks_for(A.cap), dim3(kBlock), 0,
                               s, A.buf.ipos, A.buf.ivel, A.buf.iaux, A.iblend, A.buf.cell_start,
                               A.buf.cell_end, A.buf.xbins, A.buf.i2r, A.buf.n_dev, G[r], P[r], c,
                               A.accel, wall_acc(t, r), dt, inc, A.buf.pos, A.buf.vel, A.nblend,
                               acc);
    }
    SPH_HIP(hipGetLastError());
    rc = wait_for(t, 5, sl, st);  // join
    if (rc != SPH_OK) return rc;
    return launch_split_merge(t);
}

// Each set's step parameters (with its own-set tune word), grid view and the
// cross-search view of its grid (searched to h0, the coupling cutoff).
void set_views(const sph_two_scale* t, StepParams* P, GridView* G, GridView* X) {
    for (int r = 0; r < 2; ++r) {
        const TwoScaleSet& S = t->set[r];
        P[r] = make_step_params(S.cfg, S.cap);
        P[r].tune = t->tune[r];
        G[r] = make_grid_view(P[r].ox, P[r].oy, P[r].oz, P[r].spacing, P[r].dx, P[r].dy, P[r].dz,
                              P[r].h);
        X[r] = make_grid_view(P[r].ox, P[r].oy, P[r].oz, P[r].spacing, P[r].dx, P[r].dy, P[r].dz,
                              t->set[0].cfg.h);
    }
}

int launch_two_scale_step(sph_two_scale* t, float dt, bool accel_out) {
    hipStream_t st = t->stream;
    StepParams P[2];
    GridView G[2], X[2];
    set_views(t, P, G, X);
    // The two sets' passes run side by side: the high set's on t->stream, the
    // low set's on t->stream2.  They meet where a pass reads the other set's
    // results of this step: the occupancy dilation (both sets' marks), the
    // complementary walks (the other set's sorted state and cell ranges, and
    // the dilation), the complementary forces (the other set's final A, B)
    // and the split/merge stage (everything).
    const hipStream_t sst[2] = {t->stream2, st};
    const hipStream_t sl = t->stream2;
    const uint32_t C0 = t->set[0].num_cells;
    int rc = wait_for(t, 0, st, sl);  // fork
    if (rc != SPH_OK) return rc;
    // Neighbourhoods (the reference runs high then low, Solver.cu:1576-1577;
    // the sets do not interact here).  The state the step reads is the
    // compacted state; the reorder gathers from it.  Then each set marks the
    // low-grid cells it occupies (occ_h for the high set, occ_l for the low)
    // and computes its colour values and boundary accelerations.
    for (int r = 1; r >= 0; --r) {
        TwoScaleSet& S = t->set[r];
        const hipStream_t s = sst[r];
        StepBuffers b = S.buf;
        b.pos = S.state_pos;
        b.vel = S.state_vel;
        rc = launch_pass(kPassHash, P[r], b, dt, S.key_bits, false, s);
        if (rc == SPH_OK) rc = launch_pass(kPassSort, P[r], b, dt, S.key_bits, false, s);
        if (rc == SPH_OK) rc = launch_pass(kPassReorder, P[r], b, dt, S.key_bits, false, s);
        if (rc != SPH_OK) return rc;
        if (blended(t)) {
            hipLaunchKernelGGL(k_gather_blend, blocks_for(S.cap), dim3(kBlock), 0, s, S.buf.perm,
                               S.buf.n_dev, S.state_blend, S.tblend);
            rc = launch_gather_x(S.tblend, S.buf.i2r, S.buf.n_dev, S.cap, S.iblend, s);
            if (rc != SPH_OK) return rc;
        }
        uint32_t* occ = (r == 1) ? t->occ : t->occ + C0;
        hipLaunchKernelGGL(k_mark_occupancy, blocks_for(S.cap), dim3(kBlock), 0, s, S.buf.tpos,
                           S.buf.n_dev, G[0], t->stamp, occ, walk_counts(t, r));
        if (t->debug & SPH_DEBUG_COLOR) {
            rc = launch_energy_color(S.buf.tpos, S.buf.tvel, S.buf.cell_start, S.buf.cell_end,
                                     S.buf.n_dev, S.cap, G[r], S.cfg.mass, S.tcolor, s);
            if (rc != SPH_OK) return rc;
        }
        if (t->walls.n) {  // walls on the low grid
            rc = launch_wall_accel(t->walls, S.buf.tpos, S.buf.n_dev, S.cap, t->bacc[r], s);
            if (rc != SPH_OK) return rc;
        }
    }
    rc = wait_for(t, 1, sl, st);  // the low set's sorted state and marks
    if (rc != SPH_OK) return rc;
    hipLaunchKernelGGL(k_dilate_occupancy, dim3(std::min<uint32_t>((C0 + kBlock - 1) / kBlock, 2048u)),
                       dim3(kBlock), 0, st, t->occ, t->occ + C0, G[0].dx, G[0].dy, G[0].dz,
                       t->stamp, t->occ + 2 * C0);
    // The blended coupling's passes read both sets throughout: they run on
    // t->stream after the join above.
    if (blended(t)) return launch_blended_passes(t, P, G, X, dt, accel_out);
    // Own-set densities; the complementary densities (which also finalise A,
    // B): the particles with a complementary walk are listed and walked one
    // per wave; the own-set forces (no integration); the complementary forces
    // of the listed particles; the integration.
    rc = launch_density_own_x(P[0], t->set[0].buf, sl);
    if (rc == SPH_OK) rc = wait_for(t, 2, st, sl);  // the dilation and the high set's sorted state
    if (rc == SPH_OK) rc = launch_density_own_x(P[1], t->set[1].buf, st);
    if (rc != SPH_OK) return rc;
    for (int r = 1; r >= 0; --r) {
        TwoScaleSet& A = t->set[r];
        TwoScaleSet& B = t->set[1 - r];
        const hipStream_t s = sst[r];
        hipLaunchKernelGGL(k_cross_list, blocks_for(A.cap), dim3(kBlock), 0, s, A.buf.tpos,
                           A.buf.n_dev, B.buf.n_dev, nearby(t, r), X[0], cross_params(t, r),
                           A.buf.rho, A.buf.aux, kNoXacc, walk_counts(t, r), A.walk,
                           walk_segment_cap(A.cap));
        launch_density_walk<false>(walk_lanes(X[1 - r]), s, A.buf.tpos, walk_counts(t, r),
                                   A.walk, walk_segment_cap(A.cap), B.buf.tpos, kNoBlend,
                                   B.buf.cell_start, B.buf.cell_end, B.buf.n_dev, X[1 - r],
                                   cross_params(t, r), A.buf.rho, A.buf.aux);
        SPH_HIP(hipEventRecord(t->ev[4 - r], s));  // event 3: high A, B final; 4: low
        rc = launch_gather_aux(P[r], A.buf, s);
        if (rc == SPH_OK) rc = launch_force_own_x(P[r], G[r], A.buf, A.accel, s);
        if (rc != SPH_OK) return rc;
    }
    for (int r = 1; r >= 0; --r) {
        TwoScaleSet& A = t->set[r];
        TwoScaleSet& B = t->set[1 - r];
        const hipStream_t s = sst[r];
        SPH_HIP(hipStreamWaitEvent(s, t->ev[3 + r], 0));  // the other set's final A, B
        CrossParams c = cross_params(t, r);
        if (c.tension != 0.0f)
            launch_force_walk<true, false>(walk_lanes(X[1 - r]), s, A.buf.tpos, A.buf.tvel,
                                           A.buf.aux, walk_counts(t, r), A.walk,
                                           walk_segment_cap(A.cap), B.buf.tpos, B.buf.tvel,
                                           B.buf.aux, kNoBlend, B.buf.cell_start,
                                           B.buf.cell_end, B.buf.n_dev, X[1 - r], c, A.accel);
        else
            launch_force_walk<false, false>(walk_lanes(X[1 - r]), s, A.buf.tpos, A.buf.tvel,
                                           A.buf.aux, walk_counts(t, r), A.walk,
                                           walk_segment_cap(A.cap), B.buf.tpos, B.buf.tvel,
                                           B.buf.aux, kNoBlend, B.buf.cell_start,
                                           B.buf.cell_end, B.buf.n_dev, X[1 - r], c, A.accel);
        c.vmax = velocity_limit(t, r, dt);
        hipLaunchKernelGGL(k_cross_integrate, blocks_for(A.cap), dim3(kBlock), 0, s, A.buf.tpos,
                           A.buf.tvel, A.accel, A.buf.n_dev, c, wall_acc(t, r), dt, A.buf.pos,
                           A.buf.vel, accel_out ? A.buf.accel : nullptr);
    }
    // Each set's split/merge marks at the end of its own branch, then the join
    // and the commit.
    rc = launch_split_merge_marks(t, st, sl);
    if (rc == SPH_OK) rc = wait_for(t, 5, sl, st);  // join
    if (rc != SPH_OK) return rc;
    return launch_split_merge_commit(t);
}

// Both sets' renderer buffers from the compacted state: AoS positions, the
// state words (blended coupling: the blend state; fixed: 0 low, 4 high, the
// reference's initial states) and, with SPH_RENDER_COLORS, the colour values.
int export_render_buffers2(sph_two_scale* t) {
    if (!t->render_arena) return SPH_OK;
    const bool colors = (t->render_flags & SPH_RENDER_COLORS) && (t->debug & SPH_DEBUG_COLOR);
    for (int r = 0; r < 2; ++r) {
        TwoScaleSet& S = t->set[r];
        const int rc = launch_export_render(S.state_pos, blended(t) ? S.state_blend : nullptr,
                                            S.color, S.buf.n_dev, S.cap, r == 0 ? 0u : 4u,
                                            t->render_xyz[r],
                                            colors ? t->render_color[r] : nullptr,
                                            t->render_state[r], t->stream);
        if (rc != SPH_OK) return rc;
    }
    return SPH_OK;
}

// One step, then the renderer buffers when they are enabled.
int launch_full_two_scale_step(sph_two_scale* t, float dt, bool accel) {
    const int rc = launch_two_scale_step(t, dt, accel);
    return rc == SPH_OK ? export_render_buffers2(t) : rc;
}

int capture_two_scale(sph_two_scale* t, float dt, bool accel) {
    destroy_graph2(t);
    hipGraph_t g = nullptr;
    SPH_HIP(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_full_two_scale_step(t, dt, accel);
    hipError_t e = hipStreamEndCapture(t->stream, &g);
    if (rc != SPH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
    e = hipGraphInstantiate(&t->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        t->graph = nullptr;
        return hip_error(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    t->graph_dt = dt;
    t->graph_accel = accel;
    t->graph_debug = t->debug;
    return SPH_OK;
}

size_t al(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

// Times the launch variants of each set's own-set density and force passes
// (tile and queue, block size, particles per lane, then the grid shape of the
// best) on the current state and returns the fastest word per set.  It runs
// the neighbourhood passes first; the passes it times write only the step's
// intermediate buffers (densities, the internal-order A/B copy, the own-set
// acceleration), which the next step recomputes, so the state is unchanged.
// The force is timed with the A, B those buffers hold.
int tune_own_passes(sph_two_scale* t, uint32_t reps, uint32_t* tune_out, float* ms_out) {
    hipStream_t st = t->stream;
    reps = std::max(reps, 1u);
    StepParams P[2];
    GridView G[2], X[2];
    set_views(t, P, G, X);
    SPH_HIP(hipStreamSynchronize(st));
    for (int r = 1; r >= 0; --r) {
        TwoScaleSet& S = t->set[r];
        StepBuffers b = S.buf;
        b.pos = S.state_pos;
        b.vel = S.state_vel;
        int rc = launch_pass(kPassHash, P[r], b, 0.0f, S.key_bits, false, st);
        if (rc == SPH_OK) rc = launch_pass(kPassSort, P[r], b, 0.0f, S.key_bits, false, st);
        if (rc == SPH_OK) rc = launch_pass(kPassReorder, P[r], b, 0.0f, S.key_bits, false, st);
        if (rc != SPH_OK) return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto done = [&](int code) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        return code;
    };
    if (e != hipSuccess) return done(hip_error(e, "autotune events", __FILE__, __LINE__));
    for (int r = 0; r < 2; ++r) {
        TwoScaleSet& S = t->set[r];
        // Mean milliseconds of one launch with tune word w (after one untimed).
        auto time = [&](bool force, uint32_t w, float& ms) -> int {
            StepParams p = P[r];
            p.tune = w;
            auto run = [&]() {
                return force ? launch_force_own_x(p, G[r], S.buf, S.accel, st)
                             : launch_density_own_x(p, S.buf, st);
            };
            int rc = run();
            if (rc != SPH_OK) return rc;
            SPH_HIP(hipEventRecord(e0, st));
            for (uint32_t i = 0; i < reps && rc == SPH_OK; ++i) rc = run();
            if (rc != SPH_OK) return rc;
            SPH_HIP(hipEventRecord(e1, st));
            SPH_HIP(hipEventSynchronize(e1));
            SPH_HIP(hipEventElapsedTime(&ms, e0, e1));
            ms /= static_cast<float>(reps);
            return SPH_OK;
        };
        uint32_t word[2] = {0u, 0u};
        float def[2] = {0.f, 0.f}, best[2] = {0.f, 0.f};
        const bool tension = P[r].tens != 0.0f;
        for (int pass = 0; pass < 2; ++pass) {
            const bool force = pass == 1;
            if (force && tension) {  // the tension form runs the default
                int rc = time(true, 0u, def[1]);
                if (rc != SPH_OK) return done(rc);
                best[1] = def[1];
                break;
            }
            // tile/queue field, block-size field, particles-per-lane field, grid field
            const uint32_t nvar = force ? 9u : 5u;
            const int vshift = force ? 4 : 0, bshift = force ? 18 : 16, pshift = force ? 22 : 20,
                      gshift = force ? 12 : 8;
            float ms = 0.f;
            uint32_t shape = 0u, grid = 0u;
            int rc = time(force, 0u, def[pass]);
            best[pass] = def[pass];
            for (uint32_t v = 1; v <= nvar && rc == SPH_OK; ++v) {
                rc = time(force, v << vshift, ms);
                if (ms < best[pass]) { best[pass] = ms; shape = v << vshift; }
            }
            for (uint32_t bv = 1; bv <= 2 && rc == SPH_OK; ++bv) {
                rc = time(force, bv << bshift, ms);
                if (ms < best[pass]) { best[pass] = ms; shape = bv << bshift; }
            }
            if (rc == SPH_OK) {
                rc = time(force, 1u << pshift, ms);
                if (ms < best[pass]) { best[pass] = ms; shape = 1u << pshift; }
            }
            for (uint32_t gsh = 1; gsh <= 4 && rc == SPH_OK; ++gsh) {
                rc = time(force, shape | (gsh << gshift), ms);
                if (ms < best[pass]) { best[pass] = ms; grid = gsh; }
            }
            if (rc != SPH_OK) return done(rc);
            word[pass] = shape | (grid << gshift);
            if (!force) {  // densities and the A, B copy for the force timings
                StepParams p = P[r];
                p.tune = word[0];
                rc = launch_density_own_x(p, S.buf, st);
                if (rc == SPH_OK) rc = launch_gather_aux(p, S.buf, st);
                if (rc != SPH_OK) return done(rc);
            }
        }
        tune_out[r] = word[0] | word[1];
        if (ms_out) {
            ms_out[4 * r + 0] = def[0];
            ms_out[4 * r + 1] = best[0];
            ms_out[4 * r + 2] = def[1];
            ms_out[4 * r + 3] = best[1];
        }
    }
    return done(SPH_OK);
}

}  // namespace

extern "C" {

int sph2_set_tune(sph_two_scale* t, int resolution, uint32_t tune) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    if (resolution != 0 && resolution != 1)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "resolution must be 0 (low) or 1 (high)");
    if (!tune_word_valid(tune, false)) return set_error(SPH_ERR_INVALID_ARGUMENT, "invalid tune word");
    if (t->tune[resolution] == tune) return SPH_OK;
    SPH_HIP(hipStreamSynchronize(t->stream));
    t->tune[resolution] = tune;
    destroy_graph2(t);
    return SPH_OK;
}

int sph2_get_tune(const sph_two_scale* t, int resolution, uint32_t* tune) {
    if (!t || !tune) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (resolution != 0 && resolution != 1)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "resolution must be 0 (low) or 1 (high)");
    *tune = t->tune[resolution];
    return SPH_OK;
}

int sph2_autotune(sph_two_scale* t, uint32_t reps, uint32_t tune[2], float* ms) {
    if (!t || !tune) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t words[2] = {0u, 0u};
    const int rc = tune_own_passes(t, reps, words, ms);
    if (rc != SPH_OK) return rc;
    for (int r = 0; r < 2; ++r) {
        const int rc2 = sph2_set_tune(t, r, words[r]);
        if (rc2 != SPH_OK) return rc2;
        tune[r] = words[r];
    }
    SPH_HIP(hipStreamSynchronize(t->stream));
    return SPH_OK;
}

int sph2_create(const sph_config* cfg, const float region_lo[3], const float region_hi[3],
                const float* pos, const float* vel, uint32_t n_low, uint32_t max_low,
                uint32_t max_high, sph_two_scale** out) {
    if (!out || !cfg || !region_lo || !region_hi)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (n_low > 0 && !pos) return set_error(SPH_ERR_INVALID_ARGUMENT, "positions are null");
    if (max_low < n_low) max_low = n_low;
    if (max_low == 0 || max_high == 0)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "capacities must be > 0");
    if (n_low >= (1u << 29) || max_low >= (1u << 29))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "two-scale ids need n < 2^29");
    if (!(cfg->h > 0.0f) || !(cfg->mass > 0.0f) || !(cfg->rest_density > 0.0f))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "config: h, mass, rest_density must be > 0");
    if (cfg->flags & SPH_FLAG_XSPH)
        return set_error(SPH_ERR_UNSUPPORTED, "two-scale: the XSPH integrator is single-scale only");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return set_error(SPH_ERR_NO_DEVICE, "no HIP device visible");
    }
    auto* t = new (std::nothrow) sph_two_scale;
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "out of host memory");
    t->cfg = *cfg;
    // Solver.h:36-43: h1 = h0/2, m1 = m0/8, one grid per resolution over the domain.
    t->set[0].cfg = *cfg;
    t->set[1].cfg = *cfg;
    t->set[1].cfg.h = 0.5f * cfg->h;
    t->set[1].cfg.mass = cfg->mass / 8.0f;
    int rc = sph_make_grid(cfg->grid.origin, cfg->domain_hi, t->set[1].cfg.h, &t->set[1].cfg.grid);
    if (rc != SPH_OK) {
        delete t;
        return rc;
    }
    t->set[0].cap = max_low;
    t->set[1].cap = max_high;
    t->num_groups = max_low;  // low-resolution ids stay below max_low
    t->id_base = max_low;
    for (int d = 0; d < 3; ++d) {
        (&t->region.lox)[d] = region_lo[d];
        (&t->region.hix)[d] = region_hi[d];
    }
    t->region.margin = 0.5f * cfg->h;
    t->region.child_d = 0.25f * t->set[1].cfg.h;

    size_t total = 0;
    std::vector<size_t> sz;
    auto add = [&](size_t b) {
        sz.push_back(al(b));
        total += al(b);
    };
    for (int r = 0; r < 2; ++r) {
        TwoScaleSet& S = t->set[r];
        S.num_cells = sph_grid_num_cells(&S.cfg.grid);
        S.key_bits = key_bits_for_cells(S.num_cells);
        const size_t N = S.cap, C = S.num_cells, nb = blocks_for(S.cap).x;
        for (int k = 0; k < 4; ++k) add(16 * N);  // pos vel tpos tvel
        add(4 * N);                                // rho
        add(16 * N);                               // accel (debug)
        for (int k = 0; k < 3; ++k) add(4 * N);    // keys skeys perm
        add(4 * C);
        add(4 * C);
        add(radix_sort_temp_bytes(S.cap));
        add(8 * N);                                // aux
        add(4);                                    // n_dev
        add(16 * N);                               // state_pos
        add(16 * N);                               // state_vel
        add(16 * N);                               // own-set accel
        add(4 * N);                                // flags
        add(8 * nb);                               // block counts
        for (int k = 0; k < 3; ++k) add(8 * N);    // state_blend tblend nblend
        add(4 * N);                                // tcolor
        add(4 * N);                                // color
        add(4ull * kWalkSegments * walk_segment_cap(S.cap));  // walk list segments
        add(16 * N);                               // ipos
        add(16 * N);                               // ivel
        add(8 * N);                                // iaux
        add(4 * N);                                // i2r
        add(kXBinBytes * C);                       // xbins
        add(8 * N);                                // iblend
    }
    add(128 * static_cast<size_t>(t->num_groups));  // grp_pos: 8 float4 per group
    add(128 * static_cast<size_t>(t->num_groups));  // grp_vel
    add(4 * static_cast<size_t>(t->num_groups));    // grp_cnt
    add(4 * kTotWords);                             // tot: totals, error word, stamp, ticket
    add(12 * static_cast<size_t>(t->set[0].num_cells));  // low-grid occupancy: occ_h, occ_l, near
    add(4 * static_cast<size_t>(t->num_groups));    // grp_live
    add(4 * static_cast<size_t>(t->num_groups));    // surf
    add(4ull * 2 * kWalkSegments * kWalkCounterStride);  // walk list counters

    hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&t->stream2, hipStreamNonBlocking);
    for (hipEvent_t& ev : t->ev)
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        destroy_streams2(t);
        delete t;
        return hip_error(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    e = hipMalloc(&t->arena, total);
    if (e != hipSuccess) {
        destroy_streams2(t);
        delete t;
        return hip_error(e, "hipMalloc", __FILE__, __LINE__);
    }
    char* a = static_cast<char*>(t->arena);
    size_t idx = 0;
    auto take = [&]() {
        void* p = a;
        a += sz[idx++];
        return p;
    };
    for (int r = 0; r < 2; ++r) {
        TwoScaleSet& S = t->set[r];
        S.buf.pos = static_cast<float4*>(take());
        S.buf.vel = static_cast<float4*>(take());
        S.buf.tpos = static_cast<float4*>(take());
        S.buf.tvel = static_cast<float4*>(take());
        S.buf.rho = static_cast<float*>(take());
        S.buf.accel = static_cast<float4*>(take());
        S.buf.keys = static_cast<uint32_t*>(take());
        S.buf.skeys = static_cast<uint32_t*>(take());
        S.buf.perm = static_cast<uint32_t*>(take());
        S.buf.cell_start = static_cast<uint32_t*>(take());
        S.buf.cell_end = static_cast<uint32_t*>(take());
        S.buf.sort_temp = take();
        S.buf.sort_temp_bytes = radix_sort_temp_bytes(S.cap);
        S.buf.aux = static_cast<float2*>(take());
        S.buf.n_dev = static_cast<uint32_t*>(take());
        S.state_pos = static_cast<float4*>(take());
        S.state_vel = static_cast<float4*>(take());
        S.accel = static_cast<float4*>(take());
        S.flag = static_cast<uint32_t*>(take());
        S.counts = static_cast<uint32_t*>(take());
        S.state_blend = static_cast<float2*>(take());
        S.tblend = static_cast<float2*>(take());
        S.nblend = static_cast<float2*>(take());
        S.tcolor = static_cast<float*>(take());
        S.color = static_cast<float*>(take());
        S.walk = static_cast<uint32_t*>(take());
        S.buf.ipos = static_cast<float4*>(take());
        S.buf.ivel = static_cast<float4*>(take());
        S.buf.iaux = static_cast<float2*>(take());
        S.buf.i2r = static_cast<uint32_t*>(take());
        S.buf.xbins = static_cast<uint4*>(take());
        S.iblend = static_cast<float2*>(take());
        S.buf.keep_sorted = true;
    }
    t->grp_pos = static_cast<float4*>(take());
    t->grp_vel = static_cast<float4*>(take());
    t->grp_cnt = static_cast<uint32_t*>(take());
    t->tot = static_cast<uint32_t*>(take());
    t->error_word = t->tot + kErrorWordIndex;
    t->stamp = t->tot + kStampWord;
    t->occ = static_cast<uint32_t*>(take());
    t->grp_live = static_cast<uint32_t*>(take());
    t->surf = static_cast<uint32_t*>(take());
    t->walk_count = static_cast<uint32_t*>(take());

    e = hipMemsetAsync(t->arena, 0, total, t->stream);
    if (e != hipSuccess) {
        sph2_destroy(t);
        return hip_error(e, "hipMemsetAsync", __FILE__, __LINE__);
    }
    rc = sph2_set_state(t, pos, vel, nullptr, n_low, nullptr, nullptr, nullptr, nullptr, 0);
    if (rc != SPH_OK) {
        sph2_destroy(t);
        return rc;
    }
    *out = t;
    return SPH_OK;
}

namespace {

// Every particle of set r settled with lambda = 1 (state 0 low, 4 high).
int reset_blend(sph_two_scale* t, int r) {
    TwoScaleSet& S = t->set[r];
    const uint32_t state = (r == 0) ? 0u : 4u;
    float2 w;
    w.x = 1.0f;
    std::memcpy(&w.y, &state, 4);
    hipLaunchKernelGGL(k_fill_blend, blocks_for(S.cap), dim3(kBlock), 0, t->stream, S.buf.n_dev, w,
                       S.state_blend);
    SPH_HIP(hipGetLastError());
    return SPH_OK;
}

// Uploads one resolution's state (AoS host arrays) into its compacted-state
// buffers, staging through the sorted-copy buffers.
int upload_set(sph_two_scale* t, int r, const float* pos, const float* vel, const uint32_t* ids,
               const uint32_t* tags, uint32_t n) {
    TwoScaleSet& S = t->set[r];
    if (n > S.cap) return set_error(SPH_ERR_CAPACITY, "two-scale: n exceeds capacity");
    if (n > 0 && !pos) return set_error(SPH_ERR_INVALID_ARGUMENT, "positions are null");
    SPH_HIP(hipMemcpyAsync(S.buf.n_dev, &n, sizeof(uint32_t), hipMemcpyHostToDevice, t->stream));
    if (n == 0) return SPH_OK;
    float* dp = reinterpret_cast<float*>(S.buf.tpos);
    float* dv = reinterpret_cast<float*>(S.buf.tvel);
    uint32_t* di = S.buf.keys;
    uint32_t* dt = S.buf.skeys;
    SPH_HIP(hipMemcpyAsync(dp, pos, 12ull * n, hipMemcpyHostToDevice, t->stream));
    if (vel) SPH_HIP(hipMemcpyAsync(dv, vel, 12ull * n, hipMemcpyHostToDevice, t->stream));
    if (ids) SPH_HIP(hipMemcpyAsync(di, ids, 4ull * n, hipMemcpyHostToDevice, t->stream));
    if (tags) SPH_HIP(hipMemcpyAsync(dt, tags, 4ull * n, hipMemcpyHostToDevice, t->stream));
    hipLaunchKernelGGL(k_pack_two_scale, blocks_for(n), dim3(kBlock), 0, t->stream, dp,
                       vel ? dv : nullptr, ids ? di : nullptr, tags ? dt : nullptr, n, S.state_pos,
                       S.state_vel);
    SPH_HIP(hipGetLastError());
    return reset_blend(t, r);
}

}  // namespace

int sph2_set_state(sph_two_scale* t, const float* pos_l, const float* vel_l,
                   const uint32_t* ids_l, uint32_t n_l, const float* pos_h, const float* vel_h,
                   const uint32_t* ids_h, const uint32_t* tags_h, uint32_t n_h) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    if (n_h > 0 && (!ids_h || !tags_h))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "high-resolution particles need ids and tags");
    int rc = upload_set(t, 0, pos_l, vel_l, ids_l, nullptr, n_l);
    if (rc == SPH_OK) rc = upload_set(t, 1, pos_h, vel_h, ids_h, tags_h, n_h);
    if (rc == SPH_OK) rc = export_render_buffers2(t);
    if (rc != SPH_OK) return rc;
    SPH_HIP(hipStreamSynchronize(t->stream));
    destroy_graph2(t);
    return SPH_OK;
}

int sph2_advance(sph_two_scale* t, float dt, uint32_t n_steps) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    if (!(dt > 0.0f) || !std::isfinite(dt)) return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (n_steps == 0) return SPH_OK;
    const bool accel = (t->debug & SPH_DEBUG_ACCEL) != 0u;
    if (!t->graph || t->graph_dt != dt || t->graph_accel != accel || t->graph_debug != t->debug) {
        const int rc = capture_two_scale(t, dt, accel);
        if (rc != SPH_OK) return rc;
    }
    for (uint32_t i = 0; i < n_steps; ++i) SPH_HIP(hipGraphLaunch(t->graph, t->stream));
    return SPH_OK;
}

int sph2_advance_eager(sph_two_scale* t, float dt, uint32_t n_steps) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    if (!(dt > 0.0f)) return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    const bool accel = (t->debug & SPH_DEBUG_ACCEL) != 0u;
    for (uint32_t i = 0; i < n_steps; ++i) {
        const int rc = launch_full_two_scale_step(t, dt, accel);
        if (rc != SPH_OK) return rc;
    }
    return SPH_OK;
}

int sph2_apply_split_merge(sph_two_scale* t) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    for (int r = 0; r < 2; ++r) {
        TwoScaleSet& S = t->set[r];
        hipLaunchKernelGGL(k_copy_state, blocks_for(S.cap), dim3(kBlock), 0, t->stream, S.state_pos,
                           S.state_vel, blended(t) ? S.state_blend : nullptr, S.buf.n_dev,
                           S.buf.pos, S.buf.vel, S.nblend);
    }
    // No step ran on this state: reference split offsets use the rest density.
    hipLaunchKernelGGL(k_fill_float, blocks_for(t->set[0].cap), dim3(kBlock), 0, t->stream,
                       t->set[0].buf.n_dev, t->set[0].cfg.rest_density, t->set[0].buf.rho);
    SPH_HIP(hipGetLastError());
    const int rc = launch_split_merge(t);
    return rc == SPH_OK ? export_render_buffers2(t) : rc;
}

int sph2_set_options(sph_two_scale* t, const sph2_options* o) {
    if (!t || !o) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (o->coupling > SPH2_COUPLING_BLENDED || o->split_offsets > SPH2_SPLIT_REFERENCE ||
        o->merge > 1u)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "options: unknown coupling/split/merge value");
    if (!(o->blend_increment > 0.0f && o->blend_increment <= 1.0f))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "options: blend_increment must be in (0, 1]");
    if (!(o->velocity_cap >= 0.0f) || !std::isfinite(o->velocity_cap))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "options: velocity_cap must be >= 0");
    const bool was = blended(t);
    t->opt = *o;
    if (blended(t) && !was) {
        for (int r = 0; r < 2; ++r) {
            const int rc = reset_blend(t, r);
            if (rc != SPH_OK) return rc;
        }
        SPH_HIP(hipMemsetAsync(t->grp_live, 0, 4ull * t->num_groups, t->stream));
    }
    SPH_HIP(hipStreamSynchronize(t->stream));
    destroy_graph2(t);
    return SPH_OK;
}

int sph2_get_options(const sph_two_scale* t, sph2_options* o) {
    if (!t || !o) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    *o = t->opt;
    return SPH_OK;
}

int sph2_set_blend(sph_two_scale* t, int resolution, const float* lambda, const uint32_t* state,
                   uint32_t n) {
    if (!t || resolution < 0 || resolution > 1)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "bad argument");
    if (!blended(t)) return set_error(SPH_ERR_INVALID_ARGUMENT, "blend words need the blended coupling");
    uint32_t c[4];
    int rc = sph2_counts(t, c);
    if (rc != SPH_OK) return rc;
    if (n != c[resolution]) return set_error(SPH_ERR_INVALID_ARGUMENT, "n must equal the set's count");
    if (n == 0) return SPH_OK;
    TwoScaleSet& S = t->set[resolution];
    std::vector<float2> w(n);
    SPH_HIP(hipMemcpyAsync(w.data(), S.state_blend, 8ull * n, hipMemcpyDeviceToHost, t->stream));
    SPH_HIP(hipStreamSynchronize(t->stream));
    for (uint32_t i = 0; i < n; ++i) {
        if (lambda) w[i].x = lambda[i];
        if (state) std::memcpy(&w[i].y, &state[i], 4);
    }
    SPH_HIP(hipMemcpyAsync(S.state_blend, w.data(), 8ull * n, hipMemcpyHostToDevice, t->stream));
    SPH_HIP(hipStreamSynchronize(t->stream));
    return SPH_OK;
}

int sph2_mark_split(sph_two_scale* t, const uint32_t* ids, uint32_t n) {
    if (!t || (n > 0 && !ids)) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (!blended(t)) return set_error(SPH_ERR_INVALID_ARGUMENT, "split marks need the blended coupling");
    TwoScaleSet& L = t->set[0];
    if (n > L.cap) return set_error(SPH_ERR_INVALID_ARGUMENT, "more ids than low-resolution capacity");
    if (n == 0) return SPH_OK;
    std::vector<uint32_t> sorted(ids, ids + n);
    std::sort(sorted.begin(), sorted.end());
    // The low set's hash-key buffer is free between steps: stage the ids there.
    SPH_HIP(hipMemcpyAsync(L.buf.keys, sorted.data(), 4ull * n, hipMemcpyHostToDevice, t->stream));
    hipLaunchKernelGGL(k_mark_ids, blocks_for(L.cap), dim3(kBlock), 0, t->stream, L.state_pos,
                       L.buf.n_dev, L.buf.keys, n, L.state_blend);
    SPH_HIP(hipGetLastError());
    SPH_HIP(hipStreamSynchronize(t->stream));
    return SPH_OK;
}

int sph2_export_render(sph_two_scale* t, int resolution, float* xyz, float* color,
                       int32_t* state) {
    if (!t || resolution < 0 || resolution > 1)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "bad argument");
    if (color && !(t->debug & SPH_DEBUG_COLOR))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "colour values need SPH_DEBUG_COLOR");
    TwoScaleSet& S = t->set[resolution];
    return launch_export_render(S.state_pos, blended(t) ? S.state_blend : nullptr, S.color,
                                S.buf.n_dev, S.cap, resolution == 0 ? 0u : 4u, xyz, color, state,
                                t->stream);
}

int sph2_detect_surface(sph_two_scale* t, float threshold, int split) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    if (!blended(t)) return set_error(SPH_ERR_INVALID_ARGUMENT, "surface marks need the blended coupling");
    if (!(threshold >= 0.0f) || !std::isfinite(threshold))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "threshold must be >= 0");
    uint32_t c[4];
    int rc = sph2_counts(t, c);
    if (rc != SPH_OK) return rc;
    if (c[2] == 0) return set_error(SPH_ERR_INVALID_ARGUMENT, "no step has run on the low set");
    TwoScaleSet& L = t->set[0];
    TwoScaleSet& H = t->set[1];
    const StepParams pl = make_step_params(L.cfg, L.cap), ph = make_step_params(H.cfg, H.cap);
    const GridView gl = make_grid_view(pl.ox, pl.oy, pl.oz, pl.spacing, pl.dx, pl.dy, pl.dz, pl.h);
    const GridView gh = make_grid_view(ph.ox, ph.oy, ph.oz, ph.spacing, ph.dx, ph.dy, ph.dz, pl.h);
    const CrossParams cp = cross_params(t, 0);
    SurfaceParams sp{};
    sp.h0 = cp.h0;
    sp.h1 = cp.h1;
    sp.c45_0 = cp.c45_0;
    sp.c45_1 = cp.c45_1;
    sp.m0 = L.cfg.mass;
    sp.m1 = H.cfg.mass;
    sp.threshold = threshold;
    // The last step's sorted state: the neighbour passes ran on c[2] / c[3]
    // particles, whose counts the step's n_dev no longer holds.
    uint32_t* n_step = t->tot + 6;  // tot[6], tot[7]
    rc = launch_surface_detect(L.buf.tpos, L.buf.aux, L.buf.cell_start, L.buf.cell_end, n_step,
                               L.cap, gl, H.buf.tpos, H.buf.aux, H.buf.cell_start, H.buf.cell_end,
                               n_step + 1, gh, sp, t->surf, t->num_groups, t->stream);
    if (rc == SPH_OK)
        rc = launch_apply_surface(L.state_pos, L.buf.n_dev, L.cap, t->surf, t->num_groups,
                                  split ? 1u : 0u, L.state_blend, t->stream);
    if (rc != SPH_OK) return rc;
    SPH_HIP(hipStreamSynchronize(t->stream));
    return SPH_OK;
}

int sph2_mark_surface_map(sph_two_scale* t, const uint32_t* map, uint32_t count, int split) {
    if (!t || (count > 0 && !map)) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (!blended(t)) return set_error(SPH_ERR_INVALID_ARGUMENT, "surface marks need the blended coupling");
    TwoScaleSet& L = t->set[0];
    const int rc = launch_mark_surface_map(map, count, L.buf.n_dev, split ? 1u : 0u, L.state_blend,
                                           t->stream);
    if (rc != SPH_OK) return rc;
    SPH_HIP(hipStreamSynchronize(t->stream));
    return SPH_OK;
}

int sph2_synchronize(sph_two_scale* t) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    uint32_t v[3] = {};
    int rc = read_error_word(t->error_word, t->stream, &v[0]);
    for (int r = 0; r < 2 && rc == SPH_OK; ++r)
        rc = read_error_word(radix_sort_error_word(t->set[r].buf.sort_temp), t->stream, &v[1 + r]);
    if (rc != SPH_OK) return rc;
    return device_error_status(public_error_bits(v[1] | v[2]) |
                               ((v[0] & 2u) ? SPH_DEVERR_CAPACITY : 0u));
}

int sph2_set_debug(sph_two_scale* t, uint32_t flags) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    // Renderer colour values need the colour pass.
    if (t->render_flags & SPH_RENDER_COLORS) flags |= SPH_DEBUG_COLOR;
    t->debug = flags;
    return SPH_OK;
}

int sph2_enable_render_buffers(sph_two_scale* t, uint32_t flags) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    if (flags & ~SPH_RENDER_COLORS) return set_error(SPH_ERR_INVALID_ARGUMENT, "unknown render flag");
    SPH_HIP(hipStreamSynchronize(t->stream));
    if (!t->render_arena) {
        size_t total = 0;
        for (int r = 0; r < 2; ++r) total += al(12ull * t->set[r].cap) + 2 * al(4ull * t->set[r].cap);
        SPH_HIP(hipMalloc(&t->render_arena, total));
        SPH_HIP(hipMemsetAsync(t->render_arena, 0, total, t->stream));
        char* a = static_cast<char*>(t->render_arena);
        for (int r = 0; r < 2; ++r) {
            const size_t n = t->set[r].cap;
            t->render_xyz[r] = reinterpret_cast<float*>(a);
            a += al(12 * n);
            t->render_color[r] = reinterpret_cast<float*>(a);
            a += al(4 * n);
            t->render_state[r] = reinterpret_cast<int32_t*>(a);
            a += al(4 * n);
        }
    }
    t->render_flags = flags;
    if (flags & SPH_RENDER_COLORS) t->debug |= SPH_DEBUG_COLOR;
    destroy_graph2(t);
    const int rc = export_render_buffers2(t);
    if (rc != SPH_OK) return rc;
    SPH_HIP(hipStreamSynchronize(t->stream));
    return SPH_OK;
}

int sph2_get_render_buffers(sph_two_scale* t, int resolution, sph_render_buffers* out) {
    if (!t || !out || resolution < 0 || resolution > 1)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "bad argument");
    if (!t->render_arena)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "render buffers are not enabled");
    uint32_t c[4];
    const int rc = sph2_counts(t, c);
    if (rc != SPH_OK) return rc;
    out->positions = t->render_xyz[resolution];
    out->color_values = t->render_color[resolution];
    out->states = t->render_state[resolution];
    out->n = c[resolution];
    out->max_n = t->set[resolution].cap;
    return SPH_OK;
}

int sph2_counts(sph_two_scale* t, uint32_t counts[4]) {
    if (!t || !counts) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t tot[8] = {};
    SPH_HIP(hipMemcpyAsync(&counts[0], t->set[0].buf.n_dev, 4, hipMemcpyDeviceToHost, t->stream));
    SPH_HIP(hipMemcpyAsync(&counts[1], t->set[1].buf.n_dev, 4, hipMemcpyDeviceToHost, t->stream));
    SPH_HIP(hipMemcpyAsync(tot, t->tot, sizeof(tot), hipMemcpyDeviceToHost, t->stream));
    SPH_HIP(hipStreamSynchronize(t->stream));
    counts[2] = tot[6];
    counts[3] = tot[7];
    return SPH_OK;
}

int sph2_device_errors(sph_two_scale* t, uint32_t* bits) {
    if (!t || !bits) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t v[3] = {};
    SPH_HIP(hipMemcpyAsync(&v[0], t->error_word, 4, hipMemcpyDeviceToHost, t->stream));
    SPH_HIP(hipMemcpyAsync(&v[1], radix_sort_error_word(t->set[0].buf.sort_temp), 4,
                           hipMemcpyDeviceToHost, t->stream));
    SPH_HIP(hipMemcpyAsync(&v[2], radix_sort_error_word(t->set[1].buf.sort_temp), 4,
                           hipMemcpyDeviceToHost, t->stream));
    SPH_HIP(hipStreamSynchronize(t->stream));
    *bits = public_error_bits(v[1] | v[2]) | ((v[0] & 2u) ? SPH_DEVERR_CAPACITY : 0u);
    if (*bits) {
        SPH_HIP(hipMemsetAsync(t->error_word, 0, 4, t->stream));
        SPH_HIP(hipMemsetAsync(radix_sort_error_word(t->set[0].buf.sort_temp), 0, 4, t->stream));
        SPH_HIP(hipMemsetAsync(radix_sort_error_word(t->set[1].buf.sort_temp), 0, 4, t->stream));
        SPH_HIP(hipStreamSynchronize(t->stream));
    }
    return SPH_OK;
}

int sph2_download(sph_two_scale* t, int resolution, int field, void* dst, size_t bytes) {
    if (!t || !dst || resolution < 0 || resolution > 1)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "bad argument");
    uint32_t c[4];
    int rc = sph2_counts(t, c);
    if (rc != SPH_OK) return rc;
    TwoScaleSet& S = t->set[resolution];
    const size_t n_now = c[resolution], n_step = c[2 + resolution], C = S.num_cells;
    const float4* vec = nullptr;
    const void* src = nullptr;
    size_t n = 0, need = 0;
    int mode = 0;  // 0 raw copy, 1 xyz of float4, 2 w bits of float4, 3/4 lambda/state
    switch (field) {
        case SPH_FIELD_POSITION: vec = S.state_pos; n = n_now; need = 12 * n; mode = 1; break;
        case SPH_FIELD_VELOCITY: vec = S.state_vel; n = n_now; need = 12 * n; mode = 1; break;
        case SPH_FIELD_ID: vec = S.state_pos; n = n_now; need = 4 * n; mode = 2; break;
        case SPH_FIELD_TAG: vec = S.state_vel; n = n_now; need = 4 * n; mode = 2; break;
        case SPH_FIELD_COLOR:
            if (!(t->debug & SPH_DEBUG_COLOR))
                return set_error(SPH_ERR_INVALID_ARGUMENT, "colour values need SPH_DEBUG_COLOR");
            src = S.color; n = n_now; need = 4 * n; break;
        case SPH_FIELD_BLEND:
        case SPH_FIELD_STATE:
            if (!blended(t))
                return set_error(SPH_ERR_INVALID_ARGUMENT, "blend words need the blended coupling");
            n = n_now; need = 4 * n; mode = (field == SPH_FIELD_BLEND) ? 3 : 4; break;
        case SPH_FIELD_DENSITY: src = S.buf.rho; n = n_step; need = 4 * n; break;
        case SPH_FIELD_SORTED_HASH: src = S.buf.skeys; n = n_step; need = 4 * n; break;
        case SPH_FIELD_SORTED_INDEX: src = S.buf.perm; n = n_step; need = 4 * n; break;
        case SPH_FIELD_CELL_START: src = S.buf.cell_start; need = 4 * C; break;
        case SPH_FIELD_CELL_END: src = S.buf.cell_end; need = 4 * C; break;
        case SPH_FIELD_ACCEL:
            if (!(t->debug & SPH_DEBUG_ACCEL))
                return set_error(SPH_ERR_INVALID_ARGUMENT, "accelerations need SPH_DEBUG_ACCEL");
            vec = S.buf.accel; n = n_step; need = 12 * n; mode = 1; break;
        default: return set_error(SPH_ERR_INVALID_ARGUMENT, "unknown field");
    }
    if (bytes < need) return set_error(SPH_ERR_INVALID_ARGUMENT, "destination too small");
    if (need == 0) return SPH_OK;
    if (mode == 0) {
        SPH_HIP(hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, t->stream));
        SPH_HIP(hipStreamSynchronize(t->stream));
        return SPH_OK;
    }
    if (mode >= 3) {  // one lane of the blend words
        std::vector<float> w(2 * n);
        SPH_HIP(hipMemcpyAsync(w.data(), S.state_blend, 8 * n, hipMemcpyDeviceToHost, t->stream));
        SPH_HIP(hipStreamSynchronize(t->stream));
        auto* o = static_cast<float*>(dst);
        for (size_t i = 0; i < n; ++i) std::memcpy(&o[i], &w[2 * i + (mode == 3 ? 0 : 1)], 4);
        return SPH_OK;
    }
    std::vector<float> tmp(4 * n);
    SPH_HIP(hipMemcpyAsync(tmp.data(), vec, 16 * n, hipMemcpyDeviceToHost, t->stream));
    SPH_HIP(hipStreamSynchronize(t->stream));
    if (mode == 2) {
        auto* o = static_cast<uint32_t*>(dst);
        for (size_t i = 0; i < n; ++i) std::memcpy(&o[i], &tmp[4 * i + 3], 4);
    } else {
        auto* o = static_cast<float*>(dst);
        for (size_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) o[3 * i + d] = tmp[4 * i + d];
    }
    return SPH_OK;
}

int sph2_set_boundary(sph_two_scale* t, const float* pos, uint32_t n) {
    if (!t) return set_error(SPH_ERR_INVALID_ARGUMENT, "handle is null");
    SPH_HIP(hipStreamSynchronize(t->stream));
    destroy_graph2(t);
    const sph_config& lo = t->set[0].cfg;
    const int rc = walls_create(&t->walls, lo.grid, lo.h, lo.boundary_mass, lo.mass, lo.speed_sound,
                                pos, n, t->stream);
    if (rc != SPH_OK || n == 0) return rc;
    for (int r = 0; r < 2; ++r)
        if (!t->bacc[r]) SPH_HIP(hipMalloc(&t->bacc[r], 16ull * t->set[r].cap));
    return SPH_OK;
}

int sph2_destroy(sph_two_scale* t) {
    if (!t) return SPH_OK;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    destroy_graph2(t);
    walls_destroy(&t->walls);
    for (int r = 0; r < 2; ++r)
        if (t->bacc[r]) (void)hipFree(t->bacc[r]);
    if (t->render_arena) (void)hipFree(t->render_arena);
    if (t->arena) (void)hipFree(t->arena);
    destroy_streams2(t);
    delete t;
    return SPH_OK;
}

}  // extern "C"
