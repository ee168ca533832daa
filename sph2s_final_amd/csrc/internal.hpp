// internal.hpp -- declarations shared by the translation units of libsph2s.so.
// Not installed; the public surface is include/sph2s.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/sph2s.h"

namespace sph {

constexpr uint32_t kEmptyCell = SPH_EMPTY_CELL;

// Records msg as the calling thread's last error and returns code.
int set_error(int code, const std::string& msg);
int hip_error(hipError_t e, const char* what, const char* file, int line);

#define SPH_HIP(call)                                                         \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) return ::sph::hip_error(e_, #call, __FILE__, __LINE__); \
    } while (0)

// Everything the step kernels need, passed by value as a kernel argument (the
// reference uploads an equivalent struct to __constant__ memory,
// SPH2S_FINAL_SINGLE_SCALE/Solver.cu:12,1068-1080).  Kernel coefficients are
// folded in double on the host and rounded to float once
// (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:103-167 computes them with the double
// literal M_PI).
struct StepParams {
    float ox, oy, oz;       // grid origin
    int   dx, dy, dz;       // grid dims
    float spacing;          // cell edge
    float h, h2;            // effective radius and its square
    float rest_density, stiffness;
    float c6m;              // mass * 315/(64 pi h^9): rho = c6m * sum (h^2-r^2)^3
    float c6m_h6;           // mass * 315/(64 pi h^3): rho = c6m_h6 * sum (1 - r^2/h^2)^3
    float c45m;             // mass * 45/(pi h^6): pressure acceleration factor
    float visc;             // mu * mass * 45/(pi h^6)
    float tens;             // sigma * mass * 315/(64 pi h^9)
    float xsph;             // epsilon * mass * 315/(64 pi h^9) (XSPH integrator)
    float mass;
    float gx, gy, gz;
    float restitution;
    float lox, loy, loz, hix, hiy, hiz;
    uint32_t flags;         // SPH_FLAG_*
    uint32_t n;             // launch capacity (particles); the live count is in device memory
    uint32_t num_cells;
    uint32_t tune;          // launch variant of the neighbour passes (0 = defaults; step.hip)
};

StepParams make_step_params(const sph_config& cfg, uint32_t n);
uint32_t key_bits_for_cells(uint32_t num_cells);

// computeCoordinates / computeHash (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:20-83)
// in FP32: clamp((int)((x + off - o) / spacing), 0, D-1) and
// cx + Dx (cy + Dy cz).  The library is built with
// -fhip-fp32-correctly-rounded-divide-sqrt, so the division is IEEE and hashes
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

__device__ __forceinline__ uint32_t cell_hash(float x, float y, float z, float ox, float oy,
                                              float oz, float sp, int dx, int dy, int dz) {
    const int cx = cell_coord(x, ox, sp, dx);
    const int cy = cell_coord(y, oy, sp, dy);
    const int cz = cell_coord(z, oz, sp, dz);
    return static_cast<uint32_t>(cx + dx * (cy + dy * cz));
}

// ---- radix sort (radix_sort.hip) ----
size_t radix_sort_temp_bytes(uint32_t n);
// Device word set to non-zero if a look-back wait of the sort gave up.
uint32_t* radix_sort_error_word(void* temp);
// radix_sort_pairs with the library's static scratch array as the temporary
// buffer (no allocation); n <= kStaticSortMaxKeys.
constexpr uint32_t kStaticSortMaxKeys = 1u << 21;
int radix_sort_pairs_static(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                            uint32_t* vals_out, uint32_t n, uint32_t key_bits,
                            hipStream_t stream);
// Bits of that word: the sort's look-back gave up (set by the sort), and the
// force pass's integration wrote a NaN or infinite position or velocity.
constexpr uint32_t kDevErrSortWait = 1u;
constexpr uint32_t kDevErrNonFinite = 2u;
// The SPH_DEVERR_* bits of an error word.
inline uint32_t public_error_bits(uint32_t word) {
    return ((word & kDevErrSortWait) ? SPH_DEVERR_SORT_WAIT : 0u) |
           ((word & kDevErrNonFinite) ? SPH_DEVERR_NONFINITE : 0u);
}
// The status a synchronize reports for SPH_DEVERR_* bits (the bits stay set
// until the handle's device-errors query clears them).
int device_error_status(uint32_t public_bits);
// Reads an error word after the stream's work (synchronises the stream).
int read_error_word(const uint32_t* word, hipStream_t stream, uint32_t* out);
// Stable LSD sort of (key, value) pairs by the low key_bits of key.
// vals_in == nullptr means values 0..n-1.  Launches on stream.
int radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in,
                     uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                     uint32_t key_bits, void* temp, size_t temp_bytes,
                     hipStream_t stream);

// Zeroes the sort state, resets cell_start/cell_end to SPH_EMPTY_CELL and
// hashes pos into keys while building the digit histograms.
// p.n is the launch capacity; the live count is *n_dev (device memory).
int hash_hist_cells(const StepParams& p, const uint32_t* n_dev, const float4* pos, uint32_t* keys,
                    uint32_t key_bits, void* temp, size_t temp_bytes, uint32_t* cell_start,
                    uint32_t* cell_end, hipStream_t stream);
// The digit passes: (keys, slot) -> (skeys, perm).
int sort_cells(const StepParams& p, const uint32_t* keys, uint32_t* skeys, uint32_t* perm,
               uint32_t key_bits, void* temp, hipStream_t stream);

// ---- single-scale step kernels (step.hip) ----
struct Walls;  // boundary-particle walls (walls.hpp)

struct StepBuffers {
    float4* pos;        // state, sorted order of the previous step (w = id bits)
    float4* vel;
    float4* tpos;       // sorted copy used by the neighbour passes
    float4* tvel;
    float*  rho;
    float2* aux;        // per particle (k (rho - rho0) / rho^2, 1 / rho)
    float4* accel;      // optional (debug)
    uint32_t* keys;     // unsorted cell hashes
    uint32_t* skeys;    // sorted cell hashes
    uint32_t* perm;     // sorted slot -> previous slot
    uint32_t* cell_start;
    uint32_t* cell_end;
    void*  sort_temp;
    size_t sort_temp_bytes;
    uint32_t* n_dev;    // live particle count (device word)
    // Internal x-order of the neighbour passes (all NULL: they run in the
    // sorted order).  Within each cell the particles are ordered by x, so a
    // stencil row's candidates can be trimmed to the x-window; results are
    // written back at the sorted slots.
    float4* ipos;       // sorted state in the internal order
    float4* ivel;
    float2* iaux;       // aux in the internal order (density -> force)
    uint32_t* i2r;      // internal slot -> sorted slot
    uint4* xbins;       // per cell: cumulative counts of its kXBins x-bins (8 bits each)
    // Boundary-particle walls (NULL: none); bacc holds each sorted slot's
    // boundary acceleration, computed before the force pass adds it.
    const Walls* walls;
    float4* bacc;
    // Colour value per sorted slot (Mexican-hat energy), computed with the
    // force pass when non-NULL.
    float* color;
    // The reorder pass writes the sorted copies tpos/tvel even with the
    // x-order (the two-scale step's other passes read them).
    bool keep_sorted;
};

// Passes of one step, in launch order.
enum StepPass { kPassHash = 0, kPassSort, kPassReorder, kPassDensity, kPassForce, kNumPasses };
int launch_pass(int pass, const StepParams& p, const StepBuffers& b, float dt, uint32_t key_bits,
                bool write_accel, hipStream_t stream);
int launch_step(const StepParams& p, const StepBuffers& b, float dt, uint32_t key_bits,
                bool write_accel, hipStream_t stream);
// Two-scale building blocks, in the internal x-order (b.ipos set; results at
// the sorted slots): own-set density without the A/B finalisation, the
// internal-order copy of the finalised A/B, and the own-set acceleration
// (gravity included) without integration.
struct GridView;
int launch_density_own_x(const StepParams& p, const StepBuffers& b, hipStream_t stream);
int launch_gather_aux(const StepParams& p, const StepBuffers& b, hipStream_t stream);
// dst[s] = src[i2r[s]] for the live slots s < *n_dev of a capacity of cap.
int launch_gather_x(const float2* src, const uint32_t* i2r, const uint32_t* n_dev, uint32_t cap,
                    float2* dst, hipStream_t stream);
int launch_force_own_x(const StepParams& p, const GridView& g, const StepBuffers& b,
                       float4* accel_own, hipStream_t stream);
int launch_pack_state(const float* pos_xyz_dev, const float* vel_xyz_dev, uint32_t n,
                      float4* pos, float4* vel, hipStream_t stream);
int launch_export_aos(const float4* src, uint32_t n, float* dst_xyz, hipStream_t stream);
// Candidates and in-range pairs (out[0], out[1], added to) of one neighbour
// pass of the most recent step.
int launch_count_pairs(const StepParams& p, const StepBuffers& b, unsigned long long* out,
                       hipStream_t stream);

}  // namespace sph

namespace sph {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  Neighbour passes process particles in sorted (spatial)
// order, so consecutive blocks share most of their neighbour data.  This maps
// hardware block b to a logical block so that each XCD receives chunks of
// kXcdChunk consecutive logical blocks (dealt to the XCDs in turn, which
// keeps the per-XCD work balanced when cost varies along the sorted order)
// and its L2 sees spatially local data.  Blocks past the last whole round of
// chunks keep their own index, so the map is a bijection for any grid size.
// 16-block chunks (4,096 particles) ran the 1M-particle trajectory about 1%
// faster than 4 or 8; 64 and more lost it again to imbalance at the end.
constexpr uint32_t kNumXcds = 8;
#ifndef SPH_XCD_CHUNK
#define SPH_XCD_CHUNK 16
#endif
constexpr uint32_t kXcdChunk = SPH_XCD_CHUNK;

__device__ __forceinline__ uint32_t xcd_local_block(uint32_t b, uint32_t nblocks) {
    if (kXcdChunk == 0) return b;
    const uint32_t round = kNumXcds * kXcdChunk;
    const uint32_t full = (nblocks / round) * round;
    if (b >= full) return b;
    const uint32_t x = b % kNumXcds, k = b / kNumXcds;
    return ((k / kXcdChunk) * kNumXcds + x) * kXcdChunk + (k % kXcdChunk);
}

}  // namespace sph
