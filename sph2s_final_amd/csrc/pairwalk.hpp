// pairwalk.hpp -- block-cooperative walk over a particle's reference stencil
// with each grid row's candidates staged in LDS (used by the adjoint passes,
// the XSPH integrator and the blended two-scale passes).
//
// The nine grid rows (cy + dy, cz + dz), dy, dz in {-1, 0, 1}, are block-wide
// phases: each lane finds its candidate run in the row, the block takes the
// union of its lanes' runs (one contiguous range when the block's targets
// share a row) and stages up to kTile candidates of it in LDS with coalesced
// loads; each lane then scans its own run from LDS, or from global memory when
// the run is not inside the staged range.  Rows the reference stencil reaches
// beyond the nine (rounding can put a bound two cells away) are scanned from
// global memory afterwards.  The candidate set is the same as for_each_run's.
#pragma once

#include "neighbors.hpp"

namespace sph {

// Keeps all four components of a candidate read from LDS live at its point of
// use, so the compiler reads the whole 16-byte element (ds_read_b128) instead
// of narrowing the access to the 12 bytes of x, y, z: on gfx950 the 12-byte
// form is markedly slower (the single-scale density pass took about 20%
// longer in compressed states).  The empty asm only pins a register; it
// touches no memory.
__device__ __forceinline__ void keep_lds_quad(const float4& q) {
    asm volatile("" ::"v"(q.w));
}

// stage(k, j): write candidate j (global index) into LDS slot k.
// tile_run(t0, t1, base): scan LDS slots [t0, t1) (global index base + t).
// global_run(b, e): scan global indices [b, e).
// kX: the indices are of the internal x-order (cells sorted by x inside) and
// rows are trimmed with the cells' x-bin tables `bins` (probe_row_run_x).
// kPad: LDS slots [seg_n, seg_n + kPad) after a staged segment also get a
// candidate (copies of its last one), so a scan that loads whole groups past
// the end of its run reads finite data (the tile needs kTile + kPad slots).
// LDS is not cleared between kernels: without the pad those slots hold
// whatever an earlier workgroup left there, NaN bit patterns included.
// Every lane of the block must call this (it synchronises the block);
// inactive lanes pass active = false.
template <int kBlockThreads, int kTile, bool kX = false, int kPad = 0, class Stage, class TileRun,
          class GlobalRun>
__device__ __forceinline__ void walk_stencil_staged(const GridView& g, const Probe& pr,
                                                    const uint32_t* __restrict__ cs,
                                                    const uint32_t* __restrict__ ce, uint32_t n,
                                                    bool active, uint32_t* red, Stage&& stage,
                                                    TileRun&& tile_run, GlobalRun&& global_run,
                                                    const uint4* __restrict__ bins = nullptr) {
    for (int dz = -1; dz <= 1; ++dz) {
        for (int dy = -1; dy <= 1; ++dy) {
            uint32_t b = kEmptyCell, e = 0u;
            if (active) probe_row<kX>(g, pr, cs, ce, bins, n, pr.cy + dy, pr.cz + dz, b, e);
            uint32_t ub, ue;
            block_run_union<kBlockThreads>(b, e, red, ub, ue);
            const uint32_t seg_n = (ue > ub) ? min(ue - ub, static_cast<uint32_t>(kTile)) : 0u;
            const uint32_t seg_e = ub + seg_n;
            __syncthreads();
            if constexpr (kPad == 0) {
                for (uint32_t k = threadIdx.x; k < seg_n; k += kBlockThreads) stage(k, ub + k);
            } else {
                const uint32_t slots = seg_n ? seg_n + kPad : 0u;
                for (uint32_t k = threadIdx.x; k < slots; k += kBlockThreads)
                    stage(k, ub + min(k, seg_n - 1));
            }
            __syncthreads();
            if (b < e) {
                if (b >= ub && e <= seg_e)
                    tile_run(b - ub, e - ub, ub);
                else
                    global_run(b, e);
            }
        }
    }
    if (!active) return;
    for (int rz = pr.sz; rz <= pr.ez; ++rz)
        for (int ry = pr.sy; ry <= pr.ey; ++ry) {
            if (abs(ry - pr.cy) <= 1 && abs(rz - pr.cz) <= 1) continue;
            uint32_t b, e;
            probe_row<kX>(g, pr, cs, ce, bins, n, ry, rz, b, e);
            if (b < e) global_run(b, e);
        }
}

}  // namespace sph
