// diagnostics.hpp -- colour values, renderer buffer export and surface marks
// (diagnostics.hip).
#pragma once

#include "neighbors.hpp"

namespace sph {

// Constants of the colour-field surface detector of the two-scale solver.
struct SurfaceParams {
    float h0, h1;        // low / high radii
    float c45_0, c45_1;  // 45 / (pi h^6)
    float m0, m1;        // low / high masses
    float threshold;     // |grad c| h0 above which a low particle is at the surface
};

// color[i] (sorted slot i of one set) from the step's sorted state.
int launch_energy_color(const float4* tpos, const float4* tvel, const uint32_t* cs,
                        const uint32_t* ce, const uint32_t* n_dev, uint32_t cap, const GridView& g,
                        float mass, float* color, hipStream_t st);
// xyz (3n floats), color (n floats) and state (n int32) from one set's state
// (count *n_dev); NULL outputs are skipped; blend NULL gives default_state.
int launch_export_render(const float4* pos, const float2* blend, const float* color_in,
                         const uint32_t* n_dev, uint32_t cap, uint32_t default_state, float* xyz,
                         float* color, int32_t* state, hipStream_t st);
// flags[id] = 1 for low particles (step's sorted state) at the free surface.
int launch_surface_detect(const float4* tpos_l, const float2* aux_l, const uint32_t* cs_l,
                          const uint32_t* ce_l, const uint32_t* n_l, uint32_t cap_l,
                          const GridView& gl, const float4* tpos_h, const float2* aux_h,
                          const uint32_t* cs_h, const uint32_t* ce_h, const uint32_t* n_h,
                          const GridView& gh, const SurfaceParams& sp, uint32_t* flags,
                          uint32_t num_ids, hipStream_t st);
// Bit 8 (and, with split, the split request of state-8 particles) for the
// flagged ids of the compacted low set; clears the flags it uses.
int launch_apply_surface(const float4* pos, const uint32_t* n_dev, uint32_t cap, uint32_t* flags,
                         uint32_t num_ids, uint32_t split, float2* blend, hipStream_t st);
// The same marks from a surface index map (entries slot + 1, 0 = none).
int launch_mark_surface_map(const uint32_t* map, uint32_t count, const uint32_t* n_dev,
                            uint32_t split, float2* blend, hipStream_t st);

}  // namespace sph
