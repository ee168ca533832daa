// batch.hpp -- the batched-scenes solver object, shared by batch.hip (the
// plain batched step) and adjoint.hip (the differentiable batched rollout).
#pragma once

#include <vector>

#include "neighbors.hpp"

struct sph_batch {
    uint32_t n_scenes = 0, n = 0, total_cells = 0, key_bits = 1;
    std::vector<uint32_t> part_offset;   // n_scenes + 1
    std::vector<uint32_t> cell_offset;   // n_scenes + 1
    std::vector<sph_config> cfg;
    std::vector<sph::SceneParams> table; // host copy of the device table
    sph::StepParams cap{};
    bool kahan = true, tension = false;
    sph::StepBuffers buf{};
    sph::SceneParams* scenes = nullptr;  // device table
    uint32_t* part_offset_dev = nullptr; // device copy of part_offset
    hipStream_t stream = nullptr;
    void* arena = nullptr;
    hipGraphExec_t graph = nullptr;
    float graph_dt = 0.0f;
    bool graph_accel = false;
    uint32_t debug = 0;
};

namespace sph {

// Allocates a batch with a zero state (positions and velocities must be
// written before the first step) and uploads the scene table.
int batch_alloc(const sph_config* cfgs, uint32_t n_scenes, const uint32_t* counts,
                sph_batch** out);
// Enqueues one step's passes on `stream` over the buffers `buf` (the batch's
// own or a copy pointing elsewhere).
int batch_step(const sph_batch* b, const StepBuffers& buf, float dt, bool accel,
               hipStream_t stream);
// Rewrites scene s's stiffness and viscosity (host values; NULL keeps the
// current ones) in the host and device tables.  Enqueued on b->stream.
int batch_set_parameters(sph_batch* b, const float* stiffness, const float* viscosity);

}  // namespace sph
