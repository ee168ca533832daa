// batch.hip -- many independent scenes advanced together: one segmented hash
// and sort, one launch per pass for all scenes, the whole step replayed from a
// hipGraph.  The reference runs one scene per process (its configuration is a
// __constant__ global, SPH2S_FINAL_SINGLE_SCALE/Solver.cu:12); batching is an
// addition of this library.
//
// Layout: scenes are stored scene-major.  A particle's sort key is its scene's
// first global cell plus its reference cell hash in the scene's grid, so the
// stable sort keeps every scene contiguous, in the same order as a solver
// running that scene alone, and every pass of a scene computes the same values
// in the same order as the single-scene path.  Per-scene constants (grid,
// stiffness, viscosity, domain) come from a device table indexed by the scene
// id carried in vel.w.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "batch.hpp"

using namespace sph;

namespace {

size_t al(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

void destroy_graph(sph_batch* b) {
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
}

int capture(sph_batch* b, float dt, bool accel) {
    destroy_graph(b);
    hipGraph_t g = nullptr;
    SPH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_step(b, b->buf, dt, accel, b->stream);
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc != SPH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return hip_error(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    b->graph_dt = dt;
    b->graph_accel = accel;
    return SPH_OK;
}

SceneParams scene_params(const sph_config& c, uint32_t count, uint32_t cell_offset) {
    SceneParams sp;
    sp.p = make_step_params(c, count);
    const StepParams& p = sp.p;
    sp.g = make_grid_view(p.ox, p.oy, p.oz, p.spacing, p.dx, p.dy, p.dz, p.h);
    sp.g.cell_offset = cell_offset;
    return sp;
}

}  // namespace

namespace sph {

int batch_alloc(const sph_config* cfgs, uint32_t n_scenes, const uint32_t* counts,
                sph_batch** out) {
    if (!out || !cfgs || !counts || n_scenes == 0)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "bad batch arguments");
    *out = nullptr;
    uint64_t n = 0, cells = 0;
    for (uint32_t s = 0; s < n_scenes; ++s) {
        const sph_config& c = cfgs[s];
        if (!(c.h > 0.0f) || !(c.mass > 0.0f) || !(c.rest_density > 0.0f) ||
            !(c.grid.spacing > 0.0f) || c.grid.dims[0] < 1 || c.grid.dims[1] < 1 ||
            c.grid.dims[2] < 1)
            return set_error(SPH_ERR_INVALID_ARGUMENT, "batch: invalid scene configuration");
        if (c.flags & SPH_FLAG_XSPH)
            return set_error(SPH_ERR_UNSUPPORTED, "batch: the XSPH integrator is single-solver only");
        n += counts[s];
        cells += sph_grid_num_cells(&c.grid);
    }
    if (n == 0) return set_error(SPH_ERR_INVALID_ARGUMENT, "batch: no particles");
    if (n >= (1ull << 30) || cells >= (1ull << 31))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "batch: too many particles or cells");
    int dcount = 0;
    if (hipGetDeviceCount(&dcount) != hipSuccess || dcount == 0) {
        (void)hipGetLastError();
        return set_error(SPH_ERR_NO_DEVICE, "no HIP device visible");
    }
    auto* b = new (std::nothrow) sph_batch;
    if (!b) return set_error(SPH_ERR_INVALID_ARGUMENT, "out of host memory");
    b->n_scenes = n_scenes;
    b->n = static_cast<uint32_t>(n);
    b->total_cells = static_cast<uint32_t>(cells);
    b->key_bits = key_bits_for_cells(b->total_cells);
    b->cfg.assign(cfgs, cfgs + n_scenes);
    b->part_offset.assign(n_scenes + 1, 0u);
    b->cell_offset.assign(n_scenes + 1, 0u);
    b->table.resize(n_scenes);
    b->kahan = false;
    for (uint32_t s = 0; s < n_scenes; ++s) {
        b->part_offset[s + 1] = b->part_offset[s] + counts[s];
        b->cell_offset[s + 1] = b->cell_offset[s] + sph_grid_num_cells(&cfgs[s].grid);
        b->table[s] = scene_params(cfgs[s], counts[s], b->cell_offset[s]);
        b->kahan = b->kahan || (cfgs[s].flags & SPH_FLAG_KAHAN) != 0u;
        b->tension = b->tension || cfgs[s].tension != 0.0f;
    }
    b->cap = make_step_params(cfgs[0], b->n);
    b->cap.num_cells = b->total_cells;

    const size_t N = b->n, C = b->total_cells;
    const size_t sort_temp = radix_sort_temp_bytes(b->n);
    const size_t sizes[] = {al(16 * N), al(16 * N), al(16 * N), al(16 * N), al(4 * N),
                            al(16 * N), al(4 * N),  al(4 * N),  al(4 * N),  al(4 * C),
                            al(4 * C),  al(sort_temp), al(8 * N), al(4),
                            al(sizeof(SceneParams) * n_scenes), al(4 * (n_scenes + 1)),
                            al(16 * N), al(16 * N), al(8 * N), al(4 * N), al(kXBinBytes * C)};
    size_t total = 0;
    for (size_t v : sizes) total += v;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete b;
        return hip_error(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    e = hipMalloc(&b->arena, total);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(b->stream);
        delete b;
        return hip_error(e, "hipMalloc", __FILE__, __LINE__);
    }
    char* a = static_cast<char*>(b->arena);
    size_t k = 0;
    auto take = [&]() {
        void* p = a;
        a += sizes[k++];
        return p;
    };
    b->buf.pos = static_cast<float4*>(take());
    b->buf.vel = static_cast<float4*>(take());
    b->buf.tpos = static_cast<float4*>(take());
    b->buf.tvel = static_cast<float4*>(take());
    b->buf.rho = static_cast<float*>(take());
    b->buf.accel = static_cast<float4*>(take());
    b->buf.keys = static_cast<uint32_t*>(take());
    b->buf.skeys = static_cast<uint32_t*>(take());
    b->buf.perm = static_cast<uint32_t*>(take());
    b->buf.cell_start = static_cast<uint32_t*>(take());
    b->buf.cell_end = static_cast<uint32_t*>(take());
    b->buf.sort_temp = take();
    b->buf.sort_temp_bytes = sort_temp;
    b->buf.aux = static_cast<float2*>(take());
    b->buf.n_dev = static_cast<uint32_t*>(take());
    b->scenes = static_cast<SceneParams*>(take());
    b->part_offset_dev = static_cast<uint32_t*>(take());
    b->buf.ipos = static_cast<float4*>(take());
    b->buf.ivel = static_cast<float4*>(take());
    b->buf.iaux = static_cast<float2*>(take());
    b->buf.i2r = static_cast<uint32_t*>(take());
    b->buf.xbins = static_cast<uint4*>(take());

    e = hipMemsetAsync(b->arena, 0, total, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->scenes, b->table.data(), sizeof(SceneParams) * n_scenes,
                           hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->part_offset_dev, b->part_offset.data(), 4 * (n_scenes + 1),
                           hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->buf.n_dev, &b->n, 4, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->buf.cell_start, 0xFF, 4 * C, b->stream);
    if (e == hipSuccess) e = hipMemsetAsync(b->buf.cell_end, 0xFF, 4 * C, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        sph_batch_destroy(b);
        return hip_error(e, "batch upload", __FILE__, __LINE__);
    }
    *out = b;
    return SPH_OK;
}

int batch_step(const sph_batch* b, const StepBuffers& buf, float dt, bool accel,
               hipStream_t stream) {
    for (int pass = 0; pass < kNumPasses; ++pass) {
        const int rc = launch_batch_pass(pass, b->cap, buf, b->scenes, b->kahan, b->tension, dt,
                                         b->key_bits, accel, stream);
        if (rc != SPH_OK) return rc;
    }
    return SPH_OK;
}

int batch_set_parameters(sph_batch* b, const float* stiffness, const float* viscosity) {
    if (!stiffness && !viscosity) return SPH_OK;
    for (uint32_t s = 0; s < b->n_scenes; ++s) {
        if (stiffness) {
            if (!std::isfinite(stiffness[s]))
                return set_error(SPH_ERR_INVALID_ARGUMENT, "stiffness must be finite");
            b->cfg[s].stiffness = stiffness[s];
        }
        if (viscosity) {
            if (!std::isfinite(viscosity[s]))
                return set_error(SPH_ERR_INVALID_ARGUMENT, "viscosity must be finite");
            b->cfg[s].viscosity = viscosity[s];
        }
        b->table[s] = scene_params(b->cfg[s], b->part_offset[s + 1] - b->part_offset[s],
                                   b->cell_offset[s]);
    }
    SPH_HIP(hipMemcpyAsync(b->scenes, b->table.data(), sizeof(SceneParams) * b->n_scenes,
                           hipMemcpyHostToDevice, b->stream));
    // The table is read from device memory by the captured step, so a graph
    // stays valid; the host copy must outlive the asynchronous copy.
    SPH_HIP(hipStreamSynchronize(b->stream));
    return SPH_OK;
}

}  // namespace sph

extern "C" {

int sph_batch_create(const sph_config* cfgs, uint32_t n_scenes, const uint32_t* counts,
                     const float* pos, const float* vel, sph_batch** out) {
    if (!pos) return set_error(SPH_ERR_INVALID_ARGUMENT, "batch: no particles");
    sph_batch* b = nullptr;
    const int rc = batch_alloc(cfgs, n_scenes, counts, &b);
    if (rc != SPH_OK) return rc;
    // Host-side packing: pos.w = global particle id, vel.w = scene id.
    const size_t N = b->n;
    std::vector<float4> p4(N), v4(N);
    for (uint32_t s = 0; s < n_scenes; ++s)
        for (uint32_t i = b->part_offset[s]; i < b->part_offset[s + 1]; ++i) {
            uint32_t id = i, sc = s;
            float fid, fsc;
            std::memcpy(&fid, &id, 4);
            std::memcpy(&fsc, &sc, 4);
            p4[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], fid);
            v4[i] = vel ? make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], fsc)
                        : make_float4(0.f, 0.f, 0.f, fsc);
        }
    hipError_t e = hipMemcpyAsync(b->buf.pos, p4.data(), 16 * N, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(b->buf.vel, v4.data(), 16 * N, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        sph_batch_destroy(b);
        return hip_error(e, "batch upload", __FILE__, __LINE__);
    }
    *out = b;
    return SPH_OK;
}

int sph_batch_destroy(sph_batch* b) {
    if (!b) return SPH_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    destroy_graph(b);
    if (b->arena) (void)hipFree(b->arena);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SPH_OK;
}

int sph_batch_advance(sph_batch* b, float dt, uint32_t n_steps) {
    if (!b) return set_error(SPH_ERR_INVALID_ARGUMENT, "batch is null");
    if (!(dt > 0.0f) || !std::isfinite(dt)) return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (n_steps == 0) return SPH_OK;
    const bool accel = (b->debug & SPH_DEBUG_ACCEL) != 0u;
    if (!b->graph || b->graph_dt != dt || b->graph_accel != accel) {
        const int rc = capture(b, dt, accel);
        if (rc != SPH_OK) return rc;
    }
    for (uint32_t i = 0; i < n_steps; ++i) SPH_HIP(hipGraphLaunch(b->graph, b->stream));
    return SPH_OK;
}

int sph_batch_advance_eager(sph_batch* b, float dt, uint32_t n_steps) {
    if (!b) return set_error(SPH_ERR_INVALID_ARGUMENT, "batch is null");
    if (!(dt > 0.0f)) return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    const bool accel = (b->debug & SPH_DEBUG_ACCEL) != 0u;
    for (uint32_t i = 0; i < n_steps; ++i) {
        const int rc = batch_step(b, b->buf, dt, accel, b->stream);
        if (rc != SPH_OK) return rc;
    }
    return SPH_OK;
}

int sph_batch_autotune(sph_batch* b, float dt, uint32_t reps, uint32_t* tune, float* ms) {
    if (!b) return set_error(SPH_ERR_INVALID_ARGUMENT, "batch is null");
    if (!(dt > 0.0f) || !std::isfinite(dt))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    uint32_t best = 0;
    const int rc = tune_neighbour_passes(b->cap, b->buf, b->scenes, b->kahan, b->tension, dt,
                                         b->key_bits, reps, b->stream, &best, ms);
    if (rc != SPH_OK) return rc;
    b->cap.tune = best;
    destroy_graph(b);
    if (tune) *tune = best;
    return SPH_OK;
}

int sph_batch_set_tune(sph_batch* b, uint32_t tune) {
    if (!b) return set_error(SPH_ERR_INVALID_ARGUMENT, "batch is null");
    if (!tune_word_valid(tune, true)) return set_error(SPH_ERR_INVALID_ARGUMENT, "invalid tune word");
    if (tune != b->cap.tune) {
        b->cap.tune = tune;
        destroy_graph(b);
    }
    return SPH_OK;
}

int sph_batch_get_tune(const sph_batch* b, uint32_t* tune) {
    if (!b || !tune) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    *tune = b->cap.tune;
    return SPH_OK;
}

int sph_batch_synchronize(sph_batch* b) {
    if (!b) return set_error(SPH_ERR_INVALID_ARGUMENT, "batch is null");
    uint32_t v = 0;
    const int rc = read_error_word(radix_sort_error_word(b->buf.sort_temp), b->stream, &v);
    return rc != SPH_OK ? rc : device_error_status(public_error_bits(v));
}

int sph_batch_set_debug(sph_batch* b, uint32_t flags) {
    if (!b) return set_error(SPH_ERR_INVALID_ARGUMENT, "batch is null");
    b->debug = flags;
    return SPH_OK;
}

int sph_batch_info(const sph_batch* b, uint32_t* n_scenes, uint32_t* n_particles,
                   uint32_t* n_cells, uint32_t* part_offsets, uint32_t* cell_offsets) {
    if (!b) return set_error(SPH_ERR_INVALID_ARGUMENT, "batch is null");
    if (n_scenes) *n_scenes = b->n_scenes;
    if (n_particles) *n_particles = b->n;
    if (n_cells) *n_cells = b->total_cells;
    if (part_offsets) std::copy(b->part_offset.begin(), b->part_offset.end(), part_offsets);
    if (cell_offsets) std::copy(b->cell_offset.begin(), b->cell_offset.end(), cell_offsets);
    return SPH_OK;
}

int sph_batch_device_errors(sph_batch* b, uint32_t* bits) {
    if (!b || !bits) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t* word = radix_sort_error_word(b->buf.sort_temp);
    uint32_t v = 0;
    SPH_HIP(hipMemcpyAsync(&v, word, 4, hipMemcpyDeviceToHost, b->stream));
    SPH_HIP(hipStreamSynchronize(b->stream));
    if (v) {
        SPH_HIP(hipMemsetAsync(word, 0, 4, b->stream));
        SPH_HIP(hipStreamSynchronize(b->stream));
    }
    *bits = public_error_bits(v);
    return SPH_OK;
}

int sph_batch_profile_passes(sph_batch* b, float dt, uint32_t n_steps, float* ms_out) {
    if (!b || !ms_out) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    const bool accel = (b->debug & SPH_DEBUG_ACCEL) != 0u;
    hipEvent_t ev[kNumPasses + 1] = {};
    for (auto& e : ev) SPH_HIP(hipEventCreate(&e));
    int rc = SPH_OK;
    for (uint32_t step = 0; step < n_steps && rc == SPH_OK; ++step) {
        (void)hipEventRecord(ev[0], b->stream);
        for (int pass = 0; pass < kNumPasses && rc == SPH_OK; ++pass) {
            rc = launch_batch_pass(pass, b->cap, b->buf, b->scenes, b->kahan, b->tension, dt,
                                   b->key_bits, accel, b->stream);
            (void)hipEventRecord(ev[pass + 1], b->stream);
        }
        if (rc != SPH_OK) break;
        const hipError_t e = hipEventSynchronize(ev[kNumPasses]);
        if (e != hipSuccess) {
            rc = hip_error(e, "hipEventSynchronize", __FILE__, __LINE__);
            break;
        }
        for (int pass = 0; pass < kNumPasses; ++pass) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev[pass], ev[pass + 1]);
            ms_out[pass] += ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return rc;
}

int sph_batch_download(sph_batch* b, int field, void* dst, size_t bytes) {
    if (!b || !dst) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    const size_t n = b->n, C = b->total_cells;
    const float4* vec = nullptr;
    const void* src = nullptr;
    size_t need = 0;
    int mode = 0;
    switch (field) {
        case SPH_FIELD_POSITION: vec = b->buf.pos; need = 12 * n; mode = 1; break;
        case SPH_FIELD_VELOCITY: vec = b->buf.vel; need = 12 * n; mode = 1; break;
        case SPH_FIELD_ID: vec = b->buf.pos; need = 4 * n; mode = 2; break;
        case SPH_FIELD_TAG: vec = b->buf.vel; need = 4 * n; mode = 2; break;  // scene ids
        case SPH_FIELD_DENSITY: src = b->buf.rho; need = 4 * n; break;
        case SPH_FIELD_SORTED_HASH: src = b->buf.skeys; need = 4 * n; break;
        case SPH_FIELD_SORTED_INDEX: src = b->buf.perm; need = 4 * n; break;
        case SPH_FIELD_CELL_START: src = b->buf.cell_start; need = 4 * C; break;
        case SPH_FIELD_CELL_END: src = b->buf.cell_end; need = 4 * C; break;
        case SPH_FIELD_ACCEL:
            if (!(b->debug & SPH_DEBUG_ACCEL))
                return set_error(SPH_ERR_INVALID_ARGUMENT, "accelerations need SPH_DEBUG_ACCEL");
            vec = b->buf.accel; need = 12 * n; mode = 1; break;
        default: return set_error(SPH_ERR_INVALID_ARGUMENT, "unknown field");
    }
    if (bytes < need) return set_error(SPH_ERR_INVALID_ARGUMENT, "destination too small");
    if (mode == 0) {
        SPH_HIP(hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, b->stream));
        SPH_HIP(hipStreamSynchronize(b->stream));
        return SPH_OK;
    }
    std::vector<float> tmp(4 * n);
    SPH_HIP(hipMemcpyAsync(tmp.data(), vec, 16 * n, hipMemcpyDeviceToHost, b->stream));
    SPH_HIP(hipStreamSynchronize(b->stream));
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

}  // extern "C"
