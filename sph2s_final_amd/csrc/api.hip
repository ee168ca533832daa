// api.hip -- the C ABI of include/sph2s.h: configuration helpers, the scene
// generator, the solver handle (device buffers, stream, captured step graph)
// and the standalone radix sort.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "diagnostics.hpp"
#include "neighbors.hpp"
#include "walls.hpp"

namespace sph {

namespace {
thread_local std::string g_last_error;
constexpr double kPi = 3.14159265358979323846;  // SPH2S_FINAL_SINGLE_SCALE/Solver.h:11
}  // namespace

int set_error(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hip_error(hipError_t e, const char* what, const char* file, int line) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file,
                  line);
    return set_error(SPH_ERR_HIP, buf);
}

int device_error_status(uint32_t bits) {
    if (bits & SPH_DEVERR_NONFINITE)
        return set_error(SPH_ERR_NONFINITE,
                         "a step wrote a NaN or infinite position or velocity (device error bits)");
    if (bits & SPH_DEVERR_CAPACITY)
        return set_error(SPH_ERR_CAPACITY,
                         "a split or merge exceeded a set's capacity; particles were dropped");
    if (bits & SPH_DEVERR_SORT_WAIT)
        return set_error(SPH_ERR_DEVICE,
                         "a radix-sort wait gave up: that step's particle order is invalid");
    return SPH_OK;
}

int read_error_word(const uint32_t* word, hipStream_t stream, uint32_t* out) {
    SPH_HIP(hipMemcpyAsync(out, word, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    SPH_HIP(hipStreamSynchronize(stream));
    return SPH_OK;
}

StepParams make_step_params(const sph_config& c, uint32_t n) {
    StepParams p{};
    p.ox = c.grid.origin[0];
    p.oy = c.grid.origin[1];
    p.oz = c.grid.origin[2];
    p.dx = c.grid.dims[0];
    p.dy = c.grid.dims[1];
    p.dz = c.grid.dims[2];
    p.spacing = c.grid.spacing;
    p.h = c.h;
    p.h2 = c.h * c.h;
    p.rest_density = c.rest_density;
    p.stiffness = c.stiffness;
    const double h = c.h;
    const double h3 = h * h * h;
    const double c6 = 315.0 / (64.0 * kPi * h3 * h3 * h3);
    const double c45 = 45.0 / (kPi * h3 * h3);
    p.c6m = static_cast<float>(c.mass * c6);
    p.c6m_h6 = static_cast<float>(c.mass * 315.0 / (64.0 * kPi * h3));
    p.c45m = static_cast<float>(c.mass * c45);
    p.visc = static_cast<float>(static_cast<double>(c.viscosity) * c.mass * c45);
    p.tens = static_cast<float>(static_cast<double>(c.tension) * c.mass * c6);
    p.xsph = static_cast<float>(static_cast<double>(c.xsph) * c.mass * c6);
    p.mass = c.mass;
    p.gx = c.gravity[0];
    p.gy = c.gravity[1];
    p.gz = c.gravity[2];
    p.restitution = c.restitution;
    p.lox = c.domain_lo[0];
    p.loy = c.domain_lo[1];
    p.loz = c.domain_lo[2];
    p.hix = c.domain_hi[0];
    p.hiy = c.domain_hi[1];
    p.hiz = c.domain_hi[2];
    p.flags = c.flags;
    p.n = n;
    p.num_cells = sph_grid_num_cells(&c.grid);
    return p;
}

uint32_t key_bits_for_cells(uint32_t num_cells) {
    if (num_cells <= 2) return 1;
    uint32_t bits = 0;
    uint32_t m = num_cells - 1;
    while (m) {
        ++bits;
        m >>= 1;
    }
    return bits;
}

}  // namespace sph

using namespace sph;

struct sph_solver {
    sph_config cfg{};
    uint32_t n = 0, max_n = 0, num_cells = 0, key_bits = 1;
    int device = 0;
    hipStream_t stream = nullptr;
    StepBuffers buf{};
    void* arena = nullptr;       // one allocation holds every buffer
    hipGraphExec_t graph = nullptr;
    float graph_dt = 0.0f;
    bool graph_accel = false;
    bool stepped = false;        // a step ran since the state was last set
    uint32_t debug = 0;
    uint32_t tune = 0;           // launch variant of the neighbour passes
    Walls walls{};               // boundary particles (none by default)
    float4* bacc = nullptr;      // boundary acceleration per sorted slot
    float* color = nullptr;      // colour values (SPH_DEBUG_COLOR)
    // Renderer buffers (sph_enable_render_buffers): AoS xyz, colour, state.
    float* render_xyz = nullptr;
    float* render_color = nullptr;
    int32_t* render_state = nullptr;
    uint32_t render_flags = 0;
};

namespace {

int validate_config(const sph_config* c) {
    if (!c) return set_error(SPH_ERR_INVALID_ARGUMENT, "config is null");
    if (!(c->h > 0.0f) || !std::isfinite(c->h))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "config: h must be > 0");
    if (!(c->mass > 0.0f)) return set_error(SPH_ERR_INVALID_ARGUMENT, "config: mass must be > 0");
    if (!(c->rest_density > 0.0f))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "config: rest_density must be > 0");
    if (!(c->grid.spacing > 0.0f))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "config: grid spacing must be > 0");
    for (int d = 0; d < 3; ++d)
        if (c->grid.dims[d] < 1)
            return set_error(SPH_ERR_INVALID_ARGUMENT, "config: grid dims must be >= 1");
    const uint64_t cells = static_cast<uint64_t>(c->grid.dims[0]) * c->grid.dims[1] * c->grid.dims[2];
    if (cells > 0x7FFFFFFFull)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "config: too many grid cells");
    if ((c->flags & SPH_FLAG_WALL_CLAMP) != 0u)
        for (int d = 0; d < 3; ++d)
            if (!(c->domain_hi[d] >= c->domain_lo[d]))
                return set_error(SPH_ERR_INVALID_ARGUMENT, "config: domain_hi < domain_lo");
    return SPH_OK;
}

size_t align_up(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

int ensure_device(int* dev) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return set_error(SPH_ERR_NO_DEVICE, "no HIP device visible");
    }
    SPH_HIP(hipGetDevice(dev));
    return SPH_OK;
}

void destroy_graph(sph_solver* s) {
    if (s->graph) {
        (void)hipGraphExecDestroy(s->graph);
        s->graph = nullptr;
    }
}

int upload_state(sph_solver* s, const float* pos, const float* vel, uint32_t n) {
    if (n > s->max_n) return set_error(SPH_ERR_CAPACITY, "n exceeds max_n");
    if (n > 0 && !pos) return set_error(SPH_ERR_INVALID_ARGUMENT, "positions are null");
    s->n = n;
    SPH_HIP(hipMemcpyAsync(s->buf.n_dev, &s->n, sizeof(uint32_t), hipMemcpyHostToDevice, s->stream));
    SPH_HIP(hipStreamSynchronize(s->stream));
    if (n == 0) return SPH_OK;
    // Stage AoS host arrays through the (not yet used) sorted-copy buffers and
    // pack them into float4 on the device.
    float* dpos = reinterpret_cast<float*>(s->buf.tpos);
    float* dvel = reinterpret_cast<float*>(s->buf.tvel);
    SPH_HIP(hipMemcpyAsync(dpos, pos, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s->stream));
    if (vel)
        SPH_HIP(hipMemcpyAsync(dvel, vel, sizeof(float) * 3 * n, hipMemcpyHostToDevice, s->stream));
    int rc = launch_pack_state(dpos, vel ? dvel : nullptr, n, s->buf.pos, s->buf.vel, s->stream);
    if (rc != SPH_OK) return rc;
    SPH_HIP(hipMemsetAsync(s->buf.rho, 0, sizeof(float) * s->max_n, s->stream));
    SPH_HIP(hipStreamSynchronize(s->stream));
    destroy_graph(s);
    s->stepped = false;
    return SPH_OK;
}

// The renderer buffers after a step (or a state upload): AoS positions and,
// with SPH_RENDER_COLORS, the colour values, in the state's order.
int export_render_buffers(sph_solver* s) {
    if (!s->render_xyz) return SPH_OK;
    const bool colors = (s->render_flags & SPH_RENDER_COLORS) && s->buf.color;
    return launch_export_render(s->buf.pos, nullptr, colors ? s->buf.color : nullptr,
                                s->buf.n_dev, s->max_n, 0u, s->render_xyz,
                                colors ? s->render_color : nullptr, nullptr, s->stream);
}

// One step: the passes, then the renderer buffers when they are enabled.
int launch_full_step(sph_solver* s, const StepParams& p, float dt, bool accel) {
    const int rc = launch_step(p, s->buf, dt, s->key_bits, accel, s->stream);
    return rc == SPH_OK ? export_render_buffers(s) : rc;
}

int capture_graph(sph_solver* s, float dt, bool accel) {
    destroy_graph(s);
    StepParams p = make_step_params(s->cfg, s->n);
    p.tune = s->tune;
    hipGraph_t g = nullptr;
    SPH_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_full_step(s, p, dt, accel);
    hipError_t e = hipStreamEndCapture(s->stream, &g);
    if (rc != SPH_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture", __FILE__, __LINE__);
    e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        s->graph = nullptr;
        return hip_error(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    s->graph_dt = dt;
    s->graph_accel = accel;
    return SPH_OK;
}

uint64_t splitmix64(uint64_t& state) {
    state += 0x9E3779B97F4A7C15ull;
    uint64_t z = state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t& state) {  // [0, 1)
    return static_cast<double>(splitmix64(state) >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

extern "C" {

const char* sph_last_error(void) { return g_last_error.c_str(); }

const char* sph_version(void) { return "sph2s-mi355x 0.1 (gfx950)"; }

int sph_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return count;
}

int sph_make_grid(const float origin[3], const float end[3], float spacing, sph_grid* out) {
    if (!origin || !end || !out) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (!(spacing > 0.0f)) return set_error(SPH_ERR_INVALID_ARGUMENT, "spacing must be > 0");
    for (int d = 0; d < 3; ++d) {
        out->origin[d] = origin[d];
        // static_cast<int>(std::ceil((end - origin)/spacing)) in FP32 (Grid.h:25-33)
        const float t = (end[d] - origin[d]) / spacing;
        out->dims[d] = static_cast<int32_t>(std::ceil(t));
    }
    out->spacing = spacing;
    return SPH_OK;
}

uint32_t sph_grid_num_cells(const sph_grid* g) {
    if (!g) return 0;
    return static_cast<uint32_t>(g->dims[0]) * static_cast<uint32_t>(g->dims[1]) *
           static_cast<uint32_t>(g->dims[2]);
}

int sph_config_init(sph_config* out, float h, float mass, float rest_density, float stiffness,
                    float viscosity, const float lo[3], const float hi[3]) {
    if (!out || !lo || !hi) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof(*out));
    out->h = h;
    out->mass = mass;
    out->rest_density = rest_density;
    out->stiffness = stiffness;
    out->viscosity = viscosity;
    out->tension = 0.0f;
    out->gravity[0] = 0.0f;
    out->gravity[1] = -9.81f;
    out->gravity[2] = 0.0f;
    out->restitution = 0.5f;
    for (int d = 0; d < 3; ++d) {
        out->domain_lo[d] = lo[d];
        out->domain_hi[d] = hi[d];
    }
    out->flags = SPH_FLAG_WALL_CLAMP | SPH_FLAG_KAHAN;
    out->speed_sound = 88.1472f;
    out->boundary_mass = mass;
    int rc = sph_make_grid(lo, hi, h, &out->grid);
    if (rc != SPH_OK) return rc;
    return validate_config(out);
}

int sph_make_configuration(const float origin[3], const float end[3], float volume,
                           uint32_t num_particles, uint32_t avg_particles, float rest_density,
                           float bulk_modulus, float viscosity, float speed_sound, float tension,
                           sph_config* out) {
    if (!out || num_particles == 0)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "bad MakeConfiguration arguments");
    // Solver.h:31-37: powf(0.75 V avg/(M_PI N), 1/3) with the double M_PI.
    const float h = std::pow(static_cast<float>(0.75f * volume * static_cast<float>(avg_particles) /
                                                (kPi * static_cast<float>(num_particles))),
                             1.0f / 3.0f);
    const float mass = rest_density * volume / static_cast<float>(num_particles);
    int rc = sph_config_init(out, h, mass, rest_density, bulk_modulus, viscosity, origin, end);
    if (rc != SPH_OK) return rc;
    out->tension = tension;
    out->flags = SPH_FLAG_KAHAN;  // the reference has no wall clamp
    out->speed_sound = speed_sound;
    out->boundary_mass = mass;    // Solver.h:38
    return SPH_OK;
}

int sph_config_set_xsph(sph_config* cfg, float epsilon) {
    if (!cfg) return set_error(SPH_ERR_INVALID_ARGUMENT, "config is null");
    if (!(epsilon >= 0.0f) || !std::isfinite(epsilon))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "xsph epsilon must be >= 0");
    cfg->xsph = epsilon;
    if (epsilon > 0.0f)
        cfg->flags |= SPH_FLAG_XSPH;
    else
        cfg->flags &= ~SPH_FLAG_XSPH;
    return SPH_OK;
}

int sph_scene_lattice(const int32_t n[3], float spacing, const float lo[3], float jitter_frac,
                      uint64_t seed, float* pos) {
    if (!n || !lo || !pos) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (n[0] < 0 || n[1] < 0 || n[2] < 0)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "negative lattice size");
    uint64_t st = seed;
    size_t c = 0;
    const double s = spacing;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const int idx[3] = {i, j, k};
                for (int d = 0; d < 3; ++d) {
                    const double u = uniform01(st);
                    const double jit = (2.0 * u - 1.0) * static_cast<double>(jitter_frac) * s;
                    pos[c + d] = static_cast<float>(static_cast<double>(lo[d]) +
                                                    (idx[d] + 0.5) * s + jit);
                }
                c += 3;
            }
    return SPH_OK;
}

int sph_scene_gaussian_velocity(uint32_t n, float sigma, uint64_t seed, float* vel) {
    if (!vel && n) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    uint64_t st = seed;
    for (size_t i = 0; i < 3 * static_cast<size_t>(n); ++i) {
        // Box-Muller, cosine branch; u1 in (0, 1].
        const double u1 = 1.0 - uniform01(st);
        const double u2 = uniform01(st);
        const double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
        vel[i] = static_cast<float>(sigma * z);
    }
    return SPH_OK;
}

int sph_create(const sph_config* cfg, const float* pos, const float* vel, uint32_t n,
               uint32_t max_n, sph_solver** out) {
    if (!out) return set_error(SPH_ERR_INVALID_ARGUMENT, "out is null");
    *out = nullptr;
    int rc = validate_config(cfg);
    if (rc != SPH_OK) return rc;
    if (max_n < n) max_n = n;
    if (max_n == 0) return set_error(SPH_ERR_INVALID_ARGUMENT, "max_n must be > 0");
    int dev = 0;
    rc = ensure_device(&dev);
    if (rc != SPH_OK) return rc;

    auto* s = new (std::nothrow) sph_solver;
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "out of host memory");
    s->cfg = *cfg;
    s->max_n = max_n;
    s->device = dev;
    s->num_cells = sph_grid_num_cells(&cfg->grid);
    s->key_bits = key_bits_for_cells(s->num_cells);

    const size_t N = max_n, C = s->num_cells;
    const size_t sort_temp = radix_sort_temp_bytes(max_n);
    const size_t sizes[] = {
        align_up(16 * N), align_up(16 * N), align_up(16 * N), align_up(16 * N),  // pos vel tpos tvel
        align_up(4 * N),  align_up(16 * N),                                      // rho accel
        align_up(4 * N),  align_up(4 * N),  align_up(4 * N),                     // keys skeys perm
        align_up(4 * C),  align_up(4 * C),  align_up(sort_temp),                 // cells, sort temp
        align_up(8 * N),                                                         // aux
        align_up(4),                                                             // live count
        align_up(16 * N), align_up(16 * N), align_up(8 * N),                    // ipos ivel iaux
        align_up(4 * N),  align_up(kXBinBytes * C)};                             // i2r xbins
    size_t total = 0;
    for (size_t b : sizes) total += b;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete s;
        return hip_error(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
    }
    e = hipMalloc(&s->arena, total);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(s->stream);
        delete s;
        return hip_error(e, "hipMalloc", __FILE__, __LINE__);
    }
    char* a = static_cast<char*>(s->arena);
    size_t off = 0;
    auto take = [&](size_t i) {
        void* p = a + off;
        off += sizes[i];
        return p;
    };
    s->buf.pos = static_cast<float4*>(take(0));
    s->buf.vel = static_cast<float4*>(take(1));
    s->buf.tpos = static_cast<float4*>(take(2));
    s->buf.tvel = static_cast<float4*>(take(3));
    s->buf.rho = static_cast<float*>(take(4));
    s->buf.accel = static_cast<float4*>(take(5));
    s->buf.keys = static_cast<uint32_t*>(take(6));
    s->buf.skeys = static_cast<uint32_t*>(take(7));
    s->buf.perm = static_cast<uint32_t*>(take(8));
    s->buf.cell_start = static_cast<uint32_t*>(take(9));
    s->buf.cell_end = static_cast<uint32_t*>(take(10));
    s->buf.sort_temp = take(11);
    s->buf.aux = static_cast<float2*>(take(12));
    s->buf.n_dev = static_cast<uint32_t*>(take(13));
    s->buf.ipos = static_cast<float4*>(take(14));
    s->buf.ivel = static_cast<float4*>(take(15));
    s->buf.iaux = static_cast<float2*>(take(16));
    s->buf.i2r = static_cast<uint32_t*>(take(17));
    s->buf.xbins = static_cast<uint4*>(take(18));
    s->buf.sort_temp_bytes = sort_temp;

    e = hipMemsetAsync(s->arena, 0, total, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->buf.cell_start, 0xFF, 4 * C, s->stream);
    if (e == hipSuccess) e = hipMemsetAsync(s->buf.cell_end, 0xFF, 4 * C, s->stream);
    if (e != hipSuccess) {
        sph_destroy(s);
        return hip_error(e, "hipMemsetAsync", __FILE__, __LINE__);
    }
    rc = upload_state(s, pos, vel, n);
    if (rc != SPH_OK) {
        sph_destroy(s);
        return rc;
    }
    *out = s;
    return SPH_OK;
}

int sph_destroy(sph_solver* s) {
    if (!s) return SPH_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    destroy_graph(s);
    walls_destroy(&s->walls);
    if (s->bacc) (void)hipFree(s->bacc);
    if (s->color) (void)hipFree(s->color);
    if (s->render_xyz) (void)hipFree(s->render_xyz);
    if (s->arena) (void)hipFree(s->arena);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return SPH_OK;
}

int sph_set_boundary(sph_solver* s, const float* pos, uint32_t n) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    SPH_HIP(hipStreamSynchronize(s->stream));
    destroy_graph(s);
    s->buf.walls = nullptr;
    s->buf.bacc = nullptr;
    const int rc = walls_create(&s->walls, s->cfg.grid, s->cfg.h, s->cfg.boundary_mass,
                                s->cfg.mass, s->cfg.speed_sound, pos, n, s->stream);
    if (rc != SPH_OK || n == 0) return rc;
    if (!s->bacc) SPH_HIP(hipMalloc(&s->bacc, 16ull * s->max_n));
    s->buf.walls = &s->walls;
    s->buf.bacc = s->bacc;
    return SPH_OK;
}

int sph_set_state(sph_solver* s, const float* pos, const float* vel, uint32_t n) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    int rc = upload_state(s, pos, vel, n);
    if (rc == SPH_OK && s->render_xyz) {
        rc = export_render_buffers(s);
        if (rc == SPH_OK) SPH_HIP(hipStreamSynchronize(s->stream));
    }
    return rc;
}

int sph_advance(sph_solver* s, float dt, uint32_t n_steps) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    if (!(dt > 0.0f) || !std::isfinite(dt))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    if (s->n == 0 || n_steps == 0) return SPH_OK;
    const bool accel = (s->debug & SPH_DEBUG_ACCEL) != 0u;
    if (!s->graph || s->graph_dt != dt || s->graph_accel != accel) {
        int rc = capture_graph(s, dt, accel);
        if (rc != SPH_OK) return rc;
    }
    for (uint32_t i = 0; i < n_steps; ++i) SPH_HIP(hipGraphLaunch(s->graph, s->stream));
    s->stepped = true;
    return SPH_OK;
}

int sph_advance_eager(sph_solver* s, float dt, uint32_t n_steps) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    if (!(dt > 0.0f) || !std::isfinite(dt))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    StepParams p = make_step_params(s->cfg, s->n);
    p.tune = s->tune;
    const bool accel = (s->debug & SPH_DEBUG_ACCEL) != 0u;
    for (uint32_t i = 0; i < n_steps; ++i) {
        int rc = launch_full_step(s, p, dt, accel);
        if (rc != SPH_OK) return rc;
    }
    if (n_steps > 0 && s->n > 0) s->stepped = true;
    return SPH_OK;
}

int sph_autotune(sph_solver* s, float dt, uint32_t reps, uint32_t* tune, float* ms) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    if (!(dt > 0.0f) || !std::isfinite(dt))
        return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    StepParams p = make_step_params(s->cfg, s->n);
    uint32_t best = 0;
    const int rc = tune_neighbour_passes(p, s->buf, nullptr, (s->cfg.flags & SPH_FLAG_KAHAN) != 0u,
                                         p.tens != 0.0f, dt, s->key_bits, reps, s->stream, &best,
                                         ms);
    if (rc != SPH_OK) return rc;
    s->tune = best;
    destroy_graph(s);
    if (tune) *tune = best;
    return SPH_OK;
}

int sph_set_tune(sph_solver* s, uint32_t tune) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    if (!tune_word_valid(tune, false)) return set_error(SPH_ERR_INVALID_ARGUMENT, "invalid tune word");
    if (tune != s->tune) {
        s->tune = tune;
        destroy_graph(s);
    }
    return SPH_OK;
}

int sph_get_tune(const sph_solver* s, uint32_t* tune) {
    if (!s || !tune) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    *tune = s->tune;
    return SPH_OK;
}

int sph_synchronize(sph_solver* s) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    uint32_t v = 0;
    const int rc = read_error_word(radix_sort_error_word(s->buf.sort_temp), s->stream, &v);
    return rc != SPH_OK ? rc : device_error_status(public_error_bits(v));
}

int sph_num_particles(const sph_solver* s, uint32_t* n) {
    if (!s || !n) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    *n = s->n;
    return SPH_OK;
}

int sph_get_config(const sph_solver* s, sph_config* out) {
    if (!s || !out) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    *out = s->cfg;
    return SPH_OK;
}

int sph_get_views(sph_solver* s, sph_views* v) {
    if (!s || !v) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    v->pos4 = s->buf.pos;
    v->vel4 = s->buf.vel;
    v->density = s->buf.rho;
    v->sorted_hash = s->buf.skeys;
    v->sorted_index = s->buf.perm;
    v->cell_start = s->buf.cell_start;
    v->cell_end = s->buf.cell_end;
    v->accel4 = (s->debug & SPH_DEBUG_ACCEL) ? s->buf.accel : nullptr;
    v->n = s->n;
    v->num_cells = s->num_cells;
    v->stream = s->stream;
    return SPH_OK;
}

int sph_profile_passes(sph_solver* s, float dt, uint32_t n_steps, float* ms_out) {
    if (!s || !ms_out) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (!(dt > 0.0f)) return set_error(SPH_ERR_INVALID_ARGUMENT, "dt must be > 0");
    static_assert(kNumPasses == SPH_NUM_PASSES, "pass count");
    StepParams p = make_step_params(s->cfg, s->n);
    p.tune = s->tune;
    const bool accel = (s->debug & SPH_DEBUG_ACCEL) != 0u;
    hipEvent_t ev[kNumPasses + 1] = {};
    for (auto& e : ev) SPH_HIP(hipEventCreate(&e));
    int rc = SPH_OK;
    for (uint32_t step = 0; step < n_steps && rc == SPH_OK; ++step) {
        (void)hipEventRecord(ev[0], s->stream);
        for (int pass = 0; pass < kNumPasses && rc == SPH_OK; ++pass) {
            rc = launch_pass(pass, p, s->buf, dt, s->key_bits, accel, s->stream);
            (void)hipEventRecord(ev[pass + 1], s->stream);
        }
        if (rc != SPH_OK) break;
        hipError_t e = hipEventSynchronize(ev[kNumPasses]);
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
    if (rc == SPH_OK && n_steps > 0 && s->n > 0) s->stepped = true;
    return rc;
}

int sph_count_pairs(sph_solver* s, uint64_t counts[2]) {
    if (!s || !counts) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    counts[0] = counts[1] = 0;
    if (!s->stepped || s->n == 0) return SPH_OK;
    unsigned long long* d = nullptr;
    SPH_HIP(hipMalloc(&d, 2 * sizeof(unsigned long long)));
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemsetAsync(d, 0, sizeof(h), s->stream);
    int rc = (e == hipSuccess) ? SPH_OK : hip_error(e, "hipMemsetAsync", __FILE__, __LINE__);
    if (rc == SPH_OK) rc = launch_count_pairs(make_step_params(s->cfg, s->n), s->buf, d, s->stream);
    if (rc == SPH_OK) {
        e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
        if (e != hipSuccess) rc = hip_error(e, "count pairs", __FILE__, __LINE__);
    }
    (void)hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    if (rc != SPH_OK) return rc;
    counts[0] = h[0];
    counts[1] = h[1];
    return SPH_OK;
}

int sph_device_errors(sph_solver* s, uint32_t* bits) {
    if (!s || !bits) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    uint32_t* word = radix_sort_error_word(s->buf.sort_temp);
    uint32_t v = 0;
    SPH_HIP(hipMemcpyAsync(&v, word, sizeof(v), hipMemcpyDeviceToHost, s->stream));
    SPH_HIP(hipStreamSynchronize(s->stream));
    if (v) {
        SPH_HIP(hipMemsetAsync(word, 0, sizeof(v), s->stream));
        SPH_HIP(hipStreamSynchronize(s->stream));
    }
    *bits = public_error_bits(v);
    return SPH_OK;
}

int sph_set_debug(sph_solver* s, uint32_t flags) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    // Renderer colour values need the colour pass.
    if (s->render_flags & SPH_RENDER_COLORS) flags |= SPH_DEBUG_COLOR;
    if ((flags & SPH_DEBUG_COLOR) && !s->color) {
        SPH_HIP(hipMalloc(&s->color, 4ull * s->max_n));
        SPH_HIP(hipMemsetAsync(s->color, 0, 4ull * s->max_n, s->stream));
    }
    float* const color = (flags & SPH_DEBUG_COLOR) ? s->color : nullptr;
    if (color != s->buf.color) {
        SPH_HIP(hipStreamSynchronize(s->stream));
        destroy_graph(s);
        s->buf.color = color;
    }
    s->debug = flags;
    return SPH_OK;
}

int sph_download(sph_solver* s, int field, void* dst, size_t bytes) {
    if (!s || !dst) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    const size_t n = s->n, C = s->num_cells;
    const float4* vec = nullptr;
    const void* src = nullptr;
    size_t need = 0;
    switch (field) {
        case SPH_FIELD_POSITION: vec = s->buf.pos; need = 12 * n; break;
        case SPH_FIELD_VELOCITY: vec = s->buf.vel; need = 12 * n; break;
        case SPH_FIELD_ACCEL:
            if (!(s->debug & SPH_DEBUG_ACCEL))
                return set_error(SPH_ERR_INVALID_ARGUMENT, "accelerations need SPH_DEBUG_ACCEL");
            vec = s->buf.accel; need = 12 * n; break;
        case SPH_FIELD_DENSITY: src = s->buf.rho; need = 4 * n; break;
        case SPH_FIELD_ID: vec = s->buf.pos; need = 4 * n; break;
        case SPH_FIELD_SORTED_HASH: src = s->buf.skeys; need = 4 * n; break;
        case SPH_FIELD_SORTED_INDEX: src = s->buf.perm; need = 4 * n; break;
        case SPH_FIELD_CELL_START: src = s->buf.cell_start; need = 4 * C; break;
        case SPH_FIELD_CELL_END: src = s->buf.cell_end; need = 4 * C; break;
        case SPH_FIELD_COLOR:
            if (!s->buf.color)
                return set_error(SPH_ERR_INVALID_ARGUMENT, "colour values need SPH_DEBUG_COLOR");
            src = s->buf.color; need = 4 * n; break;
        default: return set_error(SPH_ERR_INVALID_ARGUMENT, "unknown field");
    }
    if (bytes < need) return set_error(SPH_ERR_INVALID_ARGUMENT, "destination too small");
    if (need == 0) return SPH_OK;
    if (src) {
        SPH_HIP(hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, s->stream));
        SPH_HIP(hipStreamSynchronize(s->stream));
        return SPH_OK;
    }
    std::vector<float> tmp(4 * n);
    SPH_HIP(hipMemcpyAsync(tmp.data(), vec, 16 * n, hipMemcpyDeviceToHost, s->stream));
    SPH_HIP(hipStreamSynchronize(s->stream));
    if (field == SPH_FIELD_ID) {
        auto* o = static_cast<uint32_t*>(dst);
        for (size_t i = 0; i < n; ++i) std::memcpy(&o[i], &tmp[4 * i + 3], 4);
    } else {
        auto* o = static_cast<float*>(dst);
        for (size_t i = 0; i < n; ++i)
            for (int d = 0; d < 3; ++d) o[3 * i + d] = tmp[4 * i + d];
    }
    return SPH_OK;
}

int sph_export_positions_aos(sph_solver* s, float* dst, int dst_is_device) {
    if (!s || !dst) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (!dst_is_device) return sph_download(s, SPH_FIELD_POSITION, dst, 12 * static_cast<size_t>(s->n));
    return launch_export_aos(s->buf.pos, s->n, dst, s->stream);
}

int sph_export_render(sph_solver* s, float* xyz, float* color) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    if (color && !s->buf.color)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "colour values need SPH_DEBUG_COLOR");
    return launch_export_render(s->buf.pos, nullptr, s->buf.color, s->buf.n_dev, s->max_n, 0u, xyz,
                                color, nullptr, s->stream);
}

int sph_enable_render_buffers(sph_solver* s, uint32_t flags) {
    if (!s) return set_error(SPH_ERR_INVALID_ARGUMENT, "solver is null");
    if (flags & ~SPH_RENDER_COLORS) return set_error(SPH_ERR_INVALID_ARGUMENT, "unknown render flag");
    SPH_HIP(hipStreamSynchronize(s->stream));
    if (!s->render_xyz) {
        // One allocation: xyz (12 B), colour (4 B) and state (4 B) per particle.
        const size_t n = s->max_n;
        char* base = nullptr;
        SPH_HIP(hipMalloc(&base, align_up(12 * n) + align_up(4 * n) + align_up(4 * n)));
        s->render_xyz = reinterpret_cast<float*>(base);
        s->render_color = reinterpret_cast<float*>(base + align_up(12 * n));
        s->render_state = reinterpret_cast<int32_t*>(base + align_up(12 * n) + align_up(4 * n));
        SPH_HIP(hipMemsetAsync(base, 0, align_up(12 * n) + align_up(4 * n) + align_up(4 * n),
                               s->stream));
    }
    s->render_flags = flags;
    if (flags & SPH_RENDER_COLORS) {
        const int rc = sph_set_debug(s, s->debug | SPH_DEBUG_COLOR);
        if (rc != SPH_OK) return rc;
    }
    destroy_graph(s);
    const int rc = export_render_buffers(s);
    if (rc != SPH_OK) return rc;
    SPH_HIP(hipStreamSynchronize(s->stream));
    return SPH_OK;
}

int sph_get_render_buffers(sph_solver* s, sph_render_buffers* out) {
    if (!s || !out) return set_error(SPH_ERR_INVALID_ARGUMENT, "null argument");
    if (!s->render_xyz)
        return set_error(SPH_ERR_INVALID_ARGUMENT, "render buffers are not enabled");
    out->positions = s->render_xyz;
    out->color_values = s->render_color;
    out->states = s->render_state;
    out->n = s->n;
    out->max_n = s->max_n;
    return SPH_OK;
}

size_t sph_radix_sort_temp_bytes(uint32_t n) { return radix_sort_temp_bytes(n); }

int sph_radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                         uint32_t* vals_out, uint32_t n, uint32_t key_bits, void* temp,
                         size_t temp_bytes, void* stream) {
    return radix_sort_pairs(keys_in, vals_in, keys_out, vals_out, n, key_bits, temp, temp_bytes,
                            static_cast<hipStream_t>(stream));
}

int sph_radix_sort_pairs_noalloc(const uint32_t* keys_in, const uint32_t* vals_in,
                                 uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                                 uint32_t key_bits, void* stream) {
    return radix_sort_pairs_static(keys_in, vals_in, keys_out, vals_out, n, key_bits,
                                   static_cast<hipStream_t>(stream));
}

}  // extern "C"
