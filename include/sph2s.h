/*
 * sph2s.h -- C ABI of the MI355X-native SPH2S solver step (libsph2s.so).
 *
 * The library replaces the device half of the reference solver
 * (SPH2S_FINAL_SINGLE_SCALE/Solver.cu, SPH2S_FINAL_TWO_SCALE/Solver.cu) and the
 * Thrust sort it calls.  Every entry point returns an int status (SPH_OK on
 * success); no C++ exception crosses this boundary.  sph_last_error() returns
 * the message of the last failure on the calling thread.  The C++ facade in
 * include/sph2s/Solver.hpp (ParticleData / Solver / SolverConfiguration /
 * Grid, the reference's class names) turns a non-zero status into a
 * std::runtime_error, as the reference's CUDA::SafeCall does
 * (SPH2S_FINAL_TWO_SCALE/cuda.h:115-129).
 *
 * Device pointers handed out by sph_get_views() are borrowed: they stay valid
 * until the next sph_advance, sph_set_state or sph_destroy on the same handle.
 * A handle may be driven by one host thread at a time; there is no
 * process-global state (the reference keeps one __constant__ configuration per
 * process, SPH2S_FINAL_SINGLE_SCALE/Solver.cu:12).
 */
#ifndef SPH2S_H_
#define SPH2S_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define SPH_OK                    0
#define SPH_ERR_INVALID_ARGUMENT  1
#define SPH_ERR_HIP               2
#define SPH_ERR_CAPACITY          3   /* a split would exceed max particles   */
#define SPH_ERR_NONFINITE         4   /* NaN/Inf found in the particle state   */
#define SPH_ERR_UNSUPPORTED       5
#define SPH_ERR_NO_DEVICE         6
#define SPH_ERR_DEVICE            7   /* a device-side check failed (see the
                                         *_device_errors bits)              */

/* Empty cell marker of cell_start/cell_end
 * (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:862,1118-1127). */
#define SPH_EMPTY_CELL 0xFFFFFFFFu

/* ---- Grid: replaces struct Grid (SPH2S_FINAL_SINGLE_SCALE/Grid.h:9-59) ---- */
typedef struct sph_grid {
    float   origin[3];
    int32_t dims[3];     /* number of cells per axis */
    float   spacing;     /* cell edge length         */
} sph_grid;

/* Grid::MakeGrid(origin, end, spacing): dims = ceil((end-origin)/spacing) in
 * FP32 per axis (SPH2S_FINAL_SINGLE_SCALE/Grid.h:16-35). */
int sph_make_grid(const float origin[3], const float end[3], float spacing,
                  sph_grid* out);
/* Grid::ComputeNumGridCells (SPH2S_FINAL_SINGLE_SCALE/Grid.h:11-14). */
uint32_t sph_grid_num_cells(const sph_grid* grid);

/* ---- configuration: replaces SolverConfiguration + __constant__ upload ---- */
#define SPH_FLAG_WALL_CLAMP  1u  /* clamp x to [domain_lo, domain_hi], v *= -e */
#define SPH_FLAG_KAHAN       2u  /* compensated density/force accumulation   */
/* XSPH integrator (integrateXSPHD, SPH2S_FINAL_SINGLE_SCALE/Solver.cu:346-395,
 * 757-853; the reference keeps it commented out, :1185-1196):
 *   v'_i  = v_i + dt a_i
 *   v''_i = v'_i + xsph * sum_{j : r_ij < h} (m / rho_j) W(r_ij) (v'_j - v'_i)
 *   x'_i  = x_i + dt v''_i, then the wall clamp if set.
 * The reference adds the x component of the correction to all three axes
 * (:837-839); here each axis gets its own.  Single-scale solver only. */
#define SPH_FLAG_XSPH        4u

typedef struct sph_config {
    float    h;              /* EffectiveRadius (Solver.h:51)            */
    float    mass;           /* FluidParticleMass                        */
    float    rest_density;   /* RestDensity                              */
    float    stiffness;      /* BulkModulus: p = k (rho - rho0)          */
    float    viscosity;      /* Viscosity (mu)                           */
    float    tension;        /* TensionCoefficient (sigma)               */
    float    gravity[3];     /* the reference hard-codes (0,-9.81,0)     */
    float    restitution;    /* wall clamp restitution e                 */
    float    domain_lo[3];   /* wall clamp box                           */
    float    domain_hi[3];
    sph_grid grid;           /* hashing grid (Solver.h:60)               */
    uint32_t flags;          /* SPH_FLAG_*                               */
    float    speed_sound;    /* SpeedSound: boundary-particle walls only */
    float    boundary_mass;  /* BoundaryParticleMass (Solver.h:38: = m)  */
    float    xsph;           /* XSPHCoeff epsilon (SPH_FLAG_XSPH)        */
} sph_config;

/* Fill a configuration from explicit physical constants; the grid is
 * MakeGrid(domain_lo, domain_hi, h); gravity (0,-9.81,0), tension 0,
 * restitution 0.5, flags WALL_CLAMP|KAHAN, speed of sound 88.1472 (the
 * reference scenes' value, SPH2S_FINAL_SINGLE_SCALE/main.cpp:243) and
 * boundary mass = mass. */
int sph_config_init(sph_config* out, float h, float mass, float rest_density,
                    float stiffness, float viscosity, const float domain_lo[3],
                    const float domain_hi[3]);

/* SolverConfiguration::MakeConfiguration (SPH2S_FINAL_SINGLE_SCALE/Solver.h:15-48):
 * h = (0.75 V avg / (pi N))^(1/3), mass = rho0 V / N, grid over [origin,end]
 * at spacing h, boundary mass = mass; speed_sound is kept for the
 * boundary-particle walls (sph_set_boundary).  The XSPH coefficient
 * (MakeConfiguration's epsilon) is set with sph_config_set_xsph. */
int sph_make_configuration(const float origin[3], const float end[3],
                           float volume, uint32_t num_particles,
                           uint32_t avg_particles, float rest_density,
                           float bulk_modulus, float viscosity,
                           float speed_sound, float tension, sph_config* out);
/* cfg->xsph = epsilon and SPH_FLAG_XSPH set (epsilon > 0) or cleared (0). */
int sph_config_set_xsph(sph_config* cfg, float epsilon);

/* ---- scene generation (replaces ParticleData::CreateParticleBox,
 *      SPH2S_FINAL_SINGLE_SCALE/ParticleData.cpp:141-176) ------------------- */
/* Lattice of n[0]*n[1]*n[2] particles at lo + (i+1/2)*spacing, x fastest,
 * with SplitMix64(seed) jitter U[-jitter_frac, jitter_frac)*spacing per axis.
 * pos_xyz receives 3*n0*n1*n2 floats (AoS). */
int sph_scene_lattice(const int32_t n[3], float spacing, const float lo[3],
                      float jitter_frac, uint64_t seed, float* pos_xyz);
/* n velocities with components N(0, sigma^2) (SplitMix64 + Box-Muller). */
int sph_scene_gaussian_velocity(uint32_t n, float sigma, uint64_t seed,
                                float* vel_xyz);

/* ---- the single-scale solver --------------------------------------------- */
typedef struct sph_solver sph_solver;

/* Solver(fluid, boundary, cfg) (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1003-1061)
 * constructed from host particle arrays: pos_xyz/vel_xyz are AoS 3*n floats
 * (vel_xyz may be NULL = at rest).  Particle i gets persistent id i. */
int sph_create(const sph_config* cfg, const float* pos_xyz,
               const float* vel_xyz, uint32_t n, uint32_t max_n,
               sph_solver** out);
int sph_destroy(sph_solver* s);

/* Solver::Advance(dt) (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1082-1096), run
 * n_steps times: hash -> radix sort -> reorder + cell ranges -> density ->
 * force + symplectic Euler + wall clamp.  The steps are replayed from a
 * captured hipGraph on the solver's own stream; the call returns without
 * waiting for the GPU. */
int sph_advance(sph_solver* s, float dt, uint32_t n_steps);
/* Same, launching each kernel directly instead of replaying the graph. */
int sph_advance_eager(sph_solver* s, float dt, uint32_t n_steps);
/* Waits for the solver's stream, then reports device-side failures of the
 * steps since the bits were last cleared (sph_device_errors):
 * SPH_ERR_NONFINITE (SPH_DEVERR_NONFINITE), SPH_ERR_CAPACITY (two-scale,
 * SPH_DEVERR_CAPACITY) or SPH_ERR_DEVICE (SPH_DEVERR_SORT_WAIT).  The bits
 * stay set, so every later synchronize reports them until they are cleared. */
int sph_synchronize(sph_solver* s);

/* Replace the particle state from host arrays (ids restart at 0..n-1). */
int sph_set_state(sph_solver* s, const float* pos_xyz, const float* vel_xyz,
                  uint32_t n);

/* Boundary particles: the boundaryData of Solver(fluid, boundary, cfg).
 * The reference hashes and sorts these static wall particles once on the
 * fluid grid in its constructor (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:
 * 1003-1059).  From then on, every step adds each fluid particle's boundary
 * acceleration to its acceleration (a_i = f_i / rho_i + b_i + g, :687-710):
 *   b_i = sum_{j : r_ij < h} B(r_ij) m_b / (m + m_b) x_ij / r_ij
 *   B(r) = 0.02 c^2 / r * (2/3 if q < 2/3; 2q - 3/2 q^2 if q < 1;
 *          1/2 (2 - q)^2 if q < 2),  q = 2 r / h
 * (evaluateBoundaryWeight, computeBoundaryForceCell, :169-196, :311-344).
 * Here c = cfg.speed_sound and m_b = cfg.boundary_mass.  The reference divides
 * by zero at r = 0; here such a pair is skipped.  pos_xyz holds AoS 3n
 * floats.  n = 0 removes the walls.  The wall clamp of cfg.flags still
 * applies when it is set. */
int sph_set_boundary(sph_solver* s, const float* pos_xyz, uint32_t n);
/* ParticleData::NumParticles. */
int sph_num_particles(const sph_solver* s, uint32_t* n);
int sph_get_config(const sph_solver* s, sph_config* out);

/* Borrowed device views of the state, in the sorted order of the last step
 * (the reference rewrites dPositions in sorted order each step,
 * SPH2S_FINAL_SINGLE_SCALE/Solver.cu:748-754). */
typedef struct sph_views {
    void*     pos4;          /* float4[n]: x, y, z, w = persistent id bits */
    void*     vel4;          /* float4[n]: vx, vy, vz, w = 0              */
    float*    density;       /* float[n] rho of the last step             */
    uint32_t* sorted_hash;   /* uint[n] cell hash, ascending              */
    uint32_t* sorted_index;  /* uint[n] stable permutation of the last sort */
    uint32_t* cell_start;    /* uint[num_cells], SPH_EMPTY_CELL if empty  */
    uint32_t* cell_end;      /* uint[num_cells], SPH_EMPTY_CELL if empty  */
    void*     accel4;        /* float4[n] accelerations (SPH_DEBUG_ACCEL) or NULL */
    uint32_t  n;
    uint32_t  num_cells;
    void*     stream;        /* hipStream_t the solver launches on        */
} sph_views;
int sph_get_views(sph_solver* s, sph_views* out);

/* Host copies of one field (synchronises the solver's stream).  bytes must be
 * at least the field's size: POSITION/VELOCITY/ACCEL 12*n (AoS xyz),
 * DENSITY/ID/SORTED_* 4*n, CELL_* 4*num_cells. */
#define SPH_FIELD_POSITION      0
#define SPH_FIELD_VELOCITY      1
#define SPH_FIELD_DENSITY       2
#define SPH_FIELD_ID            3
#define SPH_FIELD_SORTED_HASH   4
#define SPH_FIELD_SORTED_INDEX  5
#define SPH_FIELD_CELL_START    6
#define SPH_FIELD_CELL_END      7
#define SPH_FIELD_ACCEL         8
#define SPH_FIELD_TAG           9   /* two-scale high set: (slot << 29) | group */
#define SPH_FIELD_BLEND        10   /* two-scale, blended coupling: lambda (float) */
#define SPH_FIELD_STATE        11   /* two-scale, blended coupling: state word (u32) */
#define SPH_FIELD_COLOR        12   /* colour value (float), SPH_DEBUG_COLOR        */
int sph_download(sph_solver* s, int field, void* host_dst, size_t bytes);

/* Renderer contract: positions as 3 floats per particle
 * (SPH2S_FINAL_TWO_SCALE/Renderer.cpp:40-42).  dst is a host pointer, or a
 * device pointer when dst_is_device != 0 (then the copy is stream-ordered). */
int sph_export_positions_aos(sph_solver* s, float* dst, int dst_is_device);
/* The renderer's buffers (ParticleData::dPositions, dColorValues,
 * SPH2S_FINAL_TWO_SCALE/ParticleData.h:39-45, drawn by Renderer.cpp:38-46),
 * copied into caller-owned device memory in the state's order: xyz gets 3n
 * floats, color n floats (needs SPH_DEBUG_COLOR).  Either may be NULL.
 * Stream-ordered on the solver's stream (sph_synchronize before reading). */
int sph_export_render(sph_solver* s, float* xyz, float* color);

/* Renderer buffers the solver keeps current: the device buffers of
 * ParticleData that the reference's renderer binds and its solver rewrites
 * every step -- dPositions (AoS, 3 floats per particle), dColorValues (1 float)
 * and, two-scale, dStates (1 int) (SPH2S_FINAL_TWO_SCALE/ParticleData.h:39-45,
 * Renderer.cpp:38-46,71, SSFRenderer.cu:438-439; single scale
 * SPH2S_FINAL_SINGLE_SCALE/ParticleData.h:39-41).  Once enabled, the solver
 * owns them (capacity max_n each) and every step rewrites them in the state's
 * order at its end, inside the captured step graph; sph_set_state refreshes
 * them too.  With SPH_RENDER_COLORS the colour pass (SPH_DEBUG_COLOR) runs
 * every step and fills color_values; without it they stay 0, as the
 * single-scale reference never writes them.  Single-scale states are 0.
 * The pointers stay valid until the handle is destroyed. */
#define SPH_RENDER_COLORS 1u
typedef struct sph_render_buffers {
    float*   positions;     /* float[3 max_n]: x, y, z per particle (dPositions) */
    float*   color_values;  /* float[max_n]   (dColorValues)                     */
    int32_t* states;        /* int[max_n]     (dStates)                          */
    uint32_t n;             /* NumParticles: entries written                     */
    uint32_t max_n;         /* MaxParticles                                      */
} sph_render_buffers;
int sph_enable_render_buffers(sph_solver* s, uint32_t flags);
int sph_get_render_buffers(sph_solver* s, sph_render_buffers* out);

#define SPH_DEBUG_ACCEL 1u   /* keep per-particle accelerations in accel4 */
/* Compute each particle's colour value every step, as the two-scale
 * reference's force pass does for dColorValues (SPH2S_FINAL_TWO_SCALE/
 * Solver.cu:206-213, 376-387, 897-907, 1050-1075): over the particle and its
 * own-set neighbours with 0 < r < h, psi = C (|x_ij/h|^2 - 3) exp(-|x_ij/h|^2/2)
 * (C = 2 / (sqrt(3) pi^(1/4))), velW = sum psi v_j, psiSum = sum psi, x_c =
 * mean position; ene = min(|velW|^2 / (psiSum^2 h), 300); colour = 0.3 where
 * |x_i - x_c| < 0.0025, else ene / 300.  Read it with SPH_FIELD_COLOR or the
 * render export. */
#define SPH_DEBUG_COLOR 2u
int sph_set_debug(sph_solver* s, uint32_t flags);

/* Per-pass timing: runs n_steps steps eagerly with hipEvents recorded on the
 * solver's stream between passes and adds each pass's milliseconds into
 * ms_out[0..SPH_NUM_PASSES-1] (hash+cell reset, radix sort, reorder+cell
 * ranges, density, force+integrate).  Synchronises after every step. */
#define SPH_NUM_PASSES 5
int sph_profile_passes(sph_solver* s, float dt, uint32_t n_steps, float* ms_out);

/* Device-side error bits raised since the solver was created or last
 * queried (then cleared); synchronises the solver's stream.
 * SPH_DEVERR_SORT_WAIT: a radix-sort look-back wait exceeded its bound and
 * gave up, so the step that raised it has an invalid ordering. */
#define SPH_DEVERR_SORT_WAIT 1u
#define SPH_DEVERR_CAPACITY  2u   /* two-scale: a split/merge exceeded a capacity */
/* SPH_DEVERR_NONFINITE: a step's integration wrote a NaN or infinite position
 * or velocity (the force pass checks every particle it integrates). */
#define SPH_DEVERR_NONFINITE 4u
int sph_device_errors(sph_solver* s, uint32_t* bits);

/* Pair work of the most recent step (no counterpart in the reference; a
 * diagnostic for the FLOP model of the neighbour passes): counts[0] = the
 * candidates one neighbour pass walks (every particle's stencil-row runs),
 * counts[1] = the pairs among them with 0 < r < h.  Both are 0 when no step
 * has run since the state was set.  Synchronises the solver's stream. */
int sph_count_pairs(sph_solver* s, uint64_t counts[2]);

/* ---- launch variants and the autotuner -------------------------------------
 * The reference launches fixed block shapes (Solver.cu:1082-1198).  Here the
 * density and force passes have launch variants, selected by a tune word:
 *   bits 0-3   density LDS tile: 0 = 1280, 1 = 512, 2 = 1536, 3 = 2048, 4 = 768,
 *              5 = 1024 particles
 *   bits 4-7   force LDS tile / neighbour-list capacity / entry width:
 *              0 = 640/16/16-bit, 1 = 768/16/16, 2 = 768/24/16, 3 = 768/8/32,
 *              4 = 512/16/16, 5 = 1024/16/16, 6 = 512/8/32, 7 = 640/24/16,
 *              8 = 1024/8/32, 9 = 512/24/16
 *   bits 8-11  density grid, bits 12-15 force grid: 0 = one block per
 *              block-size particles; 1..4 = persistent grid of 1, 2, 4, 8
 *              blocks per CU looping over those blocks
 *   bits 16-17 density block size, bits 18-19 force block size: 0 = 256,
 *              1 = 128, 2 = 512 threads; a non-zero block size requires that
 *              pass's tile/list field to be 0
 *   bits 20-21 density, bits 22-23 force: particles per lane, 0 = 1, 1 = 2
 *              (particles per wavefront 64 or 128; a block's targets share
 *              one staged union per row); 2 requires that pass's tile/list
 *              and block-size fields to be 0, and a batch refuses it
 *   bits 24-31 must be zero; other words are refused with SPH_ERR_ARG
 * All variants give bit-identical results.  sph_autotune times the variants
 * on the solver's current state (reps launches each), keeps the fastest and
 * returns its tune word; ms[4] (may be NULL) receives the default and tuned
 * milliseconds of density and force.  The state is unchanged; the solver's
 * intermediate fields (density, sorted arrays) then hold the tuning pass's
 * values. */
#define SPH_TUNE_DEFAULT 0u
int sph_autotune(sph_solver* s, float dt, uint32_t reps, uint32_t* tune, float* ms);
int sph_set_tune(sph_solver* s, uint32_t tune);
int sph_get_tune(const sph_solver* s, uint32_t* tune);

/* ---- the two-scale solver (SPH2S_FINAL_TWO_SCALE/Solver.h:109-144) ---------
 * A low-resolution set (h0, m0 from cfg) and a high-resolution set (h1 = h0/2,
 * m1 = m0/8, TWO_SCALE/Solver.h:36-43), each with its own grid over the
 * domain, coupled through the complementary averaged-kernel density and force
 * sums of TWO_SCALE/Solver.cu:252-290,394-499 (blend weight 1).  Every step,
 * low-resolution particles inside [region_lo, region_hi] split into 8 children
 * (mass m1, x +/- h1/4 per axis, velocity copied; TWO_SCALE/Solver.cu:541-586
 * with the benchmark's offsets), and groups of 8 siblings that are all outside
 * the region expanded by h0/2 merge back into one particle (centroid, mean
 * velocity).  Counts stay on the device; the whole step replays as a hipGraph.
 * Ids: low-resolution particles keep ids < max_low (a merged particle takes its
 * group's id); child k of particle p has id max_low + 8 p + k, group p, slot k. */
typedef struct sph_two_scale sph_two_scale;
int sph2_create(const sph_config* cfg, const float region_lo[3],
                const float region_hi[3], const float* pos_xyz,
                const float* vel_xyz, uint32_t n_low, uint32_t max_low,
                uint32_t max_high, sph_two_scale** out);
int sph2_destroy(sph_two_scale* t);
/* Replace both sets' states.  High-resolution particles need ids and tags
 * ((slot << 29) | group); ids_l may be NULL (0..n_l-1). */
int sph2_set_state(sph_two_scale* t, const float* pos_l, const float* vel_l,
                   const uint32_t* ids_l, uint32_t n_l, const float* pos_h,
                   const float* vel_h, const uint32_t* ids_h,
                   const uint32_t* tags_h, uint32_t n_h);
/* Solver::Advance(dt) of TWO_SCALE/Solver.cu:1568-1614, n_steps times. */
int sph2_advance(sph_two_scale* t, float dt, uint32_t n_steps);
int sph2_advance_eager(sph_two_scale* t, float dt, uint32_t n_steps);
/* Runs only the split/merge/compaction stage on the current state. */
int sph2_apply_split_merge(sph_two_scale* t);
/* Boundary particles of the two-scale solver: sorted once on the
 * low-resolution grid.  Both sets get the boundary acceleration of
 * sph_set_boundary with h0 and m0: the reference passes LOW_RES for both
 * (SPH2S_FINAL_TWO_SCALE/Solver.cu:454-499, :1007-1044). */
int sph2_set_boundary(sph_two_scale* t, const float* pos_xyz, uint32_t n);
/* Waits and reports device-side failures like sph_synchronize (both sets'
 * bits and the capacity bit, SPH_ERR_CAPACITY). */
int sph2_synchronize(sph_two_scale* t);
int sph2_set_debug(sph_two_scale* t, uint32_t flags);
/* counts[0..1] = current (low, high) counts; counts[2..3] = the counts the
 * last step ran with.  Synchronises. */
int sph2_counts(sph_two_scale* t, uint32_t counts[4]);
int sph2_device_errors(sph_two_scale* t, uint32_t* bits);
/* Host copy of one field of resolution 0 (low) or 1 (high).  POSITION,
 * VELOCITY, ID, TAG, BLEND, STATE: current (post split/merge) state; DENSITY,
 * ACCEL, SORTED_*: the last step's sorted order; CELL_*: the last step's grid.
 * BLEND and STATE exist under SPH2_COUPLING_BLENDED only. */
int sph2_download(sph_two_scale* t, int resolution, int field, void* host_dst,
                  size_t bytes);

/* Coupling options of the two-scale solver.
 *
 * coupling SPH2_COUPLING_FIXED (the default): blend weight 1 everywhere,
 * instantaneous split (parent removed) and merge (children removed), as
 * described above.
 *
 * coupling SPH2_COUPLING_BLENDED: the reference's own two-scale coupling.
 * Every particle carries a blend weight lambda (SPH2S_FINAL_TWO_SCALE/
 * Solver.h:89-91) and a state word (state & 3: 0 settled, 1 fading out,
 * 2 fading in, 3 split requested; state & 4: high set).  Every neighbour term
 * of the own-set and complementary density and force sums is weighted by the
 * neighbour's lambda (Solver.cu:215-499); after the integration lambda moves
 * by -/+ blend_increment in states 1/2, a fading-out particle is removed when
 * lambda <= 0 and a fading-in one settles (state & 4, lambda = 1) when
 * lambda >= 1 (integrateD, Solver.cu:1240-1289).  A settled low particle
 * inside the region, or one in state 3 (sph2_mark_split), splits into 8
 * children with lambda 0 in state 6 and fades out (state 1) instead of being
 * removed.  Merging is the reverse: a group of 8 settled siblings outside the
 * region expanded by h0/2 emits a low particle at their centroid with their
 * mean velocity, lambda 0, state 2, and the siblings fade out.  (The
 * reference has no merge.)  A split waits while a child of the same group is
 * alive, a merge while a low particle with the group's id is alive, so ids
 * stay unique.
 *
 * split_offsets SPH2_SPLIT_LATTICE: children at x +/- h1/4 per axis;
 * SPH2_SPLIT_REFERENCE: at x + r d_k with d_k the 8 diagonals (+/-1, +/-1,
 * +/-1)/sqrt(3) and r = 1/2 (3 m0 / (4 pi rho))^(1/3), rho the parent's
 * density of the step (insertHighResParticles, Solver.cu:541-586; the
 * standalone sph2_apply_split_merge uses the rest density).  Child k takes the
 * sign + on axis d when bit d of k is set.
 * merge: 0 = never merge (the reference), 1 = merge as above.
 * velocity_cap c > 0: after v += dt a, |v| is limited to c h_r / dt with h_r
 * the particle's own radius (Solver.cu:1212-1223, c = 0.8 there; |v| = 0 is
 * left alone rather than divided by); 0 disables the cap.
 * blend_increment: lambda's change per step (SolverConfiguration::
 * BlendIncrement, Solver.h:27,48,60; 0.0125 in SPH2S_FINAL_TWO_SCALE/
 * main.cpp:228-240), in (0, 1].
 * Switching to the blended coupling sets every particle to lambda = 1,
 * settled. */
#define SPH2_COUPLING_FIXED    0u
#define SPH2_COUPLING_BLENDED  1u
#define SPH2_SPLIT_LATTICE     0u
#define SPH2_SPLIT_REFERENCE   1u
typedef struct sph2_options {
    uint32_t coupling;         /* SPH2_COUPLING_*                          */
    uint32_t split_offsets;    /* SPH2_SPLIT_*                             */
    uint32_t merge;            /* 0 or 1                                   */
    float    blend_increment;  /* lambda change per step (blended)         */
    float    velocity_cap;     /* c of |v| <= c h / dt; 0 = off            */
} sph2_options;
int sph2_set_options(sph_two_scale* t, const sph2_options* options);
int sph2_get_options(const sph_two_scale* t, sph2_options* options);
/* Launch variants of the two-scale own-set passes (no counterpart in the
 * reference, like sph_autotune): each resolution (0 = low, 1 = high) has a
 * tune word with the fields of sph_set_tune's, applied to its own-set density
 * and force passes (the tension form of the force runs the default); every
 * word gives bit-identical results.  sph2_autotune times the variants of both
 * sets' own-set passes on the current state (reps launches each), keeps the
 * fastest per set and returns the two words; ms[8] (may be NULL) receives,
 * per set (low first), the default and tuned milliseconds of density and
 * force.  The state is unchanged; intermediate fields hold the tuning pass's
 * values until the next step. */
int sph2_set_tune(sph_two_scale* t, int resolution, uint32_t tune);
int sph2_get_tune(const sph_two_scale* t, int resolution, uint32_t* tune);
int sph2_autotune(sph_two_scale* t, uint32_t reps, uint32_t tune[2], float* ms);
/* Replace one set's blend weights and state words (blended coupling; n must
 * equal the set's current count; either array may be NULL to keep it). */
int sph2_set_blend(sph_two_scale* t, int resolution, const float* lambda,
                   const uint32_t* state, uint32_t n);
/* Request the split of the listed low-resolution particles (by id): settled
 * ones (mode 0, surface bit 8 allowed) get mode 3 and split at the end of the
 * next step.  The reference sets
 * this state from its screen-space surface pass (SSFRenderer.cu:196-218 and
 * the mark in Solver.cu:1077-1082).  Blended coupling only; n <= max_low. */
int sph2_mark_split(sph_two_scale* t, const uint32_t* ids, uint32_t n);

/* Renderer buffers of one set (ParticleData::dPositions, dColorValues,
 * dStates of SPH2S_FINAL_TWO_SCALE/ParticleData.h:39-45) in the current
 * (compacted) order, copied into caller-owned device memory: xyz 3n floats,
 * color n floats (SPH_DEBUG_COLOR; children carry their parent's value from
 * the step that made them, a merged particle its group's slot-0 value),
 * state n int32 (the blended coupling's state words; 0 low / 4 high under the
 * fixed coupling).  Any pointer may be NULL.  Stream-ordered on the solver's
 * stream (sph2_synchronize before reading). */
int sph2_export_render(sph_two_scale* t, int resolution, float* xyz, float* color,
                       int32_t* state);
/* Solver-owned renderer buffers of both sets, rewritten at the end of every
 * step (inside the step graph), by sph2_set_state and by
 * sph2_apply_split_merge, in the compacted order that sph2_export_render
 * copies (see sph_enable_render_buffers; the two-scale reference writes
 * dColorValues and dStates every step, TWO_SCALE/Solver.cu:1050-1083,
 * 1240-1289).  sph2_get_render_buffers reads the set's count back
 * (synchronises); capacity max_low or max_high. */
int sph2_enable_render_buffers(sph_two_scale* t, uint32_t flags);
int sph2_get_render_buffers(sph_two_scale* t, int resolution, sph_render_buffers* out);

/* Surface marks (blended coupling).  The reference's screen-space pass sets
 * state bit 8 on the low particles its surface index map shows
 * (getSurfaceIDs, SPH2S_FINAL_TWO_SCALE/SSFRenderer.cu:196-218) and, in a
 * line it leaves commented out, requests the split of those in state 8
 * (Solver.cu:1077-1082).
 * sph2_mark_surface_map does exactly that from a device index map of count
 * entries (slot + 1 of the current low set, 0 = background): bit 8, and with
 * split != 0 mode 3 for particles whose state is 8.
 * sph2_detect_surface is its view-independent replacement: from the last
 * step's sorted state, a low particle whose colour-field gradient
 *   n_i = sum_{L, 0<r<h0} (m0/rho_j) grad W_h0 + sum_{H, 0<r<h0} (m1/rho_j)
 *         1/2 (grad W_h0 + grad W_h1)     (spiky gradients)
 * has |n_i| h0 > threshold is at the free surface; it gets the same marks.
 * Both synchronise. */
int sph2_mark_surface_map(sph_two_scale* t, const uint32_t* map_dev, uint32_t count, int split);
int sph2_detect_surface(sph_two_scale* t, float threshold, int split);

/* ---- batched independent scenes --------------------------------------------
 * The reference runs one scene per process through a global __constant__
 * configuration (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:12,1068-1080).  A batch
 * advances n_scenes independent single-scale scenes, each with its own
 * sph_config (grid, stiffness, viscosity, domain), in one segmented hash/sort
 * and one launch per pass; the step replays as a hipGraph.  Scene s owns the
 * particle range [part_offsets[s], part_offsets[s+1]) and the global cell
 * range [cell_offsets[s], cell_offsets[s+1]) of the batch's arrays; every
 * scene's results equal those of an sph_solver running that scene alone.
 * pos/vel: host AoS 3 * sum(counts), scene-major; vel may be NULL. */
typedef struct sph_batch sph_batch;
int sph_batch_create(const sph_config* cfgs, uint32_t n_scenes,
                     const uint32_t* counts, const float* pos_xyz,
                     const float* vel_xyz, sph_batch** out);
int sph_batch_destroy(sph_batch* b);
int sph_batch_advance(sph_batch* b, float dt, uint32_t n_steps);
int sph_batch_advance_eager(sph_batch* b, float dt, uint32_t n_steps);
/* Waits and reports device-side failures like sph_synchronize. */
int sph_batch_synchronize(sph_batch* b);
int sph_batch_set_debug(sph_batch* b, uint32_t flags);
/* part_offsets and cell_offsets (n_scenes + 1 entries each) may be NULL. */
int sph_batch_info(const sph_batch* b, uint32_t* n_scenes, uint32_t* n_particles,
                   uint32_t* n_cells, uint32_t* part_offsets, uint32_t* cell_offsets);
int sph_batch_device_errors(sph_batch* b, uint32_t* bits);
/* Adds the milliseconds of each pass over n_steps eager steps to ms[5]. */
int sph_batch_profile_passes(sph_batch* b, float dt, uint32_t n_steps, float* ms);
/* Host copy of one field over the whole batch (fields as sph_download; TAG
 * returns each particle's scene index; CELL_* cover all scenes' cells). */
int sph_batch_download(sph_batch* b, int field, void* host_dst, size_t bytes);
/* The autotuner and tune words of sph_autotune, for a batch. */
int sph_batch_autotune(sph_batch* b, float dt, uint32_t reps, uint32_t* tune, float* ms);
int sph_batch_set_tune(sph_batch* b, uint32_t tune);
int sph_batch_get_tune(const sph_batch* b, uint32_t* tune);

/* ---- differentiable batched rollout -----------------------------------------
 * The reference has no differentiation; this differentiates n_steps of the
 * batched step (Solver::Advance, SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1082-1198,
 * with the wall clamp) by hand-written adjoint kernels with the neighbour
 * topology frozen within each step.  The state is checkpointed every
 * checkpoint_every steps; backward recomputes one segment at a time.
 * Arrays are device pointers, AoS 3 * n in global particle order (scene-major,
 * as sph_batch_create's input).  Tension must be 0.
 * All calls are ordered after the work already on `stream` (a hipStream_t,
 * NULL = the null stream) and `stream` is ordered after them; they return
 * before the work completes unless host outputs are requested. */
typedef struct sph_rollout sph_rollout;
int sph_rollout_create(const sph_config* cfgs, uint32_t n_scenes,
                       const uint32_t* counts, float dt, uint32_t n_steps,
                       uint32_t checkpoint_every, sph_rollout** out);
int sph_rollout_destroy(sph_rollout* r);
int sph_rollout_info(const sph_rollout* r, uint32_t* n_particles,
                     uint32_t* n_scenes, size_t* device_bytes);
/* Runs the rollout from (pos0, vel0) (vel0 may be NULL) with per-scene
 * stiffness and viscosity (host arrays of n_scenes; NULL keeps the current
 * values) and writes the final positions by particle to final_pos (may be
 * NULL).  Keeps the checkpoints for one sph_rollout_backward. */
int sph_rollout_forward(sph_rollout* r, const float* pos0, const float* vel0,
                        const float* stiffness, const float* viscosity,
                        float* final_pos, void* stream);
/* Given dL/d(final_pos), writes dL/dpos0 and dL/dvel0 (device, may be NULL)
 * and dL/dstiffness, dL/dviscosity per scene (host doubles, may be NULL;
 * requesting them synchronises).  Consumes the forward's state. */
int sph_rollout_backward(sph_rollout* r, const float* grad_final,
                         float* grad_pos0, float* grad_vel0,
                         double* grad_stiffness, double* grad_viscosity,
                         void* stream);
int sph_rollout_synchronize(sph_rollout* r);
int sph_rollout_device_errors(sph_rollout* r, uint32_t* bits);

/* ---- standalone stable LSD radix sort (replaces thrust::sort_by_key,
 *      SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1110-1115) ----------------------- */
/* Temporary device bytes needed to sort n pairs. */
size_t sph_radix_sort_temp_bytes(uint32_t n);
/* Stable sort of (key, value) pairs by the low key_bits bits of the key
 * (1..32).  vals_in == NULL means values 0..n-1.  All pointers are device
 * pointers; keys_out/vals_out must not alias the inputs.  stream is a
 * hipStream_t (NULL = default stream).  Asynchronous. */
int sph_radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in,
                         uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                         uint32_t key_bits, void* temp, size_t temp_bytes,
                         void* stream);
/* The same sort without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  The temporary buffer
 * is the library's own static device array (24 MiB, one per GPU, held until
 * the process ends), so n is limited to SPH_RADIX_SORT_NOALLOC_MAX_KEYS pairs;
 * a larger n returns SPH_ERR_CAPACITY before any device work.  The first call
 * on a device takes the array's address there; no call allocates or frees.
 * The array keeps the last call's contents; each call resets the sort state
 * it needs with a kernel (no memset), so it is safe to capture.  Calls that
 * overlap on one GPU (other streams, other threads) share the array and
 * race: order them on one stream.  Otherwise as sph_radix_sort_pairs. */
#define SPH_RADIX_SORT_NOALLOC_MAX_KEYS (1u << 21)
int sph_radix_sort_pairs_noalloc(const uint32_t* keys_in, const uint32_t* vals_in,
                                 uint32_t* keys_out, uint32_t* vals_out, uint32_t n,
                                 uint32_t key_bits, void* stream);

/* ---- diagnostics ----------------------------------------------------------- */
const char* sph_last_error(void);
/* Number of visible HIP devices (0 on a host without a GPU; never fails). */
int sph_device_count(void);
const char* sph_version(void);

#ifdef __cplusplus
}
#endif

#endif /* SPH2S_H_ */
