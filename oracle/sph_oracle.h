/*
 * sph_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Host (CPU) restatement of the reference SPH2S solver step, used as the
 * parity checker of the HIP library and as the CPU baseline of the bench.
 * Nothing in the product (sph2s_final_amd/, include/) links or calls it.
 *
 * Precision: hashing and cell coordinates are computed in FP32 with the exact
 * IEEE operation order of the reference (subtract, add, divide, truncate,
 * clamp; no FMA contraction -- this file is compiled with -ffp-contract=off),
 * because the reference computes them in FP32
 * (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:20-83).  Everything else (density,
 * pressure, forces, integration, wall clamp) is FP64.
 *
 * Every function names the reference lines it restates.  How each one is
 * pinned independently of this file (closed forms, brute force, invariants)
 * is documented in tests/test_oracle.py.
 */
#ifndef SPH_ORACLE_H_
#define SPH_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_grid {
    float   origin[3];
    int32_t dims[3];
    float   spacing;
} oracle_grid;

typedef struct oracle_params {
    double  h, mass, rest_density, stiffness, viscosity, tension;
    double  gravity[3];
    double  restitution;
    double  lo[3], hi[3];
    int32_t wall_clamp;
    float   h_f32;          /* the FP32 h used for stencil offsets          */
    oracle_grid grid;
    double  xsph;           /* XSPH coefficient epsilon                     */
    int32_t use_xsph;       /* 1: the XSPH integrator (integrateXSPHD)      */
} oracle_params;

/* Cell hash of each particle (FP32 positions, AoS 3n).
 * SPH2S_FINAL_SINGLE_SCALE/Solver.cu:40-83 (computeCoordinates, computeHash). */
void oracle_hash(const oracle_grid* g, const float* pos, uint32_t n,
                 uint32_t* hash_out);

/* Stable sort of indices 0..n-1 by key (ties keep input order); the
 * semantics of thrust::sort_by_key(hash, ids) with ids = 0..n-1
 * (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1110-1115, ids from :416). */
void oracle_stable_sort(const uint32_t* keys, uint32_t n, uint32_t* keys_out,
                        uint32_t* perm_out);

/* Cell start/end from sorted keys; SPH_EMPTY (0xFFFFFFFF) for empty cells.
 * SPH2S_FINAL_SINGLE_SCALE/Solver.cu:478-534, 1118-1127. */
void oracle_cell_ranges(const uint32_t* sorted_keys, uint32_t n,
                        uint32_t num_cells, uint32_t* start, uint32_t* end);

/* One full FP64 step from an FP32 input state (AoS 3n) -- Solver::Advance,
 * SPH2S_FINAL_SINGLE_SCALE/Solver.cu:1082-1198 plus the wall clamp.
 * Outputs are in sorted order: hash_sorted[n], perm[n], cell_start/end[C],
 * rho[n], acc[3n] (FP64), pos_out/vel_out[3n] (FP64, after integration). */
void oracle_step(const oracle_params* p, const float* pos, const float* vel,
                 uint32_t n, double dt, uint32_t* hash_sorted, uint32_t* perm,
                 uint32_t* cell_start, uint32_t* cell_end, double* rho,
                 double* acc, double* pos_out, double* vel_out);

/* n_steps FP64 steps.  State is kept in FP64 and in sorted order, as the
 * reference rewrites its buffers in sorted order each step; ids_out[i] is the
 * input index of the particle now at slot i. */
void oracle_rollout(const oracle_params* p, const float* pos, const float* vel,
                    uint32_t n, double dt, uint32_t n_steps, double* pos_out,
                    double* vel_out, uint32_t* ids_out);

/* Boundary particles (static walls).  Each fluid particle's acceleration gains
 *   b_i = sum_{j : 0 < r_ij < h} B(r_ij) m_b / (m + m_b) x_ij / r_ij
 * over the wall particles found through the fluid grid's stencil
 * (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:311-344, 687-710; walls sorted once,
 * :1003-1059).  The reference divides by zero at r = 0; such pairs are
 * skipped. */
typedef struct oracle_walls {
    const float* pos;       /* AoS 3n FP32 wall positions */
    uint32_t n;
    double mass;            /* boundary particle mass m_b */
    double speed_sound;     /* c */
} oracle_walls;
/* B(r) of evaluateBoundaryWeight (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:169-196)
 * in FP64: 0.02 c^2 / r times 2/3 (q < 2/3), 2q - 3/2 q^2 (q < 1),
 * 1/2 (2 - q)^2 (q < 2), else 0, with q = 2 r / h. */
double oracle_boundary_weight(double r, double h, double speed_sound);
/* oracle_step / oracle_rollout with boundary particles (w may be NULL). */
void oracle_step_walls(const oracle_params* p, const oracle_walls* w, const float* pos,
                       const float* vel, uint32_t n, double dt, uint32_t* hash_sorted,
                       uint32_t* perm, uint32_t* cell_start, uint32_t* cell_end, double* rho,
                       double* acc, double* pos_out, double* vel_out);
void oracle_rollout_walls(const oracle_params* p, const oracle_walls* w, const float* pos,
                          const float* vel, uint32_t n, double dt, uint32_t n_steps,
                          double* pos_out, double* vel_out, uint32_t* ids_out);
/* oracle_rollout_walls that also returns the last step's densities (rho_out,
 * n f64 in the order of pos_out / ids_out: the last step's sorted order). */
void oracle_rollout_rho(const oracle_params* p, const oracle_walls* w, const float* pos,
                        const float* vel, uint32_t n, double dt, uint32_t n_steps,
                        double* pos_out, double* vel_out, uint32_t* ids_out, double* rho_out);

/* Colour value of every particle of a state (FP32 AoS pos/vel, n), by input
 * index: the Mexican-hat energy of SPH2S_FINAL_TWO_SCALE/Solver.cu:206-213,
 * 376-387, 897-907, 1050-1075 over the particle and its neighbours with
 * 0 < r < h (p gives h, mass and grid).  dist_out (may be NULL) receives
 * |x_i - x_c|, the distance the 0.0025 test uses. */
void oracle_energy_colors(const oracle_params* p, const float* pos, const float* vel, uint32_t n,
                          double* color_out, double* dist_out);
/* Colour-field gradient magnitude |n_i| h0 of the two-scale surface detector
 * for each low particle (inputs in any order; rho_* per particle, FP64):
 *   n_i = sum_{L, 0<r<h0} (m0/rho_j) grad W_h0 + sum_{H, 0<r<h0} (m1/rho_j)
 *         1/2 (grad W_h0 + grad W_h1)   (spiky gradients), all pairs. */
void oracle_surface_indicator(const oracle_params* p_low, const float* pos_l, const double* rho_l,
                              uint32_t n_l, const float* pos_h, const double* rho_h, uint32_t n_h,
                              double* mag_out);

/* O(n^2) density and acceleration over all pairs (no grid).  Pins neighbour
 * completeness of the grid path. */
void oracle_density_bruteforce(const oracle_params* p, const double* pos,
                               uint32_t n, double* rho);
void oracle_accel_bruteforce(const oracle_params* p, const double* pos,
                             const double* vel, const double* rho, uint32_t n,
                             double* acc);

/* Kernel functions (SPH2S_FINAL_SINGLE_SCALE/Solver.cu:103-167) in FP64. */
double oracle_poly6(double r, double h);
double oracle_spiky_grad_mag(double r, double h);   /* d/dr of spiky W */
double oracle_visc_laplacian(double r, double h);

/* ---- FP32 OpenMP CPU step (the CPU baseline the bench times) ---------- */
typedef struct cpu_solver cpu_solver;
cpu_solver* cpu_solver_create(const oracle_params* p, const float* pos,
                              const float* vel, uint32_t n);
void cpu_solver_step(cpu_solver* s, float dt, uint32_t n_steps);
void cpu_solver_get(const cpu_solver* s, float* pos_out, float* vel_out,
                    float* rho_out, uint32_t* ids_out);
int  cpu_solver_threads(void);
void cpu_solver_destroy(cpu_solver* s);

#ifdef __cplusplus
}
#endif

#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- two-scale step (FP64), SPH2S_FINAL_TWO_SCALE/Solver.cu:1568-1614 with
 * the split/merge and coupling rules documented in two_scale.hip ---------- */
typedef struct oracle_ts oracle_ts;
/* p_low: the low-resolution parameters; the high set uses h/2, m/8 and a grid
 * at spacing h/2 over the same domain.  max_low bounds low ids (child ids
 * start at max_low). */
oracle_ts* oracle_ts_create(const oracle_params* p_low, const oracle_grid* grid_high,
                            const float region_lo[3], const float region_hi[3],
                            uint32_t max_low);
void oracle_ts_destroy(oracle_ts* t);
void oracle_ts_set_state(oracle_ts* t, const float* pos_l, const float* vel_l,
                         const uint32_t* ids_l, uint32_t n_l, const float* pos_h,
                         const float* vel_h, const uint32_t* ids_h, const uint32_t* tags_h,
                         uint32_t n_h);
void oracle_ts_step(oracle_ts* t, double dt);
void oracle_ts_split_merge(oracle_ts* t);
/* counts[0..1] current (low, high); counts[2..3] the counts the last step ran with. */
void oracle_ts_counts(const oracle_ts* t, uint32_t counts[4]);
/* field: 0 pos (3n f64), 1 vel (3n f64), 2 id (n u32), 3 tag (n u32),
 * 4 rho (n f64, last step sorted order), 5 acc (3n f64), 6 sorted hash (n u32),
 * 7 perm (n u32), 8 cell start, 9 cell end (u32 per cell), 10 blend weight
 * lambda (n f32), 11 state word (n u32). */
void oracle_ts_get(const oracle_ts* t, int res, int field, void* out);

/* Coupling options.  coupling 0: blend weight 1 and instantaneous split/merge
 * (the rules above); 1: the reference's blended coupling -- per-particle
 * blend weights lambda on every neighbour term, the fade-in/fade-out state
 * machine of integrateD (TWO_SCALE/Solver.cu:1163-1262) and the faded merge.
 * split_offsets 0: children at +/- h1/4 per axis; 1: at r/sqrt(3) along the
 * diagonals, r = 1/2 (3 m0 / (4 pi rho))^(1/3) (Solver.cu:541-586).
 * merge 0 disables merging.  velocity_cap c > 0 limits |v| to c h_r / dt
 * after the velocity update (Solver.cu:1212-1223); 0 disables it. */
typedef struct oracle_ts_options {
    uint32_t coupling, split_offsets, merge;
    float blend_increment, velocity_cap;
} oracle_ts_options;
void oracle_ts_set_options(oracle_ts* t, const oracle_ts_options* o);
/* Replaces one set's blend weights and state words (n = its current count). */
void oracle_ts_set_blend(oracle_ts* t, int res, const float* lambda, const uint32_t* state);
/* Low particles whose id is listed and whose state is 0 get state 3 (split
 * requested), the reference's external mark of TWO_SCALE/Solver.cu:1077-1082. */
void oracle_ts_mark_split(oracle_ts* t, const uint32_t* ids, uint32_t n);
/* Boundary particles of the two-scale step: sorted on the low grid, and both
 * sets get the boundary acceleration with h0 and m0
 * (SPH2S_FINAL_TWO_SCALE/Solver.cu:454-499, 1007-1044).  w may be NULL. */
void oracle_ts_set_walls(oracle_ts* t, const oracle_walls* w);

/* Plants one deliberate mistake in the complementary (cross-resolution) sums
 * of the fixed coupling, so that tests can show their checks catch it:
 * 0 none; 1 the own set's mass in place of the other set's; 2 cutoff h1 in
 * place of h0; 3 the averaged kernels without the factor 1/2; 4 a stencil
 * reaching only +/- h1 (not +/- h0) into the other set's grid.  Process-wide;
 * tests reset it to 0. */
void oracle_ts_plant_mistake(int which);

#ifdef __cplusplus
}
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- differentiable rollout (FP64) ------------------------------------------
 * n_steps single-scale steps from (pos0, vel0) (AoS 3n FP64, vel0 may be NULL)
 * and the loss L = sum_{i,d} (x_T[i,d] - target[i,d])^2 / (3n), particles
 * matched by input index.  Returns L; x_final (3n, by input index) and the
 * gradients dL/dpos0, dL/dvel0 (3n), dL/dstiffness, dL/dviscosity are
 * written where non-NULL (all NULL: forward only).  Reverse-mode adjoint with
 * the pair topology frozen within each step; tension is not differentiated. */
double oracle_rollout_loss_grad(const oracle_params* p, const double* pos0,
                                const double* vel0, uint32_t n, double dt,
                                uint32_t n_steps, const double* target,
                                double* x_final, double* g_pos0, double* g_vel0,
                                double* g_stiffness, double* g_viscosity);

#ifdef __cplusplus
}
#endif
