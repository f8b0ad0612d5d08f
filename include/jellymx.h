/* jellymx — MI355X-native batched 2D MLS-MPM engine: C ABI.
 *
 * This is the drop-in boundary for the simulation core of c-m-x-r/jellyfih (`mpm_sim.py`).
 * The reference exposes that core as a Python module surface (Taichi fields and kernels);
 * each entry point below names the reference item it replaces (file:line). A ctypes
 * binding of these functions (jellyfih_amd/_lib.py) backs a Python module with the
 * reference's names (jellyfih_amd/mpm_sim.py).
 *
 * Conventions
 *  - All functions return a jm_status; on failure jm_last_error() returns a thread-local
 *    message. Handles are not thread-safe; use one handle per thread.
 *  - Host arrays use the reference's layout: per-particle fields are (n_instances,
 *    n_particles, ...) in slot order, grids are (n_instances, n_grid, n_grid_y, ...).
 *  - Device memory is owned by the library (hipMalloc at jm_create). All work is queued on
 *    the handle's own HIP stream (jm_stream); functions that return host data synchronise it.
 *  - Arithmetic is fp32, as in the reference.
 */
#ifndef JELLYMX_H
#define JELLYMX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JM_ABI_VERSION 4

typedef enum jm_status {
  JM_OK = 0,
  JM_ERR_ARG = 1,    /* invalid argument (size, index, field id, NaN where forbidden) */
  JM_ERR_HIP = 2,    /* a HIP runtime call failed (message names it) */
  JM_ERR_STATE = 3,  /* call not valid in the handle's current state */
  JM_ERR_NOMEM = 4   /* device allocation failed */
} jm_status;

/* Step paths.
 * JM_PATH_FUSED: cell-sorted particles, one fused G2P->P2G kernel per substep with
 *   LDS-tiled gather/scatter and triple-buffered grids, blocks captured as HIP graphs.
 * JM_PATH_PLAIN: unsorted, one launch per stage (reset, P2G with global atomics, grid op,
 *   G2P) per instance — the straightforward form kept as the comparison baseline. */
typedef enum jm_path { JM_PATH_FUSED = 0, JM_PATH_PLAIN = 1 } jm_path;

/* Fields readable with jm_read_field (and, where noted, writable with jm_write_field). */
typedef enum jm_field {
  JM_FIELD_X = 0,        /* f32 (M,N,2)   mpm_sim.py:86   rw */
  JM_FIELD_V = 1,        /* f32 (M,N,2)   mpm_sim.py:87   rw */
  JM_FIELD_C = 2,        /* f32 (M,N,2,2) mpm_sim.py:88   rw */
  JM_FIELD_F = 3,        /* f32 (M,N,2,2) mpm_sim.py:89   rw */
  JM_FIELD_MATERIAL = 4, /* i32 (M,N)     mpm_sim.py:90   r (set by jm_load_particles) */
  JM_FIELD_JP = 5,       /* f32 (M,N)     mpm_sim.py:91   rw */
  JM_FIELD_FIBER = 6,    /* f32 (M,N,2)   mpm_sim.py:112  rw */
  JM_FIELD_GLOW = 7,     /* f32 (M,N)     mpm_sim.py:81   rw (needs track_glow) */
  JM_FIELD_GRID_V = 8,   /* f32 (M,nx,ny,2) mpm_sim.py:92 r: grid velocity of the last substep */
  JM_FIELD_GRID_M = 9,   /* f32 (M,nx,ny) mpm_sim.py:93  r: grid mass of the last substep */
  JM_FIELD_FITNESS = 10  /* f32 (M,3)     mpm_sim.py:109  r: [sum_y, count, sum_x] */
} jm_field;

/* Per-instance parameters (arrays of n_instances floats). */
typedef enum jm_param {
  JM_PARAM_ACTUATION = 0,       /* instance_actuation, mpm_sim.py:116 (default 500) */
  JM_PARAM_CONTRACTION = 1,     /* instance_act_contraction, mpm_sim.py:121 (default 0.2) */
  JM_PARAM_FREQ = 2,            /* instance_freq, mpm_sim.py:125 (default 1.0) */
  JM_PARAM_PAYLOAD_DENSITY = 3, /* instance_payload_density, mpm_sim.py:142 (default 2.5) */
  JM_PARAM_E_JELLY = 4          /* jelly/muscle Young's modulus, mpm_sim.py:31 (default 700);
                                   a module constant in the reference, per instance here */
} jm_param;

/* Stats modes for jm_compute_stats. */
typedef enum jm_stats_mode {
  JM_STATS_PAYLOAD = 0, /* compute_payload_stats, mpm_sim.py:746-753 (material 2) */
  JM_STATS_BODY = 1,    /* compute_body_stats, mpm_sim.py:755-762 (materials 1 and 3) */
  JM_STATS_ANY = 2      /* _get_stats_any, mpm_sim.py:779-792: payload, else body */
} jm_stats_mode;

/* Engine configuration; mirrors the reference's import-time settings (mpm_sim.py:11-22). */
typedef struct jm_config {
  int32_t abi_version;   /* must be JM_ABI_VERSION */
  int32_t n_instances;   /* JELLY_INSTANCES, mpm_sim.py:11 */
  int32_t n_particles;   /* JELLY_PARTICLES, mpm_sim.py:13: slots per instance */
  int32_t quality;       /* mpm_sim.py:12: n_grid = 128*quality, dt = 2e-5/quality */
  int32_t n_grid_y;      /* JELLY_GRID_Y, mpm_sim.py:15; 0 means n_grid */
  int32_t track_glow;    /* 1: run the render-only glow pass (mpm_sim.py:335-338) */
  int32_t path;          /* jm_path */
  int32_t axisym;        /* JELLY_AXISYM, mpm_sim.py:17: 1 = axisymmetric (r, z) variant
                            (ring-weighted P2G, hoop stress, axis BC; mpm_sim.py:127-133,
                            163-165,236-304), on either path; the fused path runs it in
                            chunks of 256 x 2 or 256 x 4 (jm_set_tuning refuses others) */
  double domain_height;  /* JELLY_DOMAIN_H, mpm_sim.py:16 */
  int32_t nan_invalid;   /* 1: jm_run_batch_headless marks an instance invalid (valid = 0)
                            when its final CoM is not finite or any of its live particles
                            ends with a non-finite v, C or F (a diverged state: advection
                            clamps positions into the domain, mpm_sim.py:331-333, so a
                            diverged instance's CoM stays finite). 0 (default): the
                            reference's gate `final_y < 0.01` alone (mpm_sim.py:830-833),
                            under which a NaN CoM counts as valid */
  int32_t reserved;      /* must be 0 */
} jm_config;

/* Launch tuning (chosen by jellyfih_amd/autotune.py; defaults are sensible). */
typedef struct jm_tuning {
  int32_t block_substeps;   /* substeps per HIP-graph block (default 100) */
  int32_t sort_interval;    /* re-sort particles by cell every K substeps (default 100) */
  int32_t threads;          /* fused kernel: threads per workgroup (128, 256 or 512) */
  int32_t ppt;              /* fused kernel: particles per thread (2, 4, 8 or 16) */
  int32_t use_graphs;       /* 1: replay captured HIP graphs per block */
  int32_t xcd_affinity;     /* 1: keep each instance's workgroups on one XCD (needs M >= 8) */
  int32_t lazy_vc;          /* 1: keep v, C in registers across fused substeps */
  int32_t sort_strip;       /* cell order of the sort: strips of this many cell rows, the
                               columns of alternate strips in opposite directions, rows
                               within a column (default 8); 0: row-major cells */
  int32_t instance_groups;  /* fused path: instances split into this many groups (1..4,
                               default 2) whose substep kernels run concurrently on their
                               own streams, so one group's launch fills the CUs that the
                               other's tail leaves idle */
} jm_tuning;

typedef struct jm_handle jm_handle;

/* Fill *cfg with the reference defaults (16 instances x 80000 slots, 128x128, H=1). */
int jm_default_config(jm_config* cfg);

/* Allocate state for cfg->n_instances instances; replaces the field allocation at
 * mpm_sim.py:81-149 (ti.init + ti.field). Parameters start at the reference defaults
 * (mpm_sim.py:143-149); every slot starts dead until loaded. */
int jm_create(const jm_config* cfg, jm_handle** out);
int jm_destroy(jm_handle* h);

int jm_set_tuning(jm_handle* h, const jm_tuning* t);
int jm_get_tuning(jm_handle* h, jm_tuning* t);

/* Load one instance: replaces load_particles / _load_particles_kernel (mpm_sim.py:715-736).
 * pos (n,2) f32, mat (n,) i32 (material < 0 marks a dead slot), fiber (n,2) f32 or NULL
 * for the reference default (0,1). n <= n_particles; slots n..N-1 become dead.
 * Resets v=0, F=I, C=0, Jp=1 for the instance. */
int jm_load_particles(jm_handle* h, int32_t instance, const float* pos, const int32_t* mat,
                      const float* fiber, int64_t n);

/* Item access of the per-instance fields (mpm_sim.py:116-149; writes as in
 * evolve.py:150-152,172). values has n_instances floats. */
int jm_set_instance_param(jm_handle* h, int32_t param, const float* values);
int jm_get_instance_param(jm_handle* h, int32_t param, float* values);

/* sim_time[None] (mpm_sim.py:95), fp32 accumulator advanced by dt per substep (:340). */
int jm_set_time(jm_handle* h, float t);
int jm_get_time(jm_handle* h, float* t);

/* Advance all instances by n substeps: replaces n calls of substep() (mpm_sim.py:153-340). */
int jm_substep(jm_handle* h, int64_t n);

/* fitness_buffer after clear_fitness_buffer + compute_payload_stats / compute_body_stats
 * (mpm_sim.py:740-762): out (M,3) f32 = [sum_y, count, sum_x] for jm_stats_mode. */
int jm_compute_stats(jm_handle* h, int32_t mode, float* out);

/* run_batch_headless (mpm_sim.py:802-834): t=0, CoM stats, `steps` substeps, stats;
 * out (M,5) f64 = [init_y, init_x, final_y, final_x, valid]. */
int jm_run_batch_headless(jm_handle* h, int64_t steps, double* out);

/* Copy a field to/from host memory in the reference layout (Taichi field.to_numpy /
 * item writes, e.g. helpers/fluid_analysis.py:26-27, helpers/fluid_test.py:145-151).
 * nbytes must equal the field's size. */
int jm_read_field(jm_handle* h, int32_t field, void* out, int64_t nbytes);
int jm_write_field(jm_handle* h, int32_t field, const void* in, int64_t nbytes);

/* Live (material >= 0) particle count of an instance. */
int jm_live_count(jm_handle* h, int32_t instance, int64_t* n);

/* Device order diagnostics: copy an instance's live particles in the engine's current
 * (cell-sorted) order. slot (n,) i32 = reference slot index; x (n,2) f32. Either may be
 * NULL. Used to check the sort permutation against a stable counting sort. */
int jm_read_device_order(jm_handle* h, int32_t instance, int32_t* slot, float* x, int64_t n);

/* Re-sort all particles by P2G base cell now (normally done every sort_interval). */
int jm_sort_now(jm_handle* h);

/* Measurement: run one block of n >= 2 fused substeps without graph capture, with HIP
 * events on the handle's stream around its n-1 fused G2P->P2G kernels. Returns their
 * average duration and the whole block's duration (ms). Advances the simulation like
 * jm_substep(h, n). No reference counterpart (the reference has no timing hooks). */
int jm_timed_block(jm_handle* h, int32_t n, float* fused_ms, float* block_ms);

/* ---- Renderer (mpm_sim.py:343-712) ----------------------------------------------------
 * The frame buffer is res x res x 3 fp32 on the device, [py][px][c] like the reference's
 * frame_buffer field (video_res = 1024 there); instance m is drawn in tile
 * (m / grid_side, m % grid_side) of res_sub pixels. All calls run on the handle's stream
 * after the pending loads are committed. */

/* clear_frame_buffer (:361-364, value 0), clear_frame_buffer_white (:483-485, 0.98): fill the
 * frame buffer with `value`; allocates (or reallocates) it at res x res. */
int jm_render_clear(jm_handle* h, int32_t res, float value);

/* render_frame_abyss (:385-477): additive splat of every live particle, coloured by
 * material (water by water_mode = water_angle_color: 0 ghost, 1 plain, 2 rainbow; jelly
 * from hue[m]; muscle from muscle_hue[m] and the activation at sim_time; payload orange).
 * hue, muscle_hue: host arrays of n_instances floats (instance_hue, instance_muscle_hue). */
int jm_render_abyss(jm_handle* h, int32_t res_sub, int32_t grid_side, float radius,
                    const float* hue, const float* muscle_hue, int32_t water_mode);

/* render_flat_pass (:487-517): particles of `material` set the pixels of their disc to
 * (r, g, b). */
int jm_render_flat(jm_handle* h, int32_t res_sub, int32_t grid_side, float radius,
                   int32_t material, float r, float g, float b);

/* tone_map_and_encode (:366-383), in place. */
int jm_render_tone_map(jm_handle* h);

/* compute_vorticity_grid (:560-578): vort (n_instances, nx, ny) = tanh(curl(grid_v) *
 * scale) at interior nodes with mass, from the grid the last substep's G2P read. */
int jm_render_vorticity_grid(jm_handle* h, float scale);

/* The vorticity passes: mode 0 render_vorticity_overlay (:520-558, param = vort_scale,
 * curl computed directly), 1 render_vorticity_rdbu (:581-624, param = alpha_max),
 * 2 render_vorticity_hueshift (:626-668, param = strength), 3 render_vorticity_white
 * (:670-712, param = alpha_max; fills the buffer with 1.0 first). Modes 1-3 read the grid of
 * the last jm_render_vorticity_grid. */
int jm_render_vorticity(jm_handle* h, int32_t mode, int32_t res_sub, int32_t grid_side,
                        float param);

/* Copy the frame buffer (res*res*3 floats) or the vorticity grid (n_instances*nx*ny floats,
 * [m][i][j] like vort_grid) to host memory; n is the number of floats. */
int jm_read_frame(jm_handle* h, float* out, int64_t n);
int jm_read_vorticity(jm_handle* h, float* out, int64_t n);

/* The frame buffer's 8-bit encode as every viewer caller writes it on the host,
 * (np.clip(img, 0, 1) * 255).astype(np.uint8) (evolve.py:396-397, helpers/view_single.py:257,
 * helpers/view_generation.py:210-211): clipped to [0, 1], multiplied by 255 in fp32,
 * truncated; NaN encodes as 0. Encoded on the device, so res*res*3 bytes cross to the host
 * instead of res*res*3 floats. n is the number of bytes (res*res*3). */
int jm_read_frame_u8(jm_handle* h, uint8_t* out, int64_t n);

/* ---- Tank building (host code, no GPU) -------------------------------------------------
 * Replaces the k-d tree query of make_jelly.py:359-366 (fill_tank, step 3): keep[i] = 1 when
 * every body point lies farther than `margin` from candidate i (Euclidean distance in
 * double, compared with >), else 0. body (n_body,2) and cand (n_cand,2) are host arrays of
 * double. */
int jm_water_mask(const double* body, int64_t n_body, const double* cand, int64_t n_cand,
                  double margin, uint8_t* keep);

/* make_jelly.py:31-58 (_polygon_self_intersects): *out = 1 when two non-adjacent edges of
 * the closed polygon poly (n,2, host double) cross properly, else 0. */
int jm_polygon_self_intersects(const double* poly, int64_t n, int32_t* out);

/* make_jelly.py:143-157,209-215 (Path(poly).contains_points over the candidate meshgrid):
 * out[j*nx + i] = 1 when (xs[i], ys[j]) is inside the polygon poly (n,2, host double, closed
 * implicitly) by the crossing-number rule of matplotlib's point-in-path test, bit for bit;
 * xs ascending. */
int jm_points_in_polygon_grid(const double* poly, int64_t n, const double* xs, int64_t nx,
                              const double* ys, int64_t ny, uint8_t* out);

/* make_jelly.py:61-334 (generate_phenotype): the robot particles of one genome (genome[0..8]
 * the morphology genes; n_genes >= 9), bit-identical to the reference's numpy evaluation:
 * pos (n,2) double with the spawn offset applied (spawn: 2 doubles), mat (n,) 1 jelly,
 * 2 payload, 3 muscle, fiber (n,2) float; *self_intersecting as make_jelly.py:139-141.
 * axisym != 0 is make_jelly's JELLY_AXISYM=1 branch (the right half alone, payload over
 * [0, PAYLOAD_WIDTH/2]). *n_out is set to the particle count; when it exceeds capacity
 * nothing is written and JM_ERR_ARG is returned (call again with capacity >= *n_out). */
int jm_generate_phenotype(const double* genome, int32_t n_genes, const double* spawn,
                          int32_t grid_res, int32_t with_payload, int32_t axisym, double* pos,
                          int32_t* mat, float* fiber, int64_t capacity, int64_t* n_out,
                          int32_t* self_intersecting);

/* make_jelly.py:336-402 (fill_tank): the fixed-size particle arrays of one instance, robot
 * first, then the water lattice minus the exclusion zone (truncated to fit), then dead slots
 * at (-1,-1) with material -1; fibres (0,1) outside muscle. pos (max_particles,2) float, mat
 * (max_particles,) int32, fiber (max_particles,2) float. spawn NULL selects the reference's
 * default ((0.5, 0.40), or (0, 0.40) when axisym != 0). info[5]: n_robot, n_water,
 * muscle_count, n_jelly, self_intersecting. JM_ERR_ARG when the robot alone exceeds
 * max_particles. */
int jm_fill_tank(const double* genome, int32_t n_genes, int64_t max_particles, int32_t grid_res,
                 const double* spawn, double water_margin, int32_t with_payload,
                 double domain_height, int32_t axisym, float* pos, int32_t* mat, float* fiber,
                 int64_t* info);

/* ---- Differentiable path -----------------------------------------------------------
 * No counterpart in the reference (it has no autodiff). A batched substep and its adjoint
 * over caller-owned device buffers (e.g. torch tensors), P particles in any order, in fp32
 * or fp64 (chosen at jm_diff_create; every scalar buffer below is of that type):
 *   state (P,12) per particle: x0 x1 v0 v1 C00 C01 C10 C11 F00 F01 F10 F11 (with
 *   JM_DIFF_MIXED the last four are F - I: F00-1 F01 F10 F11-1);
 *   inputs: inst (P,) int32 instance index, mat (P,) int32, fiber (P,2), amp_scale (P,)
 *   (per-particle multiplier on instance_actuation), params (M,5) = instance_actuation,
 *   instance_act_contraction, instance_freq, instance_payload_density, E_jelly.
 * All pointers are device pointers; work is queued on `stream` (hipStream_t, may be NULL). */
typedef struct jm_diff jm_diff;
typedef struct jm_diff_inputs {
  const int32_t* inst;
  const int32_t* mat;
  const void* fiber;
  const void* amp_scale;
  const void* params;
} jm_diff_inputs;

/* Grid buffers for cfg's grid and n_instances, and scratch for n_particles particles.
 * scalar_bits = 32: fp32 buffers and arithmetic (the reference's precision); 64: fp64 buffers
 * and arithmetic; JM_DIFF_MIXED: fp32 buffers (states, grids, inputs, gradients; every
 * "scalar" below is fp32) with each substep evaluated, and the adjoint carried between
 * substeps, in fp64, and the states' F columns holding F - I (its adjoint is F's). In fp32
 * arithmetic the small strains near F = I (the jelly's F - R, the water's J - 1 at
 * stiffness 1e5) keep only a few digits, which costs gradients percents over a few hundred
 * substeps; fp32 storage of the results, with F stored as F - I, costs ~1e-5 there. */
#define JM_DIFF_MIXED 3264
int jm_diff_create(const jm_config* cfg, int64_t n_particles, int32_t scalar_bits,
                   jm_diff** out);
int jm_diff_destroy(jm_diff* d);

/* Declares particles [water_start, n_particles) all water (material 0): jm_bwd_block then
 * runs them with a water-only instantiation of its kernel (the other materials' SVD
 * derivative compiled out: more waves per SIMD) and the particles before them with the
 * general one. The caller packs the water after the body's particles (DiffTank's
 * order="split"); a non-water particle in the range would be treated as water. Default
 * n_particles (no water range); results are the same either way. No counterpart in the
 * reference (which has no adjoint). */
int jm_diff_set_water_start(jm_diff* d, int64_t water_start);

/* nsteps substeps from state_in at time t0 (the clock advances by dt per substep in the
 * handle's precision; mpm_sim.py:336-340 does it in fp32). grid_out = NULL: out (P,12) is
 * the final state. Otherwise out (nsteps,P,12) receives the state after every substep and
 * grid_out (nsteps,M,3,nx*ny) the momentum x, momentum y and mass planes of every
 * substep's grid: the trajectory the backward pass reads. */
int jm_fwd_block(jm_diff* d, const jm_diff_inputs* in, const void* state_in, double t0,
                 int32_t nsteps, void* out, void* grid_out, void* stream);

/* Reverse pass: traj (nsteps+1,P,12) holds the states at substeps 0..nsteps and grids the
 * nsteps grids, as jm_fwd_block stores them (state_in, then out, from t0); grad_out (P,12)
 * is the adjoint of the last state. Writes grad_in (P,12), the adjoint of the first;
 * accumulates grad_amp (P,), the adjoint of amp_scale, and grad_params (M,4, always fp64),
 * the adjoints of (E_jelly, instance_actuation, contraction, freq). grad_out and grad_in
 * may not alias. */
int jm_bwd_block(jm_diff* d, const jm_diff_inputs* in, const void* traj, const void* grids,
                 double t0, int32_t nsteps, const void* grad_out, void* grad_in,
                 void* grad_amp, double* grad_params, void* stream);

/* jm_fwd_block (final-state mode, grid_out = NULL) without any device allocation. No
 * counterpart in the reference; it is for callers that cannot allocate: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc. There is no handle:
 * cfg (n_instances, quality, n_grid_y, domain_height) and n_particles / scalar_bits are
 * what jm_diff_create takes, and the inputs, state_in and out are the caller's device
 * buffers as for jm_fwd_block (work on `stream`, on the current device). The grid ring and
 * the states between kernels live in one fixed array of JM_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends: 9*M*nx*ny + 32*P scalars must fit (each part rounded up to
 * 256 bytes), else JM_ERR_ARG and nothing runs (e.g. fp32: one 128x128 instance up to about
 * 257K particles). Every call on one GPU uses that GPU's array: calls that overlap on one
 * GPU (two streams, two threads) race; calls queued on one stream do not. The kernels clear
 * every grid they accumulate into, so earlier calls' contents do not matter. Returns the
 * same results as jm_fwd_block (the same kernels). */
#define JM_NOALLOC_SCRATCH_BYTES (32u << 20)
int jm_fwd_block_noalloc(const jm_config* cfg, int64_t n_particles, int32_t scalar_bits,
                         const jm_diff_inputs* in, const void* state_in, double t0,
                         int32_t nsteps, void* out, void* stream);

/* ti.sync(): wait for all queued work. */
int jm_sync(jm_handle* h);

/* The handle's HIP stream (hipStream_t), for event timing and interop. */
void* jm_stream(jm_handle* h);

/* Thread-local message of the last failure ("" if none). */
const char* jm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* JELLYMX_H */
