/*
 * ludwig_hip.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann time step
 * (libludwig_hip.so).
 *
 * Drop-in seam: the Julia host of OpenLUDWIG would `ccall` these in place of
 *   - execute_timestep_batch!  (src/solver/solver_control.jl:199-219), which
 *     drives recursive_step! (:6-197) -> perform_timestep_v2!
 *     (src/collision/timestep.jl:1-159) -> stream_collide_cumulant!
 *     (src/collision/kernel_cumulant.jl:29-627), apply_bouzidi_correction!
 *     (src/geometry/bouzidi_kernel.jl:99-123) and stream_collide_thermal!
 *     (src/collision/kernel_thermal.jl:14-202);
 *   - compute_aerodynamics!    (src/forces/surface.jl:913-919);
 *   - adapt(backend, BlockLevel) (src/main.jl:103), init_eq! and
 *     init_thermal_eq! (src/main.jl:113-172);
 *   - the downloads of io_vtk.jl:46-51 and diagnostics.jl:43-93.
 *
 * Conventions
 *   - Host arrays are borrowed for the duration of a call and copied; the
 *     caller keeps ownership.  The context owns all device memory.
 *   - Tables use the reference's own values and shapes (1-based, 0 = absent;
 *     column-major), so Julia arrays pass unchanged.
 *   - Every function returns 0 on success or a negative LB_E* code; the message
 *     is in lb_last_error(ctx).  No C++ exception crosses the ABI.
 *   - A context is externally synchronised (one host thread at a time) and
 *     issues all device work on one HIP stream.
 */
#ifndef LUDWIG_HIP_H
#define LUDWIG_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LB_OK 0
#define LB_EINVAL (-1)
#define LB_EHIP (-2)
#define LB_ENOMEM (-3)
#define LB_ESTATE (-4)
#define LB_ENONFINITE (-5)

#define LB_STORAGE_FP32 0
#define LB_STORAGE_FP16 1 /* populations stored as f - w in binary16 */
#define LB_STORAGE_BF16 2 /* populations stored as f - w in bfloat16 */

typedef struct lb_ctx lb_ctx;

/* Case-wide configuration fixed at creation (config_loader.jl globals that the
 * launcher reads: TEMPORAL_INTERPOLATION, DDF_ENABLED, Q_MIN_THRESHOLD).     */
typedef struct {
    int32_t n_levels;
    int32_t domain_nx, domain_ny, domain_nz; /* coarse cells (physics_scaling.jl:119-121) */
    int32_t temporal_interpolation;          /* TEMPORAL_INTERPOLATION                   */
    int32_t ddf_enabled;                     /* DDF_ENABLED                              */
    float q_min_threshold;                   /* Q_MIN_THRESHOLD                          */
    int32_t storage;                         /* LB_STORAGE_*                             */
    int32_t strict_fp;                       /* 1: no FMA contraction (bitwise-comparable) */
    int32_t wide_offsets;                    /* 1: 64-bit population offsets on every level.
                                              * A level switches to them by itself once its
                                              * population buffer reaches 4 GiB; forcing them
                                              * runs those kernel instances on any case.     */
    int32_t reserved[6];
} lb_config;

/* One BlockLevel (src/geometry/blocks.jl:16-74) as the reference holds it. */
typedef struct {
    int32_t level_id; /* 1-based                                              */
    int32_t n_blocks;
    const int32_t* map_x; /* (n_b) 1-based block coords, sorted (bx,by,bz)   */
    const int32_t* map_y;
    const int32_t* map_z;
    const int32_t* neighbor_table; /* (n_b,27) column-major, 1-based, 0 = none */
    const int32_t* block_pointer;  /* (ptr_dim_x,ptr_dim_y,ptr_dim_z) col-major */
    int32_t ptr_dim_x, ptr_dim_y, ptr_dim_z;
    const uint8_t* obstacle; /* (8,8,8,n_b) Bool                             */
    const float* sponge;     /* (8,8,8,n_b)                                   */
    const float* wall_dist;  /* (8,8,8,n_b) or NULL when the wall model is off */
    float tau;               /* level.tau                                     */
    float tau_g;             /* level.tau_g (thermal)                          */
    double dx;               /* level.dx (metres), used by lb_forces           */
    /* Bouzidi (bouzidi_enabled levels): the sparse cell list of
     * compute_bouzidi_qmap_sparse and either the dense q map
     * (8,8,8,n_b,27) or q per listed cell (n_bc,27); FP16 bits.          */
    int64_t n_boundary_cells;
    const int32_t* bouzidi_cell_block; /* (n_bc) 1-based                       */
    const int8_t* bouzidi_cell_x;      /* (n_bc) 1-based                       */
    const int8_t* bouzidi_cell_y;
    const int8_t* bouzidi_cell_z;
    const uint16_t* bouzidi_q_dense;  /* may be NULL if bouzidi_q_cells is set */
    const uint16_t* bouzidi_q_cells;  /* may be NULL if bouzidi_q_dense is set */
} lb_level_desc;

/* Per-call scalars of execute_timestep_batch! and the globals read inside
 * perform_timestep_v2! (timestep.jl:40-86, kernel_cumulant.jl:29-69).      */
typedef struct {
    float c_wale;           /* c_wale (C_SMAGO)                          */
    float nu_sgs_background;
    float inlet_turbulence;
    int32_t symmetric;      /* SYMMETRIC_ANALYSIS                         */
    int32_t wall_model;     /* params.wall_model_active                   */
    int32_t sponge_blend_distributions;
    float omega_bulk, omega_3, omega_4;
    int32_t adaptive_omega4;
    float lambda_param;
    int32_t limiter;        /* 0 none, 1 factored, 2 positivity            */
    int32_t compressible;   /* COMPRESSIBILITY_CORRECTION                  */
    float T_inlet, T_wall;  /* DDF_T_INLET, DDF_T_WALL                      */
    int32_t wall_bc_adiabatic;
    int32_t reserved[8];
} lb_step_params;

/* compute_aerodynamics! inputs (surface.jl:766-919, structs.jl:44-74).     */
typedef struct {
    double mesh_offset[3];
    double velocity_scale; /* params.velocity_scale                         */
    double rho_physical;
    double rho_ref, u_ref, area_ref, chord_ref;
    double moment_center[3];
    float c_wale;            /* C_SMAGO                                      */
    float nu_sgs_background; /* NU_SGS_BACKGROUND                            */
    int32_t search_radius;   /* 5 at every call site (main.jl:238)           */
    int32_t symmetric;
    int32_t reserved[8];
} lb_force_params;

/* lb_forces output slots */
#define LB_F_FXP 0
#define LB_F_FYP 1
#define LB_F_FZP 2
#define LB_F_FXV 3
#define LB_F_FYV 4
#define LB_F_FZV 5
#define LB_F_MX 6
#define LB_F_MY 7
#define LB_F_MZ 8
#define LB_F_FX 9
#define LB_F_FY 10
#define LB_F_FZ 11
#define LB_F_CD 12
#define LB_F_CL 13
#define LB_F_CS 14
#define LB_F_CMX 15
#define LB_F_CMY 16
#define LB_F_CMZ 17
#define LB_F_N 18

/* lb_download_level / lb_upload_level fields (reference layout, FP32) */
#define LB_FIELD_F 0   /* (8,8,8,n_b,27) latest populations          */
#define LB_FIELD_RHO 1 /* (8,8,8,n_b) level.rho                        */
#define LB_FIELD_VEL 2 /* (8,8,8,n_b,3) latest velocity                */
#define LB_FIELD_G 3   /* (8,8,8,n_b,27) latest thermal populations    */
#define LB_FIELD_T 4   /* (8,8,8,n_b) level.temperature                */

const char* lb_last_error(const lb_ctx* ctx);
const char* lb_version(void);

/* Create a context on HIP device `device`.  Replaces the backend selection and
 * allocation of solve_main (main.jl:49-51, 103).                          */
int lb_create(lb_ctx** ctx, int device, const lb_config* cfg);
void lb_destroy(lb_ctx* ctx);

/* adapt(backend, BlockLevel) for one level (main.jl:103, blocks.jl:76-100).
 * Levels must be uploaded in order 1..n_levels.                           */
int lb_level_upload(lb_ctx* ctx, const lb_level_desc* desc);

/* init_eq! + init_thermal_eq! for every level (main.jl:113-172).          */
int lb_init_equilibrium(lb_ctx* ctx, float T0);

/* execute_timestep_batch!(grids, t_start, n_steps, u_curr, ...)
 * (solver_control.jl:199-219): n_steps coarse steps t_start..t_start+n-1,
 * each the full 2:1 multilevel recursion.  Asynchronous on the context's
 * stream; lb_synchronize waits.                                            */
int lb_step_batch(lb_ctx* ctx, int64_t t_start, int32_t n_steps, float u_curr, const lb_step_params* p);
int lb_synchronize(lb_ctx* ctx);

/* upload_mesh_to_gpu (geometry.jl:50-73): FP32 centres, normals, areas.    */
int lb_mesh_upload(lb_ctx* ctx, int64_t n_tri, const float* cx, const float* cy, const float* cz, const float* nx,
                   const float* ny, const float* nz, const float* areas);

/* compute_aerodynamics! on the finest level (surface.jl:913-919), with the
 * integration reduced in FP64 in a fixed order (deterministic).  out has
 * LB_F_N doubles.  Optional maps (may be NULL): per-triangle p, tau_x/y/z. */
int lb_forces(lb_ctx* ctx, const lb_force_params* fp, double* out, float* p_map, float* sx_map, float* sy_map,
              float* sz_map);

/* The per-triangle surface maps (pressure and wall shear, Float32) of one
 * side of the surface on the finest level: side = +1 is map_surface_stresses!
 * (probe along the outward normal, surface.jl:350-507, 766-826; the maps
 * lb_forces returns), side = -1 is map_minus_side_stresses! (surface.jl:
 * 511-662, 923-954), used by compute_net_aerodynamics! for open surfaces.  The four
 * arrays have n_tri floats each.                                           */
int lb_surface_maps(lb_ctx* ctx, const lb_force_params* fp, int32_t side, float* p_map, float* sx_map,
                    float* sy_map, float* sz_map);

/* Field transfer in the reference layout (io_vtk.jl:46-51).               */
int lb_download_level(lb_ctx* ctx, int level_id, int field, float* dst);
int lb_upload_level(lb_ctx* ctx, int level_id, int field, const float* src);

/* FP64 diagnostics over fluid cells of a level (compute_flow_stats,
 * diagnostics.jl:43-93, plus the global mass/energy check): out[0] sum rho,
 * [1] sum 0.5 rho u^2, [2] rho min, [3] rho max, [4] number of non-finite
 * populations, [5] fluid cell count, [6] sum of all populations (mass),
 * [7] v_max = max |u| over fluid cells.                                     */
int lb_stats(lb_ctx* ctx, int level_id, double* out);

/* lb_stats without allocating: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under stream capture, tight
 * serving loops).  The same outputs as lb_stats, to the bit; lb_stats
 * allocates its 64 KiB of partial sums on first use, this variant keeps them
 * and the reduced result in a fixed per-device array of 65,600 bytes
 * (1,024 partial rows of 8 doubles, the most the statistics kernel
 * produces for any level, plus 8).  Calls that overlap on one GPU share that
 * array and race; the caller serialises them.                             */
int lb_stats_noalloc(lb_ctx* ctx, int level_id, double* out);

/* Divergence check (check_stability, diagnostics.jl:95-121, which only
 * prints warnings; here a diverged case is an error): counts non-finite
 * density and velocity values of every level after the last completed
 * coarse step.  Returns LB_ENONFINITE (the count in lb_last_error) when any
 * is found, else LB_OK; *n_nonfinite (may be NULL) receives the count.
 * Reads 16 bytes per cell; meant for the diagnostic cadence.               */
int lb_check_finite(lb_ctx* ctx, int64_t* n_nonfinite);

/* Device bytes the context holds (serving admission).                    */
int64_t lb_device_bytes(const lb_ctx* ctx);

/* Device bytes a case would hold, computed before anything is allocated
 * (no GPU needed): what lb_create, lb_level_upload of every level and
 * lb_mesh_upload of n_tri triangles allocate, plus the engine's ramp table
 * of n_ramp_batches values (lb_engine_add).  Used to admit serving requests
 * against the HBM budget (no counterpart in the reference).                */
int lb_footprint(const lb_config* cfg, const lb_level_desc* levels, int32_t n_levels, int64_t n_tri,
                 int32_t n_ramp_batches, int64_t* bytes);

/* Free and total device memory (hipMemGetInfo), for the admission budget. */
int lb_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes);

/* Kernel timing (no counterpart in the reference, which only times the host
 * loop, main.jl:232): when enabled, HIP events are recorded on the context's
 * stream around every stream-collide launch.  lb_profile_read synchronises,
 * returns the summed kernel time (ms) and launch count of level_id since the
 * last read, and resets them.                                              */
int lb_profile_enable(lb_ctx* ctx, int enable);
int lb_profile_read(lb_ctx* ctx, int level_id, double* ms, int64_t* launches);
/* The same for kernel kind `kernel`: 0 stream-collide, 1 thermal DDF.     */
int lb_profile_read_kernel(lb_ctx* ctx, int level_id, int kernel, double* ms, int64_t* launches);

/* ------------------------------------------------------------------------
 * Batched engine (no counterpart in the reference, whose run_all_cases,
 * main.jl:425-469, runs one case after another).  Cases are built as usual
 * (lb_create, lb_level_upload, lb_mesh_upload, lb_init_equilibrium) and then
 * attached; one engine step advances every attached case together:
 * the 2:1 recursion is run over depths, and each launch covers the same depth
 * of all cases over a concatenated block list (cases with fewer levels are
 * aligned on their finest level and advance 2^(Lmax - L) coarse steps per
 * engine step).  The step is captured as a hipGraph and replayed.  Per case,
 * the sequence of coarse steps, seeds and inlet velocities is that of
 * execute_timestep_batch! with batches of `batch_len` steps, so a case's
 * results do not depend on which other cases share the engine.
 * While attached, a case may only be used through the engine.            */
typedef struct lb_engine lb_engine;

int lb_engine_create(lb_engine** e, int device, int32_t storage, int32_t strict_fp);
void lb_engine_destroy(lb_engine* e);
const char* lb_engine_last_error(const lb_engine* e);

/* Attach an initialised case for coarse steps 1..n_coarse_steps.  u_batches
 * holds the inlet velocity of each ramp batch of batch_len coarse steps
 * (n_batches >= ceil(n_coarse_steps / batch_len)), as the host loop computes
 * it (main.jl:214-218).  p: the physics parameters of lb_step_batch.       */
int lb_engine_add(lb_engine* e, lb_ctx* ctx, int64_t n_coarse_steps, const float* u_batches, int32_t n_batches,
                  int32_t batch_len, const lb_step_params* p);

/* Detach a case (waits for the engine's work); its fields then hold the state
 * after its last completed coarse step and lb_forces etc. apply again.     */
int lb_engine_remove(lb_engine* e, lb_ctx* ctx);

/* Enqueue n engine steps (asynchronous). */
int lb_engine_step(lb_engine* e, int32_t n_steps);
int lb_engine_synchronize(lb_engine* e);

/* Coarse steps a case has completed once the enqueued work has run. */
int lb_engine_case_progress(const lb_engine* e, const lb_ctx* ctx, int64_t* t_done);

/* Force sampling of an attached case inside the engine step
 * (accumulate_forces!, surface.jl:36-56, at the samples main.jl:236-243
 * takes): at the end of every coarse step t with t % diag_freq == 0 and
 * t > ramp_steps, the finest level's forces -- as lb_forces computes them
 * with fp -- are integrated and added to the case's FP64 accumulator.  The
 * case needs a mesh (lb_mesh_upload).  A case's samples do not depend on the
 * other attached cases.                                                     */
int lb_engine_sample_forces(lb_engine* e, lb_ctx* ctx, const lb_force_params* fp, int64_t diag_freq,
                            int64_t ramp_steps);
/* The accumulator (waits for the engine): sums[0..8] = sums over the samples
 * of Fp xyz, Fv xyz, M xyz (the LB_F_FXP..LB_F_MZ slots of lb_forces, with
 * its symmetric doubling), *n_samples their number.  Read it before
 * lb_engine_remove.                                                        */
int lb_engine_case_forces(lb_engine* e, const lb_ctx* ctx, double* sums, int64_t* n_samples);

/* The divergence check of lb_check_finite over every attached case, after
 * the enqueued work (waits for it).  The first `cap` diverged cases are
 * written to diverged[] and their number to *n_diverged; returns
 * LB_ENONFINITE if there is any.  The other cases are not affected: a
 * diverged case can be removed (lb_engine_remove) while they go on.       */
int lb_engine_check_finite(lb_engine* e, lb_ctx** diverged, int32_t cap, int32_t* n_diverged);

/* Launch options (picked by the autotuner; results do not depend on them):
 * graph (1: replay a captured hipGraph per step), the work-group schedule
 * of the stream-collide launches (LB_SCHED_*; default LB_SCHED_LINEAR, the
 * only one lb_step_batch uses) and the occupancy of the interior-block
 * stream-collide kernel, 5 or 6 waves per SIMD (the persistent form runs
 * at 5; default and lb_step_batch: 5 for FP32 storage, 6 for 16-bit).    */
#define LB_SCHED_XCD 0        /* work-groups dealt to XCDs in contiguous runs  */
#define LB_SCHED_LINEAR 1     /* work-group i takes half-tile i                */
#define LB_SCHED_PERSISTENT 2 /* 5 work-groups per CU, each looping over the
                               * half-tiles in order (stream-collide only; the
                               * thermal kernel runs LB_SCHED_LINEAR)          */
int lb_engine_set_options(lb_engine* e, int32_t use_graph, int32_t schedule, int32_t waves);

/* ------------------------------------------------------------------------
 * Setup on the GPU.  compute_bouzidi_qmap_sparse (src/geometry/
 * bouzidi_setup.jl:64-167, compute_q_for_cell bouzidi_math.jl:53-102): the
 * same arguments and outputs as lbs_bouzidi_qmap (include/ludwig_setup.h),
 * bit for bit (FP64, no contraction), on HIP device `device`.  The three
 * arrays are malloc'd; release them with lb_free.  Errors: negative LB_E*,
 * message in lb_setup_gpu_last_error().                                     */
int lb_bouzidi_qmap_gpu(int device, int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy,
                        int gdz, const double* tris, int64_t ntri, double dx, const double* offset,
                        int32_t** cells_out, uint16_t** q_out, int32_t** tri_out, int64_t* n_out);
/* voxelize_blocks! (domain_generation.jl:74-112) on the GPU: sets obst[b*512
 * + cell] = 1 for every cell whose 1.5 dx box a candidate triangle crosses
 * (SAT test, FP64, no contraction); other entries are left as they are.
 * Same arguments and the same result as lbs_voxelize of the setup library. */
int lb_voxelize_gpu(int device, int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy, int gdz,
                    const double* tris, int64_t ntri, double dx, const double* offset, uint8_t* obst);
/* compute_wall_distances! (domain_generation.jl:371-434) on the GPU: the
 * distance (FP32 metres) from each fluid cell to its nearest solid
 * neighbour among the 26, written into wall_dist for near-wall cells only;
 * *count_out (may be NULL) = number of such cells.  Same result as
 * lbs_wall_distances.                                                       */
int lb_wall_distances_gpu(int device, int64_t nb, const int32_t* coords, const int32_t* grid, int gdx, int gdy,
                          int gdz, const uint8_t* obst, double dx, float* wall_dist, int64_t* count_out);
void lb_free(void* p);
const char* lb_setup_gpu_last_error(void);

/* HBM bandwidth probe (no counterpart in the reference): streams `bytes`
 * with 16-byte accesses, one per work-item over a grid sized to the buffer;
 * gbps_out[0] copy (read + write counted), [1] read-only, [2] write-only,
 * [3] copy with non-temporal loads and stores, [4] the same copy as a
 * grid-stride loop over 8 work-groups per CU (five doubles), each averaged
 * over `reps` launches.  The practical ceiling reported beside the nominal
 * 8 TB/s.                                                                  */
int lb_bandwidth_probe(int device, int64_t bytes, int reps, double* gbps_out);

#ifdef __cplusplus
}
#endif
#endif
