/*
 * ffcm_hip.h: C ABI of the MI355X (gfx950) backend for the FFCM mobility operator.
 *
 * This library is the `gpu_acceleration = true` backend of ForceCouplingMethod.jl v0.1.0. Every
 * entry point below names the reference interface it replaces (paths relative to the reference
 * source tree). All calls return 0 on success or one of the FFCM_ERR_* codes; the message of
 * the last failure is available from ffcm_last_error().
 *
 * Conventions (shared with the host-side bindings):
 *  - Scalars are computed by the host in the working precision T exactly as
 *    FFCMConfig{T} derives them (src/config.jl:326-359) and passed verbatim in ffcm_params,
 *    as doubles; the library converts them back to T, which is exact.
 *  - Particle arrays are 3xN column-major matrices (x,y,z of particle 1, then particle 2, ...)
 *    in T, in the caller's original particle order.
 *  - One context owns one HIP stream and every device buffer; nothing is allocated after
 *    ffcm_create. A context is not thread-safe; distinct contexts may run concurrently.
 *  - Index buffers read back in the reference's Julia conventions: original_index and
 *    cell_start are 1-based, cell_end is the 1-based inclusive end (empty cell: end =
 *    start - 1), cell_hash is 0-based.
 */
#ifndef FFCM_HIP_H
#define FFCM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FFCM_OK = 0,
  FFCM_ERR_ARGUMENT = 1,  /* Julia ArgumentError (src/config.jl:267-312, :461-464) */
  FFCM_ERR_DIMENSION = 2, /* Julia DimensionMismatch (src/mobility.jl:29-34) */
  FFCM_ERR_DEVICE = 3,    /* no usable gfx950 device, or a HIP runtime failure */
  FFCM_ERR_MEMORY = 4,    /* device allocation failed at ffcm_create */
  FFCM_ERR_INTERNAL = 5
};

enum { FFCM_F32 = 0, FFCM_F64 = 1 };

/* Buffer identifiers for ffcm_set_buffer / ffcm_get_buffer / ffcm_buffer_bytes. Each maps to
 * one field of the reference's buffer sub-structs (src/config.jl:24-125); the host-side
 * layout is the reference's (column-major Julia arrays, SoA grid components), whatever the
 * device layout. */
enum {
  FFCM_BUF_Y_INPUT = 0,         /* ParticleBuffers.Y_input   3xN T (staging, original order) */
  FFCM_BUF_F_INPUT = 1,         /* ParticleBuffers.F_input   3xN T */
  FFCM_BUF_V_OUTPUT = 2,        /* ParticleBuffers.V_output  3xN T */
  FFCM_BUF_Y_WRAPPED = 3,       /* ParticleBuffers.Y_wrapped 3xN T */
  FFCM_BUF_Y_SORTED = 4,        /* ParticleBuffers.Y_sorted  3xN T */
  FFCM_BUF_F_SORTED = 5,        /* ParticleBuffers.F_sorted  3xN T */
  FFCM_BUF_CELL_HASH = 6,       /* CellBuffers.cell_hash      N int32, 0-based */
  FFCM_BUF_ORIGINAL_INDEX = 7,  /* CellBuffers.original_index N int32, 1-based */
  FFCM_BUF_CELL_START = 8,      /* CellBuffers.cell_start     C int32, 1-based */
  FFCM_BUF_CELL_END = 9,        /* CellBuffers.cell_end       C int32, 1-based inclusive */
  FFCM_BUF_FORCE_DENSITY = 10,  /* GridBuffers.force_density  3 x (Mx,My,Mz) T, SoA */
  FFCM_BUF_FLUID_VELOCITY = 11, /* GridBuffers.fluid_velocity 3 x (Mx,My,Mz) T, SoA */
  FFCM_BUF_FLUID_HAT = 12,      /* SolverState.fluid_hat      3 x (Mx/2+1,My,Mz) Complex{T} */
  FFCM_BUF_NEIGHBOR_MAP = 13,   /* CellBuffers.neighbor_map   13C int32, 0-based (read only) */
  FFCM_BUF_COUNTING_SORT_SCRATCH = 14, /* CellBuffers.counting_sort_scratch C int32 (workspace; the
                                        * radix sort does not read it) */
  FFCM_BUF_COUNT = 15
};

/* Every scalar of FFCMConfig{T} (src/config.jl:196-215) that the hot path reads. */
typedef struct ffcm_params {
  int32_t dtype; /* FFCM_F32 or FFCM_F64 */
  int32_t pad0;
  int64_t N;
  int32_t M[3]; /* num_grid_points */
  int32_t M_G;
  int32_t num_cells[3];
  int32_t pad1;
  double L[3];
  double R_c, a, sigma, Sigma, h, inv_h, mu;
  double cell_size[3];
  double inv_cell_size[3];
  const void *k_x, *k_y, *k_z; /* host wavevectors in T: lengths Mx/2+1, My, Mz */
} ffcm_params;

typedef struct ffcm_ctx ffcm_ctx;

/* Replaces _assemble_gpu_buffers (src/config.jl:461-464, called at :495-502; CUDA method
 * ext/ForceCouplingMethodCUDAExt/config.jl:19-103). Validates the parameters with the checks
 * of _validate_config_parameters (src/config.jl:267-312) before touching any device, then
 * allocates every buffer and plans the transforms. Backend limits on top of the reference's:
 * grid sizes must factor into primes <= 13, M_G <= 64, N <= 2^26 (N * max_rhs <= 2^28). */
int ffcm_create(ffcm_ctx **out, const ffcm_params *params);

/* The seam exactly as the reference calls it: _assemble_gpu_buffers(T, N, num_grid_points, M_G,
 * num_cells_total, k_x, k_y, k_z, neighbor_map_host) (src/config.jl:495-502; CUDA method
 * ext/ForceCouplingMethodCUDAExt/config.jl:19-103). dtype is FFCM_F32 or FFCM_F64 for T;
 * k_x, k_y, k_z are the host wavevectors in T; neighbor_map_host is the host half-shell map
 * of _build_neighbor_map (13 num_cells_total int32, src/cell_list.jl:258-289), from which the
 * per-axis cell counts are read and which is copied to the device (FFCM_BUF_NEIGHBOR_MAP).
 * The seam passes no box, kernel or fluid scalars: the context learns them from the
 * per-barrier entry points below (each takes the scalars its reference barrier receives) or,
 * for the fused entry points, from ffcm_set_scalars; until then those fail with
 * FFCM_ERR_ARGUMENT. An invalid map fails with FFCM_ERR_ARGUMENT before any device is touched. */
int ffcm_assemble(ffcm_ctx **out, int32_t dtype, int64_t N, const int32_t num_grid_points[3], int32_t M_G,
                  int64_t num_cells_total, const void *k_x, const void *k_y, const void *k_z,
                  const int32_t *neighbor_map_host);

/* Every scalar of the FFCMConfig (src/config.jl:196-215) for a context built by ffcm_assemble
 * (N, num_grid_points, M_G and num_cells must be the assembled ones; the wavevector pointers
 * are not read again). A fused mobility! method holding the config calls it once. */
int ffcm_set_scalars(ffcm_ctx *ctx, const ffcm_params *params);

/* The L1 barriers with the reference's arguments, buffers named by FFCM_BUF_* id (the
 * config fields the reference passes). Each stores the scalars it receives, runs its step on
 * the context stream and synchronises.
 *   ffcm_wrap_positions_kernel        wrap_positions!(Y_wrapped, Y_input, L)     src/cell_list.jl:22-28
 *   ffcm_assign_cells_kernel          _assign_cells_kernel!(cell_hash, Y_wrapped, inv_cell_size, num_cells)
 *                                                                                src/cell_list.jl:222-238
 *   ffcm_build_cell_list_kernel       _build_cell_list_kernel!(original_index, cell_start, cell_end,
 *                                     counting_sort_scratch, cell_hash)          src/cell_list.jl:64-93
 *   ffcm_gather_particles_kernel      _gather_particles_kernel!(Y_sorted, F_sorted, Y_wrapped, F_input,
 *                                     original_index)                            src/cell_list.jl:159-176
 *   ffcm_spread_forces_kernel         _spread_forces_kernel!(force_density, Y_sorted, F_sorted, σ, Σ, h,
 *                                     inv_h, num_grid_points, M_G, ...)          src/spread_forces.jl:73-130
 *   ffcm_forward_transform            mul!(fluid_hat_c, forward plan, force_density_c)  src/stokes_solve.jl:24-26
 *   ffcm_apply_inverse_stokes_kernel  _apply_inverse_stokes_kernel!(..., μ, inv_M)      src/stokes_solve.jl:70-110
 *   ffcm_inverse_transform            mul!(fluid_velocity_c, inverse plan, fluid_hat_c) src/stokes_solve.jl:38-40
 *   ffcm_interpolate_velocities_kernel _interpolate_velocities_kernel!(V_output, fluid_velocity, Y_sorted,
 *                                     original_index, σ, Σ, h, inv_h, num_grid_points, M_G, ...)
 *                                                                                src/interpolate.jl:82-136
 *   ffcm_correct_velocities_kernel    _correct_velocities_kernel!(V_output, Y_sorted, F_sorted, cell_start,
 *                                     cell_end, original_index, neighbor_map, num_cells, inv_cell_size, L,
 *                                     σ, Σ, a, μ, R_c)                            src/correct_velocities.jl:216-286
 * The transforms act on the three components together: the call for component 0 transforms
 * all three, and the calls for components 1 and 2 that directly follow it find them done. */
int ffcm_wrap_positions_kernel(ffcm_ctx *ctx, int dst_buffer, int src_buffer, const double L[3]);
int ffcm_assign_cells_kernel(ffcm_ctx *ctx, int cell_hash_buffer, int Y_buffer, const double inv_cell_size[3],
                             const int32_t num_cells[3]);
int ffcm_build_cell_list_kernel(ffcm_ctx *ctx);
int ffcm_gather_particles_kernel(ffcm_ctx *ctx, int Y_buffer, int F_buffer);
int ffcm_spread_forces_kernel(ffcm_ctx *ctx, double sigma, double Sigma, double h, double inv_h,
                              const int32_t num_grid_points[3], int32_t M_G);
int ffcm_forward_transform(ffcm_ctx *ctx, int component);
int ffcm_apply_inverse_stokes_kernel(ffcm_ctx *ctx, double mu, double inv_M);
int ffcm_inverse_transform(ffcm_ctx *ctx, int component);
int ffcm_interpolate_velocities_kernel(ffcm_ctx *ctx, int V_buffer, double sigma, double Sigma, double h, double inv_h,
                                       const int32_t num_grid_points[3], int32_t M_G);
int ffcm_correct_velocities_kernel(ffcm_ctx *ctx, int V_buffer, const int32_t num_cells[3],
                                   const double inv_cell_size[3], const double L[3], double sigma, double Sigma,
                                   double a, double mu, double R_c);

/* Releases the context (the finalizer of the config's device arrays). */
int ffcm_destroy(ffcm_ctx *ctx);

/* Message of the last failure on ctx; ctx may be NULL for the last ffcm_create failure of the
 * calling thread. */
const char *ffcm_last_error(const ffcm_ctx *ctx);

/* Replaces mobility! (src/mobility.jl:23-45) with its staging seam _stage_mobility_io! /
 * _retrieve_mobility_output! (ext/ForceCouplingMethodCUDAExt/mobility.jl:10-23): V, Y, F are
 * host 3xN arrays in T; Y and F are read only; V is overwritten. Synchronous. The host arrays are
 * staged through page-locked buffers the context allocates at creation. */
int ffcm_mobility(ffcm_ctx *ctx, void *V, const void *Y, const void *F, int64_t N);

/* The device-resident form of mobility! (the device->device branch of _stage_mobility_io!,
 * ext/ForceCouplingMethodCUDAExt/mobility.jl:1-8): V, Y, F are device pointers (3xN, T).
 * Enqueued on the context stream and not synchronised. */
int ffcm_mobility_device(ffcm_ctx *ctx, void *V, const void *Y, const void *F, int64_t N);

/* Allocation-free form of ffcm_mobility_device. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No context is created: params is validated as by
 * ffcm_create, and every intermediate buffer is placed in one static device array of
 * ffcm_noalloc_scratch_bytes() bytes (128 MiB; each device holds its own copy until the
 * process ends), which fits grids up to about 128^3 in Float32. Configurations that need more
 * fail with FFCM_ERR_ARGUMENT. V, Y, F are device pointers (3xN, T); the wavevectors are
 * recomputed on the device from L and the grid with the reference's formula, so
 * params->k_x, k_y, k_z are not read. Enqueued on `stream` (a hipStream_t; NULL is the null
 * stream), not synchronised. All calls on one device share its array: calls that overlap in
 * time on one device race, so order them (e.g. on one stream). Errors: the message is
 * ffcm_noalloc_last_error() (per thread). */
int ffcm_mobility_device_noalloc(const ffcm_params *params, void *V, const void *Y, const void *F, int64_t N,
                                 void *stream);
const char *ffcm_noalloc_last_error(void);
size_t ffcm_noalloc_scratch_bytes(void);

/* The position-cached form of the iterative-solver operator FFCMMobility (src/mobility.jl:88-147;
 * the reference re-uploads and re-sorts Y on every mul!, docs/src/devdocs/gpu-architecture.md:
 * 159-161). ffcm_set_positions runs steps 1-2 of mobility! for Y alone (wrap, hash, sort, the
 * slot-ordered positions and the spread's tile binning) and keeps them in the context;
 * ffcm_apply_forces then runs the rest of mobility!(V, config, Y, F) for that Y: it gathers F
 * into slot order, spreads, solves, interpolates and corrects. Any other call that rebuilds
 * the cell list (a mobility apply, the pullback, the per-step entry points, ffcm_set_buffer)
 * drops the cached positions, and ffcm_apply_forces then fails with FFCM_ERR_ARGUMENT until
 * ffcm_set_positions is called again. The plain forms take host 3xN arrays and synchronise;
 * the _device forms take device pointers and are enqueued on the context stream. */
int ffcm_set_positions(ffcm_ctx *ctx, const void *Y, int64_t N);
int ffcm_set_positions_device(ffcm_ctx *ctx, const void *Y, int64_t N);
int ffcm_apply_forces(ffcm_ctx *ctx, void *V, const void *F, int64_t N);
int ffcm_apply_forces_device(ffcm_ctx *ctx, void *V, const void *F, int64_t N);

/* Several right-hand sides sharing Y (no reference function; each right-hand side r gives
 * exactly mobility!(V_r, config, Y, F_r), src/mobility.jl:23-45). ffcm_create_multi builds a
 * context whose buffers hold up to max_rhs right-hand sides (ffcm_create is max_rhs = 1).
 * V and F are 3 x N x R arrays (right-hand side r is the 3xN block at offset 3 N r); the cell
 * list is built once and the R force grids are transformed together. ffcm_mobility_multi
 * takes host arrays and synchronises; the _device form takes device pointers and is
 * enqueued on the context stream. */
int ffcm_create_multi(ffcm_ctx **out, const ffcm_params *params, int max_rhs);
int ffcm_mobility_multi(ffcm_ctx *ctx, int R, void *V, const void *Y, const void *F, int64_t N);
int ffcm_mobility_multi_device(ffcm_ctx *ctx, int R, void *V, const void *Y, const void *F, int64_t N);

/* Batched ensemble (no reference function): n_systems independent suspensions, each with
 * n_rhs right-hand sides sharing its positions. Semantics per (system k, right-hand side r):
 * V_k[:, :, r] = mobility!(config_k, Y_k, F_k[:, :, r]) (src/mobility.jl:23-45). params is an
 * array of n_systems parameter sets of one dtype. The ensemble apply is captured into one
 * hipGraph over up to four stream lanes; the lane count is autotuned on the first apply (each
 * candidate graph is timed over two replays). V, Y, F are arrays of n_systems pointers:
 * Y[k] is 3 x N_k, V[k] and F[k] are 3 x N_k x n_rhs. */
typedef struct ffcm_batch ffcm_batch;
int ffcm_batch_create(ffcm_batch **out, int n_systems, const ffcm_params *params, int n_rhs);
int ffcm_batch_destroy(ffcm_batch *batch);
const char *ffcm_batch_last_error(const ffcm_batch *batch); /* batch may be NULL after a create failure */
/* Device pointers; enqueued on the batch stream (ffcm_batch_get_stream), not synchronised.
 * A call with pointers other than the previous call's re-captures the graph. */
int ffcm_batch_mobility_device(ffcm_batch *batch, void *const *V, const void *const *Y, const void *const *F);
/* Host arrays; synchronous. */
int ffcm_batch_mobility(ffcm_batch *batch, void *const *V, const void *const *Y, const void *const *F);
int ffcm_batch_get_stream(ffcm_batch *batch, void **stream);
int ffcm_batch_synchronize(ffcm_batch *batch);
/* lanes[0] = execution in use: l > 0 per-system launches on l stream lanes (a captured graph),
 * -l grouped launches (one launch per pipeline step over a lane's systems) on l lanes, 0 before
 * the first apply; lanes[1] = lanes allocated; graph_nodes = nodes of the captured graph, or
 * the grouped launches of one apply; tune_ms[8]: time of each candidate timed on the first
 * apply, [l-1] per-system on l lanes and [3+l] grouped on l lanes (0 where not a candidate). */
int ffcm_batch_get_info(const ffcm_batch *batch, int32_t *lanes, int64_t *graph_nodes, double *tune_ms);

/* Pullback of mobility! for the ChainRulesCore rrule (no counterpart in the reference; its
 * semantics are the derivative of mobility!, src/mobility.jl:23-45). For L = <G, M(Y) F>:
 *   dF = M(Y) G   (M is symmetric, src/mobility.jl:157-162)
 *   dY = dL/dY    (gradient-of-Gaussian stencil weights of src/stencil.jl:80-112 and the
 *                  r-derivatives of _correction_scalars, src/correct_velocities.jl:71-104)
 * The cell list and weights are recomputed from Y. All five arrays are host 3xN in T; dF and
 * dY are overwritten. Synchronous. dY is the derivative of the smooth pieces: it does not see
 * the jumps where a stencil anchor round(Y/h) switches or a pair crosses r = R_c. */
int ffcm_mobility_vjp(ffcm_ctx *ctx, void *dF, void *dY, const void *Y, const void *F, const void *G,
                      int64_t N);

/* Device-resident form of ffcm_mobility_vjp (device pointers, enqueued, not synchronised). */
int ffcm_mobility_vjp_device(ffcm_ctx *ctx, void *dF, void *dY, const void *Y, const void *F, const void *G,
                             int64_t N);

/* The L1 step barriers as per-step entry points, reading and writing the context buffers as
 * the reference steps read and write the config's buffers. All are synchronous.
 *   ffcm_wrap_positions          wrap_positions!(Y_wrapped, Y_input, L)   src/cell_list.jl:22-31
 *   ffcm_assign_cells            assign_cells!(config, Y_wrapped)        src/cell_list.jl:193-238
 *   ffcm_sort_particles_by_cell  sort_particles_by_cell!(config, Y_wrapped, F_input)
 *                                                                       src/cell_list.jl:119-176
 *   ffcm_spread_forces           spread_forces!(config)                  src/spread_forces.jl:17-130
 *   ffcm_stokes_solve            stokes_solve!(config)                   src/stokes_solve.jl:21-110
 *   ffcm_interpolate_velocities  interpolate_velocities!(V_output, config) src/interpolate.jl:23-136
 *   ffcm_correct_velocities      correct_velocities!(V_output, config)   src/correct_velocities.jl:154-286
 */
int ffcm_wrap_positions(ffcm_ctx *ctx);
int ffcm_assign_cells(ffcm_ctx *ctx);
int ffcm_sort_particles_by_cell(ffcm_ctx *ctx);
int ffcm_spread_forces(ffcm_ctx *ctx);
int ffcm_stokes_solve(ffcm_ctx *ctx);
int ffcm_interpolate_velocities(ffcm_ctx *ctx);
int ffcm_correct_velocities(ffcm_ctx *ctx);

/* The two halves of stokes_solve! (src/stokes_solve.jl:21-42). The forward half runs the r2c
 * transforms of force_density and the projection _apply_inverse_stokes_kernel!
 * (src/stokes_solve.jl:70-110), leaving the projected spectrum in fluid_hat; the backward half
 * runs the unnormalised c2r transforms of fluid_hat into fluid_velocity
 * (src/stokes_solve.jl:38-40). */
int ffcm_stokes_forward_project(ffcm_ctx *ctx);
int ffcm_stokes_backward(ffcm_ctx *ctx);

/* Host <-> device copies of one context buffer in the reference layout (the `Array(x)` and
 * `copyto!` a Julia test applies to a config field, e.g. test/cuda/test_gpu_correct.jl:22-32).
 * `bytes` must equal ffcm_buffer_bytes. */
int ffcm_buffer_bytes(const ffcm_ctx *ctx, int buffer_id, size_t *bytes);
int ffcm_set_buffer(ffcm_ctx *ctx, int buffer_id, const void *host, size_t bytes);
int ffcm_get_buffer(ffcm_ctx *ctx, int buffer_id, void *host, size_t bytes);

/* The context's HIP stream (hipStream_t), for callers that order their own work with it (the
 * CUDA extension's task-local stream, docs/src/devdocs/gpu-architecture.md). */
int ffcm_get_stream(ffcm_ctx *ctx, void **stream);
int ffcm_synchronize(ffcm_ctx *ctx);

/* Launch-shape report: spread/interpolation tile extent and tile count, for the four FFT
 * kernels (x forward, y, z, x inverse) the (width, threads, LDS bytes) triple, and the
 * interpolation kernel an apply would run now (0 per particle, 1 column ring filled through
 * registers, 2 column ring filled by LDS DMA). Any pointer may be NULL. No reference
 * counterpart (profiling aid). */
int ffcm_get_info(const ffcm_ctx *ctx, int32_t tile[3], int32_t *n_tiles, int32_t fft_shape[12],
                  int32_t *interp_kernel);

/* Per-stage device time of the mobility applies, measured with HIP events recorded on the
 * context stream between the stages: 0 cell list (wrap, hash, sort, gather, tile binning),
 * 1 spread, 2 Stokes solve (all FFT passes), 3 interpolation, 4 correction. Enabling resets
 * the totals; ffcm_get_timing returns the summed milliseconds and the number of applies.
 * No reference counterpart (the reference times with BenchmarkTools,
 * benchmark/benchmarks.jl:103-187). */
int ffcm_set_timing(ffcm_ctx *ctx, int enable);
int ffcm_get_timing(ffcm_ctx *ctx, double *ms_per_stage, int n_stages, int64_t *calls);

/* Captured device-resident applies (no counterpart in the reference, whose CUDA extension
 * launches every kernel on each mobility! call). With enable != 0, ffcm_mobility_device runs
 * the first apply with a set of (V, Y, F) pointers as launched, captures the second into a
 * hipGraph, and replays that graph while the pointers stay the same; other pointers start over.
 * Scalar changes (ffcm_set_scalars, the per-barrier entry points) drop the graph; so does
 * enable = 0. Not used while stage timing is on. */
int ffcm_set_graph(ffcm_ctx *ctx, int enable);

#ifdef __cplusplus
}
#endif

#endif /* FFCM_HIP_H */
