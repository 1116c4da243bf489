// Internal declarations shared by the translation units of libffcm_hip.so.
//
// Device data layout (all buffers owned by one context, allocated once at ffcm_create):
//   Y_in, F_in, V_out   3xN T, the caller's column-major layout, original particle order
//   Y_wrapped           3xN T, original order
//   hash                N int32, 0-based cell index per particle (original order)
//   keys/perm           N int32 ping-pong pairs for the radix sort; after the sort
//                       keys_sorted[s] is the cell of slot s and perm[s] the 0-based
//                       original index of slot s
//   cell_start          C int32, 0-based first slot of each cell
//   cell_end            C int32, one past the last slot (equal to the reference's
//                       1-based inclusive end)
//   Y_sorted, F_sorted  N x T4 (x, y, z, 0) per slot, 16/32-byte aligned for vector loads
//   tile_*              per grid tile: particle-slot lists binned by stencil low corner
//   force, velocity     real grids, x-fastest with the three components interleaved per
//                       point: g[((z*My + y)*Mx + x)*3 + c]
//   spectrum            half spectrum, components interleaved per mode:
//                       s[((kz*My + ky)*Mhp + kx)*3 + c], complex T, rows padded to
//                       Mhp = Mh rounded up to 16 so every row starts on a 128-byte line
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ffcm_hip.h"

namespace ffcm {

template <typename T>
struct Vec;
template <>
struct Vec<float> {
  using c2 = float2;
  using v4 = float4;
};
template <>
struct Vec<double> {
  using c2 = double2;
  using v4 = double4;
};

// Per-context scalars in the working precision, passed by value to kernels. Constants the
// reference recomputes per call or per pair (kernel coefficients, self term, Gaussian
// prefactors) are computed once on the host in T with the reference's operation order.
template <typename T>
struct Scalars {
  int N;
  int M[3];
  int MG, half;
  int nc[3];
  int C;
  T L[3];
  T inv_cs[3];
  T h, inv_h, h3;
  // _modified_kernel_coefficients (src/stencil.jl:22-31)
  T a0, a2, inv_norm, inv_2S2;
  // _self_correction / _correction_scalars (src/correct_velocities.jl:21-104)
  T Rc2, sigma, Sigma, mu, delta;
  T s2, s4, S2, S4, dd, dq, fourpimu, eightpimu, two_sigma, two_Sigma, Ds_pref, DS_pref, four_s2,
      four_S2, two_S2, two_mu_S2, four_mu_S4;
  // adjoint: 1/Sigma^2 (gradient of the stencil weight), 1/sqrt(pi), 2 sigma^2 (derivatives of
  // the correction scalars)
  T inv_S2, inv_sqrtpi, two_s2;
  // _apply_inverse_stokes_kernel! (src/stokes_solve.jl:80): inv_M / mu
  T inv_mu_M;
};

// One axis of the hand-written FFT: the length and its radix sequence.
constexpr int kMaxStages = 16;
struct FftAxis {
  int n;
  int nstages;
  int radix[kMaxStages];
  unsigned long long packed;  // radix of stage s in bits [4s, 4s+4): read without indexing
};

// Launch geometry of one FFT pass (chosen at ffcm_create).
struct FftPass {
  int pencils;     // pencils per tile
  int tp;          // threads per pencil
  int threads;     // pencils * tp
  int lds_bytes;   // dynamic LDS
  int width;       // columns (x passes: row pairs) per tile
  int log2w;       // log2(width) for the y and z passes
  int tiles;       // tiles of the pass
  int blocks;      // persistent grid size of one grid (resident blocks, at most tiles)
  int resident;    // blocks resident on the device at once (occupancy)
  bool covered;    // the prefetch registers of the block's threads hold a whole tile
  bool wide = false;  // z pass: the wide kernel (more than 512 threads)
};

// Grid tiling shared by the spread and interpolation kernels.
struct Tiling {
  int T[3];     // tile extent per axis (divides M on that axis)
  int nt[3];    // tiles per axis
  int span[3];  // tiles one stencil can reach per axis
  int ntiles;
  int cand_cap;  // contributors staged in LDS per batch
  int lds_bytes;
  int mfma;      // spread on the matrix cores (Float32, one right-hand side): 1 M_G = 8, 2 M_G 9..16
  int blocks;    // persistent grid of the matrix-core spread
  int wave_ranks;  // binning ranks slots with one atomic per tile and wave (dense suspensions)
};

// Launch plan of the column-streamed interpolation (a block per column of spread tiles and
// run of tile layers).
struct RingPlan {
  int seg_layers;  // tile layers per block
  int nseg;        // blocks per column
  int ncols;       // tile columns, nt[0] * nt[1]
  int patch;       // columns visited in 8 x 4 patches
  int lds_bytes;   // dynamic LDS
  int blocks;      // grid size, a multiple of 8
  int threads;     // threads per block (512 or 1024)
  int dma;         // ring filled by LDS DMA: planes per step (2, 4, 8), or 0 for register staging
};

struct DevBuffers {
  void *Y_in, *F_in, *V_out, *Y_wrapped;
  void *G_in, *G_sorted, *dY_out;  // adjoint: upstream gradient (3xN, N x T4 sorted), dL/dY
  void *Y_sorted, *F_sorted;
  int *hash;
  int *keys_a, *keys_b, *perm_a, *perm_b;
  int *keys_sorted, *perm;  // aliases of the radix sort's final ping-pong buffers
  int *radix_hist;
  size_t radix_hist_len;
  int *cell_start, *cell_end;
  int4 *anchor;  // per slot: wrapped stencil low corner (x, y, z) and home tile (spread binning)
  int *tile_count, *tile_start, *tile_cursor, *tile_list;
  int *spread_queue;  // matrix-core spread: one tile counter per XCD
  int *scan_partials;
  void *force, *velocity;  // 3 * M3 T each, components interleaved
  void *spectrum;          // 3 * Mz * My * Mhp complex T, components interleaved
  void *twiddle_all;       // the three axes' twiddles, packed
  void *twiddle[3];        // forward twiddles exp(-2 pi i m / n) per axis, complex T (into twiddle_all)
  void *k_x, *k_y, *k_z;
  int *neighbor_map;  // device copy of the caller's half-shell map (ffcm_assemble contexts)
  int *sort_scratch;  // CellBuffers.counting_sort_scratch: C int32 (the radix sort keeps its own workspace)
  void *pair_table;   // Float32 contexts: cubic pieces over r^2 of the pair scalars A, B, then of A'/r, B'/r
                      // (2 x kPairTableIntervals x 8 floats)
};

// Scalar groups of a context (ffcm_assemble contexts learn them call by call).
enum {
  kScalarsL = 1,       // L
  kScalarsCells = 2,   // inv_cell_size (num_cells are fixed at assembly)
  kScalarsKernel = 4,  // sigma, Sigma, h, inv_h
  kScalarsMu = 8,      // mu
  kScalarsPair = 16,   // a, R_c
  kScalarsAll = 31
};

}  // namespace ffcm

namespace ffcm {
struct Recorder;
}

struct ffcm_ctx {
  int dtype;
  ffcm_params host;  // copy of the creation parameters (pointers cleared)
  int64_t N;
  int64_t M3;
  int Mh;   // Mx / 2 + 1
  int Mhp;  // Mh rounded up to a multiple of 16 (row stride of the spectrum)
  int64_t C;
  int key_bits;
  hipStream_t stream;
  int device;
  ffcm::DevBuffers d;
  ffcm::FftAxis fft[3];
  ffcm::FftPass pass_xf, pass_y, pass_z, pass_xi;
  ffcm::Tiling tiling;
  bool tiles_valid;      // tile lists match the current Y_sorted
  bool positions_valid;  // the cell list, slot positions and tiles are those of the last ffcm_set_positions
  void *pinned[3];       // page-locked staging of the host entry points: Y (3N), F and V (3N x nrhs)
  int scalar_groups;     // kScalars* bits known (all, except for ffcm_assemble contexts)
  int fft_state;         // per-component transform sequence of the seam calls (0: none)
  bool timing;
  int lds_limit;  // LDS bytes one workgroup may use
  void *events[16];
  double stage_ms[8];
  int64_t stage_calls;
  int nrhs;              // right-hand sides the force, spectrum and velocity buffers hold
  bool borrowed_grids;   // force, velocity, spectrum belong to a batch (not freed here)
  bool borrowed_stream;  // stream belongs to a batch
  std::string err;
  void *scalars;  // ffcm::Scalars<T>*
  ffcm::Recorder *rec = nullptr;  // when set, launches are recorded for a grouped batch, not run
  // ffcm_set_graph: device-resident applies replayed from a captured graph
  bool use_graph = false;
  int graph_state = 0;             // 0: no apply yet; 1: one uncaptured apply with graph_key; 2: graph
  hipGraphExec_t graph = nullptr;  // the captured apply of graph_key
  const void *graph_key[3] = {nullptr, nullptr, nullptr};
};

namespace ffcm {
// Releases a context's captured apply (after its stream has finished with it).
void drop_graph(ffcm_ctx *c);

// Pair scalars A(r), B(r) of the correction as cubic pieces in s = r^2 over [0, R_c^2]: interval
// i holds (a0, a1, a2, a3, b0, b1, b2, b3) with A = a0 + t (a1 + t (a2 + t a3)), t = s / ds - i.
constexpr int kPairTableIntervals = 256;
// Writes the table of a Float32 context's scalars to its device buffer (nothing for Float64
// contexts or contexts without the buffer). Called whenever the scalars change.
int upload_pair_table(ffcm_ctx *c);
}

namespace ffcm {

// Stage launchers, instantiated for float and double.
template <typename T>
hipError_t launch_wrap(ffcm_ctx *c, const T *Y);
template <typename T>
hipError_t launch_assign_cells(ffcm_ctx *c);
template <typename T>
hipError_t launch_wrap_assign(ffcm_ctx *c, const T *Y);
// offsets: also build cell_start / cell_end (false when launch_gather_bin follows, which
// builds them).
hipError_t launch_sort_cells(ffcm_ctx *c, bool offsets = true);
// Radix sort geometry for n keys: keys per thread of a 256-thread tile, and tiles (blocks).
int sort_items(int64_t n);
int sort_blocks(int64_t n);
template <typename T>
hipError_t launch_gather(ffcm_ctx *c, const T *F);
// The apply's fused step after the sort: cell ranges, the slot-ordered gather and the spread
// binning counts (tile_count zeroed by launch_wrap_assign); launch_bin_tiles(c, true) follows.
template <typename T>
hipError_t launch_gather_bin(ffcm_ctx *c, const T *F);
// Gathers nrhs consecutive 3xN force-like arrays into slot order (nrhs x N x T4) with the
// current permutation.
template <typename T>
hipError_t launch_gather_forces(ffcm_ctx *c, const T *F, void *dst, int nrhs = 1);
template <typename T>
hipError_t launch_bin_tiles(ffcm_ctx *c, bool counted = false);
// Fs: slot-ordered forces (N x T4); nullptr means the context's F_sorted. grid: output force
// grid; nullptr means the context's force grid.
// R right-hand sides: forces Fs[r * N + slot], grids grid + r * 3 M^3.
template <typename T>
hipError_t launch_spread(ffcm_ctx *c, const void *Fs = nullptr, T *grid = nullptr, int R = 1);
template <typename T>
hipError_t launch_fft_forward_project(ffcm_ctx *c);
template <typename T>
hipError_t launch_fft_inverse(ffcm_ctx *c);
// Steps of the Stokes solve (bits): forward r2c transforms, projection, backward c2r transforms.
enum { kFftForward = 1, kFftInverse = 2, kFftProject = 8 };
template <typename T>
hipError_t launch_fft_steps(ffcm_ctx *c, int steps);
// out: velocity grid written by the solve; nullptr means the context's velocity grid. The
// force grid may be passed: it is fully read before the last pass writes.
// nb: number of consecutive grids (right-hand sides) transformed together.
template <typename T>
hipError_t launch_fft_stokes(ffcm_ctx *c, T *out = nullptr, int nb = 1);
// grid: interleaved velocity grid to interpolate; nullptr means the context's velocity grid.
template <typename T>
hipError_t launch_interpolate(ffcm_ctx *c, T *V, const T *grid = nullptr, int tag = 10000);
// Interpolation kernel an apply runs with the context's configuration and the environment:
// 0 per particle, 1 column ring filled through registers, 2 column ring filled by LDS DMA.
int interp_kernel_kind(const ffcm_ctx *c);
// Adjoint kernels: grid part of dL/dY from the two solved grids, and the pair correction's
// contribution to dL/dF (+=) and dL/dY (+=).
template <typename T>
hipError_t launch_interp_grad(ffcm_ctx *c, T *dY, const T *uF, const T *uG);
// Float32 grid part through the DMA ring of the interpolation; false (nothing launched) where
// the interpolation does not use that ring, else true with the launch status in *err.
bool launch_interp_grad_ring(ffcm_ctx *c, float *dY, const float *uF, const float *uG, hipError_t *err);
template <typename T>
hipError_t launch_correct_vjp(ffcm_ctx *c, T *dF, T *dY);
// Fs: slot-ordered forces; nullptr means the context's F_sorted.
template <typename T>
hipError_t launch_correct(ffcm_ctx *c, T *V, const void *Fs = nullptr);
// R right-hand sides at once: Fs[r * N + slot] forces, V + r * 3N velocities.
template <typename T>
hipError_t launch_correct_multi(ffcm_ctx *c, T *V, const void *Fs, int R);

// Entry-point plumbing shared with seam.cpp: argument check and error reset, error message,
// stream synchronisation, and the host build of the half-shell neighbour map.
int call_begin(ffcm_ctx *c);
int call_fail(ffcm_ctx *c, int code, const std::string &msg);
int call_sync(ffcm_ctx *c);
std::vector<int32_t> host_neighbor_map(const int32_t nc[3]);

// Context creation shared by ffcm_create, ffcm_create_multi and the batch runtime: nrhs is the
// number of right-hand sides the buffers hold; without alloc_grids the force, velocity and
// spectrum pointers are left for the owner (a batch) to set. Errors go to ffcm_last_error(NULL).
int create_context(ffcm_ctx **out, const ffcm_params *params, int nrhs, bool alloc_grids);
// Enqueues the mobility apply for R right-hand sides (device pointers) on c->stream; only
// kernel launches, so it can be captured into a graph. Returns an FFCM_* code.
int enqueue_mobility(ffcm_ctx *c, void *V, const void *Y, const void *F, int R);
// Grid buffer sizes (bytes) of one right-hand side: force/velocity grid and spectrum.
size_t grid_bytes(const ffcm_ctx *c);
size_t spectrum_bytes(const ffcm_ctx *c);

// Context planning and buffer placement, shared by ffcm_create and the allocation-free entry
// point: plan_context fills the launch plans from c->host (no allocation); buffer_list names
// every device buffer with its size; bind_aliases points the derived pointers after placement;
// twiddle_table builds the host twiddles the FFT reads.
struct BufReq {
  void **slot;
  size_t bytes;
  const char *name;
};
template <typename T>
int plan_context(ffcm_ctx *c);
template <typename T>
void buffer_list(ffcm_ctx *c, bool staging, std::vector<BufReq> *out);
void bind_aliases(ffcm_ctx *c);
template <typename T>
void twiddle_table(const ffcm_ctx *c, std::vector<T> *out);
// Parameter checks of ffcm_create (FFCM_OK or FFCM_ERR_ARGUMENT with *msg set) and the
// kernels' scalars from the parameters.
int validate_params(const ffcm_params *p, std::string *msg);
template <typename T>
void fill_scalars(const ffcm_params *p, Scalars<T> *s);

// FFT planning helpers (host).
bool plan_axis(int n, FftAxis *ax);
template <typename T>
void plan_passes(ffcm_ctx *c, int lds_budget);
// Grid tiling (host); false when no tiling fits the LDS budget.
template <typename T>
bool plan_tiling(ffcm_ctx *c, int lds_budget);

hipError_t launch_fill_i32(hipStream_t s, int *p, int64_t n, int v);
template <typename T>
hipError_t launch_fill(hipStream_t s, T *p, int64_t n, T v);
// Exclusive scan of n int32 counts into out[0..n]; out[n] receives the total.
hipError_t launch_exclusive_scan(hipStream_t s, const int *in, int *out, int64_t n, int *partials);
hipError_t launch_exclusive_scan(ffcm_ctx *c, int tag, const int *in, int *out, int64_t n, int *partials);
size_t scan_partials_needed(int64_t n);

}  // namespace ffcm
