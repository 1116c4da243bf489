// C ABI of libffcm_hip.so: context creation (parameter validation, buffer allocation, FFT and
// tiling plans), the mobility apply, the per-step entry points and the buffer copies.
//
// Reference seams replaced here (paths relative to the ForceCouplingMethod.jl source tree):
//   _assemble_gpu_buffers                     src/config.jl:461-464, :495-502
//   _validate_config_parameters               src/config.jl:267-312 (checks restated)
//   mobility! + _stage_mobility_io! /
//   _retrieve_mobility_output!                src/mobility.jl:23-56, ext/.../mobility.jl:10-23
//   the L1 step barriers                      src/cell_list.jl, spread_forces.jl, stokes_solve.jl,
//                                             interpolate.jl, correct_velocities.jl
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "ffcm_internal.hpp"
#include "group.hpp"

using namespace ffcm;

template <typename T>
using T4s = typename Vec<T>::v4;

namespace {

thread_local std::string g_create_error;

constexpr int kStageCount = 5;
constexpr int kTimingSlots = 32;

struct Timing {
  hipEvent_t ev[kTimingSlots][kStageCount + 1];
  int pending[kTimingSlots];
  int head;
  double ms[kStageCount];
  int64_t calls;
};

int fail(ffcm_ctx* c, int code, const std::string& msg) {
  if (c)
    c->err = msg;
  else
    g_create_error = msg;
  return code;
}

int hip_fail(ffcm_ctx* c, hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(c, FFCM_ERR_DEVICE, m);
}

#define FFCM_HIP(ctx, expr)                                    \
  do {                                                         \
    hipError_t _e = (expr);                                    \
    if (_e != hipSuccess) return hip_fail((ctx), _e, #expr);   \
  } while (0)

bool smooth13(int n) {
  for (int f : {2, 3, 5, 7, 11, 13})
    while (n % f == 0) n /= f;
  return n == 1;
}

}  // namespace

// The checks of _validate_config_parameters (src/config.jl:267-312) plus the isotropy check of
// _derive_scalars (src/config.jl:341-348), in the same order, then this backend's limits.
int ffcm::validate_params(const ffcm_params* p, std::string* msg) {
  if (!p) {
    *msg = "params must not be NULL";
    return FFCM_ERR_ARGUMENT;
  }
  if (p->dtype != FFCM_F32 && p->dtype != FFCM_F64) {
    *msg = "dtype must be FFCM_F32 or FFCM_F64";
    return FFCM_ERR_ARGUMENT;
  }
  const double eps = p->dtype == FFCM_F32 ? (double)std::numeric_limits<float>::epsilon()
                                          : std::numeric_limits<double>::epsilon();
  auto arg = [&](bool ok, const std::string& m) {
    if (!ok) *msg = m;
    return ok;
  };
  if (!arg(p->R_c > 0, "R_c must be positive")) return FFCM_ERR_ARGUMENT;
  if (!arg(p->L[0] > 0 && p->L[1] > 0 && p->L[2] > 0, "L components must be positive")) return FFCM_ERR_ARGUMENT;
  const double minL = std::fmin(p->L[0], std::fmin(p->L[1], p->L[2]));
  if (!arg(p->R_c <= minL / 2, "R_c must be at most half the smallest box length")) return FFCM_ERR_ARGUMENT;
  if (!arg(p->N > 0, "N must be positive")) return FFCM_ERR_ARGUMENT;
  if (!arg(p->a == 1.0, "non-unit particle radius not yet implemented; only a = 1 is supported"))
    return FFCM_ERR_ARGUMENT;
  if (!arg(p->Sigma >= p->sigma, "kernel widths ratio must be at least 1")) return FFCM_ERR_ARGUMENT;
  if (!arg(p->M_G >= 2, "M_G must be at least 2")) return FFCM_ERR_ARGUMENT;
  if (!arg(p->M[0] >= 1 && p->M[1] >= 1 && p->M[2] >= 1, "num_grid_points components must each be >= 1"))
    return FFCM_ERR_ARGUMENT;
  const int minM = std::min(p->M[0], std::min(p->M[1], p->M[2]));
  if (!arg(p->M_G <= minM, "M_G must not exceed the grid on any axis")) return FFCM_ERR_ARGUMENT;
  if (!arg(p->mu > 0, "viscosity must be positive")) return FFCM_ERR_ARGUMENT;
  const double h0 = p->L[0] / p->M[0], h1 = p->L[1] / p->M[1], h2 = p->L[2] / p->M[2];
  const double rel = std::sqrt(eps);
  if (!arg(std::fabs(h1 - h0) <= rel * h0 * 1.0000001 && std::fabs(h2 - h0) <= rel * h0 * 1.0000001,
           "anisotropic grid spacing not supported"))
    return FFCM_ERR_ARGUMENT;
  for (int i = 0; i < 3; ++i)
    if (!arg(p->num_cells[i] >= 3, "num_cells must be at least 3 on every axis")) return FFCM_ERR_ARGUMENT;
  if (!arg(!(p->h <= 0) && !(p->inv_h <= 0), "h and inv_h must be positive")) return FFCM_ERR_ARGUMENT;
  if (!arg(p->k_x && p->k_y && p->k_z, "k_x, k_y, k_z must be provided")) return FFCM_ERR_ARGUMENT;
  // backend limits
  // slot indices, 3 n + c element offsets and lane counts (up to 8 lanes per slot) are 32-bit
  if (!arg(p->N <= (int64_t)(1 << 26), "N above 2^26 (67,108,864) is not supported by this backend"))
    return FFCM_ERR_ARGUMENT;
  if (!arg(p->M_G <= 64, "M_G above 64 is not supported by this backend")) return FFCM_ERR_ARGUMENT;
  for (int i = 0; i < 3; ++i)
    if (!arg(smooth13(p->M[i]), "num_grid_points must factor into primes <= 13 for the HIP FFT"))
      return FFCM_ERR_ARGUMENT;
  const int64_t C = (int64_t)p->num_cells[0] * p->num_cells[1] * p->num_cells[2];
  if (!arg(C < (int64_t)(1 << 30), "too many cells")) return FFCM_ERR_ARGUMENT;
  return FFCM_OK;
}

namespace {

}  // namespace

template <typename T>
void ffcm::fill_scalars(const ffcm_params* p, Scalars<T>* s) {
  const T pi = (T)3.141592653589793238462643383279502884;
  s->N = (int)p->N;
  for (int i = 0; i < 3; ++i) {
    s->M[i] = p->M[i];
    s->nc[i] = p->num_cells[i];
    s->L[i] = (T)p->L[i];
    s->inv_cs[i] = (T)p->inv_cell_size[i];
  }
  s->MG = p->M_G;
  s->half = p->M_G / 2;
  s->C = p->num_cells[0] * p->num_cells[1] * p->num_cells[2];
  s->h = (T)p->h;
  s->inv_h = (T)p->inv_h;
  s->h3 = s->h * s->h * s->h;
  const T sigma = (T)p->sigma, Sigma = (T)p->Sigma, mu = (T)p->mu, a = (T)p->a;
  // _modified_kernel_coefficients (src/stencil.jl:22-31)
  {
    const T S2 = Sigma * Sigma;
    const T S4 = S2 * S2;
    const T d = sigma * sigma - S2;
    s->a0 = T(1) - T(3) * d / (T(2) * S2);
    s->a2 = d / (T(2) * S4);
    s->inv_norm = T(1) / std::sqrt(T(2) * pi * S2);
    s->inv_2S2 = T(1) / (T(2) * S2);
  }
  // _self_correction (src/correct_velocities.jl:21-27). Julia evaluates the integer powers
  // Sigma^5 and pi^3 of a Float32 in Float64 and rounds once; std::pow in double does the same.
  {
    const T d = sigma * sigma - Sigma * Sigma;
    const T Ssp = Sigma * std::sqrt(pi);
    const T S5 = (T)std::pow((double)Sigma, 5.0);
    const T pi3 = (T)std::pow((double)pi, 3.0);
    s->delta = T(1) / (T(6) * pi * mu * a) - T(1) / (T(6) * pi * mu * Ssp) + d / (T(12) * mu * (Ssp * Ssp * Ssp)) -
               (d * d) / (T(32) * mu * S5 * std::sqrt(pi3));
  }
  // _correction_scalars constants (src/correct_velocities.jl:71-104)
  s->sigma = sigma;
  s->Sigma = Sigma;
  s->mu = mu;
  s->Rc2 = (T)p->R_c * (T)p->R_c;
  s->s2 = sigma * sigma;
  s->s4 = s->s2 * s->s2;
  s->S2 = Sigma * Sigma;
  s->S4 = s->S2 * s->S2;
  s->dd = s->s2 - s->S2;
  s->dq = s->dd * s->dd / T(4);
  s->fourpimu = T(4) * pi * mu;
  s->eightpimu = T(8) * pi * mu;
  s->two_sigma = T(2) * sigma;
  s->two_Sigma = T(2) * Sigma;
  s->Ds_pref = std::pow(T(4) * pi * s->s2, -T(3) / T(2));
  s->DS_pref = std::pow(T(4) * pi * s->S2, -T(3) / T(2));
  s->four_s2 = T(4) * s->s2;
  s->four_S2 = T(4) * s->S2;
  s->two_S2 = T(2) * s->S2;
  s->two_mu_S2 = T(2) * mu * s->S2;
  s->four_mu_S4 = T(4) * mu * s->S4;
  // adjoint constants (derivatives of the stencil weight and of A, B)
  s->inv_S2 = T(1) / s->S2;
  s->inv_sqrtpi = T(1) / std::sqrt(pi);
  s->two_s2 = T(2) * s->s2;
  // stokes_solve! (src/stokes_solve.jl:28-30) and _apply_inverse_stokes_kernel! (:80)
  const int64_t Mtot = (int64_t)p->M[0] * p->M[1] * p->M[2];
  const T inv_M = T(1) / (T)Mtot;
  s->inv_mu_M = inv_M / mu;
}
template void ffcm::fill_scalars<float>(const ffcm_params*, Scalars<float>*);
template void ffcm::fill_scalars<double>(const ffcm_params*, Scalars<double>*);

namespace {

template <typename T>
Scalars<T>* make_scalars(const ffcm_params* p) {
  auto* s = new Scalars<T>();
  fill_scalars<T>(p, s);
  return s;
}

size_t tsize(const ffcm_ctx* c) { return c->dtype == FFCM_F32 ? 4 : 8; }

void free_all(ffcm_ctx* c) {
  DevBuffers& d = c->d;
  if (c->borrowed_grids) d.force = d.velocity = d.spectrum = nullptr;
  void* ptrs[] = {d.Y_in, d.F_in, d.V_out, d.G_in, d.G_sorted, d.dY_out, d.Y_wrapped, d.Y_sorted, d.F_sorted, d.hash, d.keys_a, d.keys_b,
                  d.perm_a, d.perm_b, d.radix_hist, d.cell_start, d.cell_end, d.tile_count, d.tile_start,
                  d.tile_cursor, d.tile_list, d.anchor, d.scan_partials, d.force, d.velocity, d.spectrum,
                  d.twiddle_all, d.k_x, d.k_y, d.k_z, d.neighbor_map, d.sort_scratch, d.pair_table, d.spread_queue};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  std::memset(&c->d, 0, sizeof(c->d));
  for (void*& q : c->pinned) {
    if (q) (void)hipHostFree(q);
    q = nullptr;
  }
}

}  // namespace

// Forward twiddles exp(-2 pi i m / n), m < n, of the three axes: computed in double and
// rounded to T, packed axis after axis (2 n T values per axis).
template <typename T>
void ffcm::twiddle_table(const ffcm_ctx* c, std::vector<T>* out) {
  out->clear();
  for (int axis = 0; axis < 3; ++axis) {
    const int n = c->fft[axis].n;
    for (int m = 0; m < n; ++m) {
      const double ang = 2.0 * 3.141592653589793238462643383279502884 * (double)m / (double)n;
      out->push_back((T)std::cos(ang));
      out->push_back((T)-std::sin(ang));
    }
  }
}
template void ffcm::twiddle_table<float>(const ffcm_ctx*, std::vector<float>*);
template void ffcm::twiddle_table<double>(const ffcm_ctx*, std::vector<double>*);

// Launch plans of a context from its parameters (c->host, c->nrhs, c->device): the FFT
// passes, the spread tiling and the derived sizes. No allocation.
template <typename T>
int ffcm::plan_context(ffcm_ctx* c) {
  const ffcm_params* p = &c->host;
  const int Mx = p->M[0];
  c->Mh = Mx / 2 + 1;
  c->Mhp = (c->Mh + 15) / 16 * 16;
  c->M3 = (int64_t)Mx * p->M[1] * p->M[2];
  c->C = (int64_t)p->num_cells[0] * p->num_cells[1] * p->num_cells[2];
  c->key_bits = 1;
  while ((int64_t(1) << c->key_bits) < c->C) ++c->key_bits;
  for (int a = 0; a < 3; ++a)
    if (!plan_axis(p->M[a], &c->fft[a])) return fail(c, FFCM_ERR_ARGUMENT, "grid size not 13-smooth");
  // LDS per block: the device limit (64 KiB unless the runtime reports otherwise)
  int lds_limit = 65536;
  if (hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess ||
      lds_limit <= 0)
    lds_limit = 65536;
  (void)hipGetLastError();
  c->lds_limit = lds_limit;
  plan_passes<T>(c, lds_limit);
  // spread: 52 KiB per block holds a 16x16x8 tile and 64 staged contributors with three
  // 512-thread blocks resident per CU (measured best among 38..108 KiB and tile shapes from
  // 8x8x8 to 64x32x1: 0.19 ms at 256^3, 2.4-2.5 ms at 512^3); a context for several right-hand
  // sides accumulates up to four per pass and takes 76 KiB
  int spread_budget = std::min(lds_limit - 1024, (c->nrhs > 1 ? 76 : 52) * 1024);
  if (const char* e = std::getenv("FFCM_SPREAD_LDS")) spread_budget = std::min(lds_limit - 1024, std::atoi(e));
  // wide stencils (large M_G) or Float64 may need the whole LDS of a workgroup
  if (!plan_tiling<T>(c, spread_budget) && !plan_tiling<T>(c, lds_limit - 1024))
    return fail(c, FFCM_ERR_ARGUMENT, "no spread tiling of this grid fits the LDS of one workgroup");
  int maxlds = std::max(std::max(c->pass_xf.lds_bytes, c->pass_y.lds_bytes), c->pass_z.lds_bytes);
  if (maxlds > lds_limit) return fail(c, FFCM_ERR_ARGUMENT, "grid too large for the LDS-staged FFT");
  if (!c->pass_xf.covered || !c->pass_xi.covered || !c->pass_y.covered || !c->pass_z.covered)
    return fail(c, FFCM_ERR_ARGUMENT, "grid too large for the FFT tiles");
  return FFCM_OK;
}
template int ffcm::plan_context<float>(ffcm_ctx*);
template int ffcm::plan_context<double>(ffcm_ctx*);

// Every device buffer of a planned context, as (slot, bytes, name). `staging` adds the
// caller-order staging buffers of the host entry points and the pullback.
template <typename T>
void ffcm::buffer_list(ffcm_ctx* c, bool staging, std::vector<BufReq>* out) {
  DevBuffers& d = c->d;
  const int64_t N = c->N;
  const size_t R = (size_t)c->nrhs;
  const size_t ts = sizeof(T);
  const int My = c->host.M[1], Mz = c->host.M[2];
  const int nsort_blocks = sort_blocks(N);
  d.radix_hist_len = (size_t)256 * nsort_blocks;
  const size_t scan_n = std::max<size_t>(d.radix_hist_len, (size_t)c->tiling.ntiles);
  out->clear();
  auto add = [&](void* slot, size_t bytes, const char* name) { out->push_back({(void**)slot, bytes, name}); };
  if (staging) {
    add(&d.Y_in, 3 * N * ts, "Y_input");
    add(&d.F_in, 3 * N * R * ts, "F_input");
    add(&d.V_out, 3 * N * R * ts, "V_output");
    add(&d.G_in, 3 * N * ts, "G_input");
    add(&d.G_sorted, 4 * N * ts, "G_sorted");
    add(&d.dY_out, 3 * N * ts, "dY_output");
  }
  add(&d.Y_wrapped, 3 * N * ts, "Y_wrapped");
  add(&d.Y_sorted, 4 * N * ts, "Y_sorted");
  add(&d.F_sorted, 4 * N * R * ts, "F_sorted");
  add(&d.hash, N * 4, "cell_hash");
  add(&d.keys_a, N * 4, "keys_a");
  add(&d.keys_b, N * 4, "keys_b");
  add(&d.perm_a, N * 4, "perm_a");
  add(&d.perm_b, N * 4, "perm_b");
  add(&d.radix_hist, (2 * d.radix_hist_len + 1) * 4, "radix_hist");
  add(&d.cell_start, c->C * 4, "cell_start");
  add(&d.cell_end, c->C * 4, "cell_end");
  if (staging) add(&d.sort_scratch, c->C * 4, "counting_sort_scratch");
  if (staging && sizeof(T) == 4) add(&d.pair_table, (size_t)kPairTableIntervals * 16 * sizeof(float), "pair_table");
  add(&d.tile_count, (size_t)c->tiling.ntiles * 4, "tile_count");
  add(&d.spread_queue, 8 * 4, "spread_queue");
  add(&d.tile_start, ((size_t)c->tiling.ntiles + 1) * 4, "tile_start");
  add(&d.tile_cursor, 2 * N * 4, "tile_cursor");
  add(&d.tile_list, N * 4, "tile_list");
  add(&d.anchor, N * sizeof(int4), "anchor");
  add(&d.scan_partials, (scan_partials_needed((int64_t)scan_n) + 1) * 4, "scan_partials");
  if (!c->borrowed_grids) {
    add(&d.force, 3 * (size_t)c->M3 * R * ts, "force_density");
    add(&d.velocity, 3 * (size_t)c->M3 * R * ts, "fluid_velocity");
    add(&d.spectrum, (size_t)2 * 3 * c->Mhp * My * Mz * R * ts, "fluid_hat");
  }
  add(&d.twiddle_all, 2 * (size_t)(c->fft[0].n + c->fft[1].n + c->fft[2].n) * ts, "twiddles");
  add(&d.k_x, (size_t)c->Mh * ts, "k_x");
  add(&d.k_y, (size_t)My * ts, "k_y");
  add(&d.k_z, (size_t)Mz * ts, "k_z");
}
template void ffcm::buffer_list<float>(ffcm_ctx*, bool, std::vector<BufReq>*);
template void ffcm::buffer_list<double>(ffcm_ctx*, bool, std::vector<BufReq>*);

// Points the per-axis twiddle pointers into the packed table and the sort's aliases at their
// first buffers (after the buffers are placed).
void ffcm::bind_aliases(ffcm_ctx* c) {
  DevBuffers& d = c->d;
  const size_t ts = c->dtype == FFCM_F32 ? 4 : 8;
  unsigned char* tw = (unsigned char*)d.twiddle_all;
  for (int a = 0; a < 3; ++a) {
    d.twiddle[a] = tw;
    tw += 2 * (size_t)c->fft[a].n * ts;
  }
  d.keys_sorted = d.keys_a;
  d.perm = d.perm_a;
}

namespace {

template <typename T>
int create_typed(ffcm_ctx* c) {
  const ffcm_params* p = &c->host;
  if (int rc = plan_context<T>(c)) return rc;
  std::vector<BufReq> reqs;
  buffer_list<T>(c, true, &reqs);
  for (const BufReq& r : reqs) {
    void* q = nullptr;
    const size_t bytes = r.bytes ? r.bytes : sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      free_all(c);
      return fail(c, FFCM_ERR_MEMORY,
                  std::string("device allocation of ") + r.name + " (" + std::to_string(bytes) + " bytes) failed");
    }
    *r.slot = q;
  }
  bind_aliases(c);
  std::vector<T> tw;
  twiddle_table<T>(c, &tw);
  FFCM_HIP(c, hipMemcpy(c->d.twiddle_all, tw.data(), tw.size() * sizeof(T), hipMemcpyHostToDevice));
  FFCM_HIP(c, hipMemcpy(c->d.k_x, p->k_x, sizeof(T) * c->Mh, hipMemcpyHostToDevice));
  FFCM_HIP(c, hipMemcpy(c->d.k_y, p->k_y, sizeof(T) * p->M[1], hipMemcpyHostToDevice));
  FFCM_HIP(c, hipMemcpy(c->d.k_z, p->k_z, sizeof(T) * p->M[2], hipMemcpyHostToDevice));
  c->scalars = make_scalars<T>(p);
  c->host.k_x = c->host.k_y = c->host.k_z = nullptr;
  return upload_pair_table(c);
}

template <typename T>
Scalars<T>& S(ffcm_ctx* c) {
  return *(Scalars<T>*)c->scalars;
}

Timing* timing_of(ffcm_ctx* c) { return (Timing*)c->events[0]; }

void timing_collect(ffcm_ctx* c, int slot) {
  Timing* t = timing_of(c);
  if (!t || !t->pending[slot]) return;
  (void)hipEventSynchronize(t->ev[slot][kStageCount]);
  for (int s = 0; s < kStageCount; ++s) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, t->ev[slot][s], t->ev[slot][s + 1]) == hipSuccess) t->ms[s] += ms;
  }
  t->calls += 1;
  t->pending[slot] = 0;
}

// The six steps of mobility! (src/mobility.jl:36-42) enqueued on the context stream, for R
// right-hand sides sharing Y: V and F hold R consecutive 3xN blocks. The cell list, the tile
// binning and the stencil anchors are built once; spread, interpolation and correction run
// per right-hand side; the R force grids are transformed together. Y == nullptr reuses the
// cell list of the last ffcm_set_positions (steps 1-2 skipped; only the forces are gathered).
template <typename T>
int run_pipeline(ffcm_ctx* c, T* V, const T* Y, const T* F, int R = 1) {
  Timing* t = c->timing ? timing_of(c) : nullptr;
  int slot = 0;
  if (t) {
    slot = t->head;
    t->head = (t->head + 1) % kTimingSlots;
    timing_collect(c, slot);
    FFCM_HIP(c, hipEventRecord(t->ev[slot][0], c->stream));
  }
  const size_t n3 = 3 * (size_t)c->N, g3 = 3 * (size_t)c->M3;
  T4s<T>* Fs = (T4s<T>*)c->d.F_sorted;
  T* force = (T*)c->d.force;
  T* vel = (T*)c->d.velocity;
  if (Y) {
    c->positions_valid = false;
    FFCM_HIP(c, launch_wrap_assign<T>(c, Y));
    FFCM_HIP(c, launch_sort_cells(c, false));
    FFCM_HIP(c, launch_gather_bin<T>(c, F));
    if (R > 1) FFCM_HIP(c, launch_gather_forces<T>(c, F + n3, Fs + c->N, R - 1));
    FFCM_HIP(c, launch_bin_tiles<T>(c, true));
  } else {
    FFCM_HIP(c, launch_gather_forces<T>(c, F, Fs, R));
  }
  if (t) FFCM_HIP(c, hipEventRecord(t->ev[slot][1], c->stream));
  FFCM_HIP(c, launch_spread<T>(c, Fs, force, R));
  if (t) FFCM_HIP(c, hipEventRecord(t->ev[slot][2], c->stream));
  FFCM_HIP(c, launch_fft_stokes<T>(c, vel, R));
  if (t) FFCM_HIP(c, hipEventRecord(t->ev[slot][3], c->stream));
  for (int r = 0; r < R; ++r) FFCM_HIP(c, launch_interpolate<T>(c, V + r * n3, vel + r * g3, kTagInterp + r));
  if (t) FFCM_HIP(c, hipEventRecord(t->ev[slot][4], c->stream));
  FFCM_HIP(c, R == 1 ? launch_correct<T>(c, V, Fs) : launch_correct_multi<T>(c, V, Fs, R));
  if (t) {
    FFCM_HIP(c, hipEventRecord(t->ev[slot][5], c->stream));
    t->pending[slot] = 1;
  }
  return FFCM_OK;
}

// Steps 1-2 of mobility! for positions Y alone (wrap, hash, sort, slot-ordered positions, the
// spread's tile binning), kept for later run_pipeline(c, V, nullptr, F) calls.
template <typename T>
int run_positions(ffcm_ctx* c, const T* Y) {
  c->positions_valid = false;
  FFCM_HIP(c, launch_wrap_assign<T>(c, Y));
  FFCM_HIP(c, launch_sort_cells(c, false));
  FFCM_HIP(c, launch_gather_bin<T>(c, (const T*)nullptr));
  FFCM_HIP(c, launch_bin_tiles<T>(c, true));
  c->positions_valid = true;
  return FFCM_OK;
}

// Vector-Jacobian product of mobility! for L = <G, M(Y) F>: dF = M(Y) G and dY = dL/dY. The cell
// list and the stencil weights are recomputed here from Y (nothing is kept from a forward call).
// u_F is solved into the velocity grid; u_G is solved into the force grid, which the solve has
// finished reading before its last pass writes.
template <typename T>
int run_vjp(ffcm_ctx* c, T* dF, T* dY, const T* Y, const T* F, const T* G) {
  T* uF = (T*)c->d.velocity;
  T* uG = (T*)c->d.force;
  c->positions_valid = false;
  FFCM_HIP(c, launch_wrap_assign<T>(c, Y));
  FFCM_HIP(c, launch_sort_cells(c, false));
  FFCM_HIP(c, launch_gather_bin<T>(c, F));
  FFCM_HIP(c, launch_gather_forces<T>(c, G, c->d.G_sorted));
  FFCM_HIP(c, launch_bin_tiles<T>(c, true));
  FFCM_HIP(c, launch_spread<T>(c, c->d.F_sorted));
  FFCM_HIP(c, launch_fft_stokes<T>(c, uF));
  FFCM_HIP(c, launch_spread<T>(c, c->d.G_sorted));
  FFCM_HIP(c, launch_fft_stokes<T>(c, uG));
  FFCM_HIP(c, launch_interpolate<T>(c, dF, uG));
  FFCM_HIP(c, launch_interp_grad<T>(c, dY, uF, uG));
  FFCM_HIP(c, launch_correct_vjp<T>(c, dF, dY));
  return FFCM_OK;
}

}  // namespace

int ffcm::enqueue_mobility(ffcm_ctx* c, void* V, const void* Y, const void* F, int R) {
  c->err.clear();
  if (c->dtype == FFCM_F32) return run_pipeline<float>(c, (float*)V, (const float*)Y, (const float*)F, R);
  return run_pipeline<double>(c, (double*)V, (const double*)Y, (const double*)F, R);
}

size_t ffcm::grid_bytes(const ffcm_ctx* c) { return 3 * (size_t)c->M3 * (c->dtype == FFCM_F32 ? 4 : 8); }
size_t ffcm::spectrum_bytes(const ffcm_ctx* c) {
  return (size_t)2 * 3 * c->Mhp * c->host.M[1] * c->host.M[2] * (c->dtype == FFCM_F32 ? 4 : 8);
}

namespace {

int check_ctx(ffcm_ctx* c) {
  if (!c) return FFCM_ERR_ARGUMENT;
  c->err.clear();
  (void)hipGetLastError();  // launch errors are checked per launch; drop any stale state
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, FFCM_ERR_DEVICE, "hipSetDevice failed");
  return FFCM_OK;
}

int sync(ffcm_ctx* c) {
  FFCM_HIP(c, hipStreamSynchronize(c->stream));
  return FFCM_OK;
}

// check_ctx for the compute entry points: every scalar must be known (a context built by
// ffcm_assemble learns them from ffcm_set_scalars or from the per-barrier calls).
int check_ready(ffcm_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  if (c->scalar_groups != ffcm::kScalarsAll)
    return fail(c, FFCM_ERR_ARGUMENT, "the config's scalars are not set: call ffcm_set_scalars after ffcm_assemble");
  c->fft_state = 0;
  return FFCM_OK;
}

// Host-to-host copy of `bytes` in 4 MiB pieces between the caller's pageable arrays and the
// page-locked buffers, each piece split over up to kCopyThreads threads (the caller and workers
// started once per call): one core's memcpy (about 33 GB/s) is slower than PCIe (about 57 GB/s
// each way), four reach about 75 GB/s. Pieces 0 .. ready are free to copy; the caller runs
// wait_piece(i) before making piece i free (none when all are free from the start) and
// piece_done(i) once piece i is copied, so transfers and copies overlap.
constexpr size_t kStagePiece = size_t(4) << 20;
constexpr int kCopyThreads = 4;

template <typename Wait, typename Done>
int staged_copy(void* dst, const void* src, size_t bytes, bool all_free, Wait wait_piece, Done piece_done) {
  const size_t np = (bytes + kStagePiece - 1) / kStagePiece;
  const int nt = bytes >= 2 * (size_t(1) << 20) ? kCopyThreads : 1;
  std::atomic<long> ready{all_free ? (long)np - 1 : -1L};
  std::unique_ptr<std::atomic<int>[]> done(new std::atomic<int>[np]);
  for (size_t i = 0; i < np; ++i) done[i].store(0, std::memory_order_relaxed);
  auto part = [&](size_t i, int t) {
    const size_t o = i * kStagePiece, n = std::min(kStagePiece, bytes - o);
    const size_t per = ((n + nt - 1) / nt + 63) & ~size_t(63);  // nt slices cover n
    const size_t b = std::min(n, per * (size_t)t), e = std::min(n, b + per);
    if (e > b) std::memcpy((unsigned char*)dst + o + b, (const unsigned char*)src + o + b, e - b);
  };
  std::thread th[kCopyThreads];
  for (int t = 1; t < nt; ++t)
    th[t] = std::thread([&, t] {
      for (size_t i = 0; i < np; ++i) {
        while (ready.load(std::memory_order_acquire) < (long)i) std::this_thread::yield();
        part(i, t);
        done[i].fetch_add(1, std::memory_order_release);
      }
    });
  int rc = FFCM_OK;
  for (size_t i = 0; i < np; ++i) {
    if (!all_free) {
      if (rc == FFCM_OK) rc = wait_piece(i);
      ready.store((long)i, std::memory_order_release);  // also on failure: the workers finish
    }
    part(i, 0);
    while (done[i].load(std::memory_order_acquire) < nt - 1) std::this_thread::yield();
    if (rc == FFCM_OK) rc = piece_done(i);
  }
  for (int t = 1; t < nt; ++t) th[t].join();
  return rc;
}

// Host -> device through the context's page-locked buffer `which`: each 4 MiB piece's DMA is
// enqueued as soon as the piece is in place, while the copy threads go on with the next ones.
int stage_in(ffcm_ctx* c, void* dev, const void* host, size_t bytes, int which) {
  unsigned char* pin = (unsigned char*)c->pinned[which];
  if (!pin) return fail(c, FFCM_ERR_INTERNAL, "no staging buffer");
  return staged_copy(
      pin, host, bytes, true, [](size_t) -> int { return FFCM_OK; },
      [&](size_t i) -> int {
        const size_t o = i * kStagePiece, n = std::min(kStagePiece, bytes - o);
        FFCM_HIP(c, hipMemcpyAsync((unsigned char*)dev + o, pin + o, n, hipMemcpyHostToDevice, c->stream));
        return FFCM_OK;
      });
}

// Device -> host through page-locked buffer `which`, in 4 MiB pieces: each piece is copied out
// of the page-locked buffer once its transfer has completed, while the next ones transfer.
// Returns with the stream drained.
int stage_out(ffcm_ctx* c, void* host, const void* dev, size_t bytes, int which) {
  unsigned char* pin = (unsigned char*)c->pinned[which];
  if (!pin) return fail(c, FFCM_ERR_INTERNAL, "no staging buffer");
  const size_t np = (bytes + kStagePiece - 1) / kStagePiece;
  std::vector<hipEvent_t> ev(np, nullptr);
  struct Guard {
    std::vector<hipEvent_t>& e;
    ~Guard() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } guard{ev};
  for (size_t i = 0; i < np; ++i) {
    const size_t o = i * kStagePiece, n = std::min(kStagePiece, bytes - o);
    FFCM_HIP(c, hipMemcpyAsync(pin + o, (const unsigned char*)dev + o, n, hipMemcpyDeviceToHost, c->stream));
    FFCM_HIP(c, hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
    FFCM_HIP(c, hipEventRecord(ev[i], c->stream));
  }
  if (int rc = staged_copy(
          host, pin, bytes, false,
          [&](size_t i) -> int {
            FFCM_HIP(c, hipEventSynchronize(ev[i]));
            return FFCM_OK;
          },
          [](size_t) -> int { return FFCM_OK; }))
    return rc;
  FFCM_HIP(c, hipStreamSynchronize(c->stream));
  return FFCM_OK;
}

// Neighbour map of _build_neighbor_map (src/cell_list.jl:258-289), 0-based.
std::vector<int32_t> neighbor_map(const int32_t nc[3]) {
  static const int off[13][3] = {{1, 0, 0},  {1, 1, 0},  {0, 1, 0},  {-1, 1, 0},  {1, 0, -1}, {1, 1, -1}, {0, 1, -1},
                                 {-1, 1, -1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}, {-1, 1, 1}, {0, 0, 1}};
  const int mx = nc[0], my = nc[1], mz = nc[2];
  std::vector<int32_t> m((size_t)13 * mx * my * mz);
  auto md = [](int a, int b) { return ((a % b) + b) % b; };
  for (int cz = 0; cz < mz; ++cz)
    for (int cy = 0; cy < my; ++cy)
      for (int cx = 0; cx < mx; ++cx) {
        size_t base = (size_t)13 * (cx + (cy + (size_t)cz * my) * mx);
        for (int k = 0; k < 13; ++k)
          m[base + k] = md(cx + off[k][0], mx) + (md(cy + off[k][1], my) + md(cz + off[k][2], mz) * my) * mx;
      }
  return m;
}

}  // namespace

extern "C" {

int ffcm_create(ffcm_ctx** out, const ffcm_params* params) { return ffcm_create_multi(out, params, 1); }

int ffcm_create_multi(ffcm_ctx** out, const ffcm_params* params, int max_rhs) {
  return ffcm::create_context(out, params, max_rhs, true);
}

int ffcm_mobility_multi_device(ffcm_ctx* c, int R, void* V, const void* Y, const void* F, int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "V, Y and F must be 3xN (xR) for the config's N");
  if (R < 1 || R > c->nrhs)
    return fail(c, FFCM_ERR_DIMENSION, "R must be between 1 and the context's max_rhs (" + std::to_string(c->nrhs) + ")");
  if (!V || !Y || !F) return fail(c, FFCM_ERR_ARGUMENT, "V, Y and F must not be NULL");
  if (c->dtype == FFCM_F32) return run_pipeline<float>(c, (float*)V, (const float*)Y, (const float*)F, R);
  return run_pipeline<double>(c, (double*)V, (const double*)Y, (const double*)F, R);
}

int ffcm_mobility_multi(ffcm_ctx* c, int R, void* V, const void* Y, const void* F, int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "V, Y and F must be 3xN (xR) for the config's N");
  if (R < 1 || R > c->nrhs)
    return fail(c, FFCM_ERR_DIMENSION, "R must be between 1 and the context's max_rhs (" + std::to_string(c->nrhs) + ")");
  if (!V || !Y || !F) return fail(c, FFCM_ERR_ARGUMENT, "V, Y and F must not be NULL");
  const size_t bytes = 3 * (size_t)c->N * tsize(c);
  if (int rc = stage_in(c, c->d.Y_in, Y, bytes, 0)) return rc;
  if (int rc = stage_in(c, c->d.F_in, F, bytes * R, 1)) return rc;
  int rc = c->dtype == FFCM_F32
               ? run_pipeline<float>(c, (float*)c->d.V_out, (const float*)c->d.Y_in, (const float*)c->d.F_in, R)
               : run_pipeline<double>(c, (double*)c->d.V_out, (const double*)c->d.Y_in, (const double*)c->d.F_in, R);
  if (rc) return rc;
  return stage_out(c, V, c->d.V_out, bytes * R, 2);
}

}  // extern "C"

namespace ffcm {
int create_context(ffcm_ctx** out, const ffcm_params* params, int nrhs, bool alloc_grids) {
  g_create_error.clear();
  if (!out) return fail(nullptr, FFCM_ERR_ARGUMENT, "out must not be NULL");
  *out = nullptr;
  if (nrhs < 1 || nrhs > 4096) return fail(nullptr, FFCM_ERR_ARGUMENT, "max_rhs must be between 1 and 4096");
  std::string msg;
  if (int rc = validate_params(params, &msg)) return fail(nullptr, rc, msg);
  // right-hand side r of slot s is element r N + s of the sorted forces (32-bit)
  if (params->N * (int64_t)nrhs > (int64_t(1) << 28))
    return fail(nullptr, FFCM_ERR_ARGUMENT, "N * max_rhs above 2^28 is not supported by this backend");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return fail(nullptr, FFCM_ERR_DEVICE, "no HIP device is available (gpu_acceleration needs an MI355X)");
  }
  (void)hipGetLastError();
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return fail(nullptr, FFCM_ERR_DEVICE, "hipGetDeviceProperties failed");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(nullptr, FFCM_ERR_DEVICE, std::string("device architecture ") + prop.gcnArchName +
                                              " is not gfx950; this library carries gfx950 code only");
  auto* c = new ffcm_ctx();
  c->dtype = params->dtype;
  c->host = *params;
  c->N = params->N;
  c->device = dev;
  std::memset(&c->d, 0, sizeof(c->d));
  std::memset(c->events, 0, sizeof(c->events));
  c->timing = false;
  c->scalars = nullptr;
  c->nrhs = nrhs;
  c->borrowed_grids = !alloc_grids;
  c->borrowed_stream = false;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(nullptr, FFCM_ERR_DEVICE, "hipStreamCreate failed");
  }
  c->positions_valid = false;
  c->scalar_groups = kScalarsAll;
  c->fft_state = 0;
  for (void*& q : c->pinned) q = nullptr;
  int rc = c->dtype == FFCM_F32 ? create_typed<float>(c) : create_typed<double>(c);
  if (rc == FFCM_OK && alloc_grids) {
    // page-locked staging for the host entry points (callers' arrays are pageable)
    const size_t n3 = 3 * (size_t)c->N * tsize(c);
    const size_t sizes[3] = {n3, n3 * (size_t)nrhs, n3 * (size_t)nrhs};
    for (int i = 0; i < 3 && rc == FFCM_OK; ++i)
      if (hipHostMalloc(&c->pinned[i], sizes[i], hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        c->pinned[i] = nullptr;
        rc = fail(c, FFCM_ERR_MEMORY, "page-locked host allocation for the staging buffers failed");
      }
  }
  if (rc != FFCM_OK) {
    g_create_error = c->err;
    free_all(c);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }
  *out = c;
  return FFCM_OK;
}

// The pair scalars of _correction_scalars (src/correct_velocities.jl:71-104) in double
// precision, from a Float32 context's sigma, Sigma and mu.
static void pair_scalars_f64(double r, double sigma, double Sigma, double mu, double* A, double* B) {
  const double pi = 3.14159265358979323846;
  const double s2 = sigma * sigma, S2 = Sigma * Sigma, s4 = s2 * s2, S4 = S2 * S2;
  const double dd = s2 - S2, dq = dd * dd / 4;
  const double E = std::erf(r / (2 * sigma)) - std::erf(r / (2 * Sigma));
  const double Ds = std::pow(4 * pi * s2, -1.5) * std::exp(-r * r / (4 * s2));
  const double DS = std::pow(4 * pi * S2, -1.5) * std::exp(-r * r / (4 * S2));
  const double r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r;
  *A = E * (1 / (8 * pi * mu * r) + s2 / (4 * pi * mu * r3)) - 2 * s4 / (mu * r2) * Ds +
       (2 * S4 / (mu * r2) + dd / mu * (1 + 2 * S2 / r2) - dq * (2 - r2 / (2 * S2)) / (2 * mu * S2)) * DS;
  *B = E * (1 / (8 * pi * mu * r3) - 3 * s2 / (4 * pi * mu * r5)) + 6 * s4 / (mu * r4) * Ds +
       (-6 * S4 / (mu * r4) - dd / (mu * r2) * (1 + 6 * S2 / r2) - dq / (4 * mu * S4)) * DS;
}

// A'(r) / r and B'(r) / r in double precision: the product rule over the three products of
// A and B (as the pullback's correction_scalars_d), divided by r. Both are even in r, so
// smooth in r^2 like A and B.
static void pair_derivs_over_r_f64(double r, double sigma, double Sigma, double mu, double* dAr, double* dBr) {
  const double pi = 3.14159265358979323846;
  const double s2 = sigma * sigma, S2 = Sigma * Sigma, s4 = s2 * s2, S4 = S2 * S2;
  const double dd = s2 - S2, dq = dd * dd / 4;
  const double r2 = r * r, ir = 1 / r, ir2 = ir * ir, ir3 = ir2 * ir, ir4 = ir2 * ir2, ir5 = ir4 * ir, ir6 = ir3 * ir3;
  const double es = std::exp(-r2 / (4 * s2)), eS = std::exp(-r2 / (4 * S2));
  const double E = std::erf(r / (2 * sigma)) - std::erf(r / (2 * Sigma));
  const double dE = (es / sigma - eS / Sigma) / std::sqrt(pi);
  const double Ds = std::pow(4 * pi * s2, -1.5) * es, DS = std::pow(4 * pi * S2, -1.5) * eS;
  const double dDs = -r * Ds / (2 * s2), dDS = -r * DS / (2 * S2);
  const double i8 = 1 / (8 * pi * mu), i4 = 1 / (4 * pi * mu), imu = 1 / mu;
  const double cA1 = i8 * ir + s2 * i4 * ir3, dcA1 = -i8 * ir2 - 3 * s2 * i4 * ir4;
  const double cA2 = -2 * s4 * imu * ir2, dcA2 = 4 * s4 * imu * ir3;
  const double cA3 = 2 * S4 * imu * ir2 + dd * imu * (1 + 2 * S2 * ir2) - dq * (2 - r2 / (2 * S2)) / (2 * mu * S2);
  const double dcA3 = -4 * S4 * imu * ir3 - 4 * dd * S2 * imu * ir3 + dq * r / (S2 * 2 * mu * S2);
  const double cB1 = i8 * ir3 - 3 * s2 * i4 * ir5, dcB1 = -3 * i8 * ir4 + 15 * s2 * i4 * ir6;
  const double cB2 = 6 * s4 * imu * ir4, dcB2 = -24 * s4 * imu * ir5;
  const double cB3 = -6 * S4 * imu * ir4 - dd * imu * ir2 * (1 + 6 * S2 * ir2) - dq / (4 * mu * S4);
  const double dcB3 = 24 * S4 * imu * ir5 + dd * imu * (2 * ir3 + 24 * S2 * ir5);
  *dAr = (dE * cA1 + E * dcA1 + dcA2 * Ds + cA2 * dDs + dcA3 * DS + cA3 * dDS) * ir;
  *dBr = (dE * cB1 + E * dcB1 + dcB2 * Ds + cB2 * dDs + dcB3 * DS + cB3 * dDS) * ir;
}

// Each interval's cubic interpolates A and B (in double) at the four Chebyshev points of the
// interval, which keeps every sample away from r = 0, where the terms of A and B cancel. The
// second half of the table holds the pieces of A'(r) / r and B'(r) / r for the pullback.
int ffcm::upload_pair_table(ffcm_ctx* c) {
  if (c->dtype != FFCM_F32 || !c->d.pair_table || !c->scalars) return FFCM_OK;
  const Scalars<float>& sc = *(const Scalars<float>*)c->scalars;
  const int K = kPairTableIntervals;
  std::vector<float> tab((size_t)K * 16, 0.f);
  const double smax = sc.Rc2, sigma = sc.sigma, Sigma = sc.Sigma, mu = sc.mu;
  if (smax > 0 && sigma > 0 && Sigma > 0 && mu > 0 && std::isfinite(smax) && std::isfinite(sigma) &&
      std::isfinite(Sigma) && std::isfinite(mu)) {
    const double ds = smax / K, pi = 3.14159265358979323846;
    double node[4], inv[4][4];
    for (int j = 0; j < 4; ++j) node[j] = (1 - std::cos((2 * j + 1) * pi / 8)) / 2;
    // inverse of the Vandermonde matrix V[j][k] = node[j]^k (Gauss-Jordan)
    double m[4][8];
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 8; ++k) m[j][k] = k < 4 ? std::pow(node[j], k) : (k - 4 == j ? 1.0 : 0.0);
    for (int col = 0; col < 4; ++col) {
      int piv = col;
      for (int r = col + 1; r < 4; ++r)
        if (std::fabs(m[r][col]) > std::fabs(m[piv][col])) piv = r;
      for (int k = 0; k < 8; ++k) std::swap(m[col][k], m[piv][k]);
      const double d = m[col][col];
      for (int k = 0; k < 8; ++k) m[col][k] /= d;
      for (int r = 0; r < 4; ++r)
        if (r != col) {
          const double f = m[r][col];
          for (int k = 0; k < 8; ++k) m[r][k] -= f * m[col][k];
        }
    }
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 4; ++k) inv[j][k] = m[j][k + 4];
    auto fit = [&](const double* v, float* out) {  // cubic coefficients from the node values
      for (int k = 0; k < 4; ++k) {
        double a = 0;
        for (int j = 0; j < 4; ++j) a += inv[k][j] * v[j];
        out[k] = (float)a;
      }
    };
    for (int i = 0; i < K; ++i) {
      double A[4], B[4], dA[4], dB[4];
      for (int j = 0; j < 4; ++j) {
        const double r = std::sqrt((i + node[j]) * ds);
        pair_scalars_f64(r, sigma, Sigma, mu, &A[j], &B[j]);
        pair_derivs_over_r_f64(r, sigma, Sigma, mu, &dA[j], &dB[j]);
      }
      fit(A, &tab[(size_t)i * 8]);
      fit(B, &tab[(size_t)i * 8 + 4]);
      fit(dA, &tab[(size_t)(K + i) * 8]);
      fit(dB, &tab[(size_t)(K + i) * 8 + 4]);
    }
  }
  // The context's stream does not wait for the null stream's copy, so an apply still in flight
  // would read a half-written table.
  if (c->stream) FFCM_HIP(c, hipStreamSynchronize(c->stream));
  FFCM_HIP(c, hipMemcpy(c->d.pair_table, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
  return FFCM_OK;
}

void drop_graph(ffcm_ctx* c) {
  if (c->graph) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipGraphExecDestroy(c->graph);
  }
  c->graph = nullptr;
  c->graph_state = 0;
  c->graph_key[0] = c->graph_key[1] = c->graph_key[2] = nullptr;
}

}  // namespace ffcm

extern "C" {

int ffcm_destroy(ffcm_ctx* c) {
  if (!c) return FFCM_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  if (Timing* t = timing_of(c)) {
    for (int i = 0; i < kTimingSlots; ++i)
      for (int s = 0; s <= kStageCount; ++s) (void)hipEventDestroy(t->ev[i][s]);
    delete t;
  }
  free_all(c);
  if (!c->borrowed_stream) (void)hipStreamDestroy(c->stream);
  if (c->scalars) {
    if (c->dtype == FFCM_F32)
      delete (Scalars<float>*)c->scalars;
    else
      delete (Scalars<double>*)c->scalars;
  }
  delete c;
  return FFCM_OK;
}

const char* ffcm_last_error(const ffcm_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int ffcm_mobility_device(ffcm_ctx* c, void* V, const void* Y, const void* F, int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "V, Y and F must each be 3xN for the config's N");
  if (!V || !Y || !F) return fail(c, FFCM_ERR_ARGUMENT, "V, Y and F must not be NULL");
  auto pipeline = [&]() {
    return c->dtype == FFCM_F32 ? run_pipeline<float>(c, (float*)V, (const float*)Y, (const float*)F)
                                : run_pipeline<double>(c, (double*)V, (const double*)Y, (const double*)F);
  };
  if (!c->use_graph || c->timing) return pipeline();
  // captured applies: the first apply with a set of pointers runs as launched (so that every
  // kernel has run outside a capture before it is captured), the second is captured and
  // replayed, later ones replay the graph
  const bool same = c->graph_key[0] == V && c->graph_key[1] == Y && c->graph_key[2] == F;
  if (same && c->graph_state == 2) {
    c->positions_valid = false;  // the replay rebuilds the cell list from Y, as run_pipeline does
    FFCM_HIP(c, hipGraphLaunch(c->graph, c->stream));
    return FFCM_OK;
  }
  if (!same || c->graph_state == 0) {
    drop_graph(c);
    c->graph_key[0] = V, c->graph_key[1] = Y, c->graph_key[2] = F;
    const int rc = pipeline();
    c->graph_state = rc == FFCM_OK ? 1 : 0;
    return rc;
  }
  FFCM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int rc = pipeline();
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc != FFCM_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    c->graph_state = 0;
    return rc != FFCM_OK ? rc : fail(c, FFCM_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
  }
  const hipError_t ei = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) {
    c->graph = nullptr;
    c->graph_state = 0;
    (void)hipGetLastError();
    return fail(c, FFCM_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  }
  c->graph_state = 2;
  FFCM_HIP(c, hipGraphLaunch(c->graph, c->stream));
  return FFCM_OK;
}

int ffcm_set_graph(ffcm_ctx* c, int enable) {
  if (int rc = check_ctx(c)) return rc;
  drop_graph(c);
  c->use_graph = enable != 0;
  return FFCM_OK;
}

int ffcm_mobility(ffcm_ctx* c, void* V, const void* Y, const void* F, int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "V, Y and F must each be 3xN for the config's N");
  if (!V || !Y || !F) return fail(c, FFCM_ERR_ARGUMENT, "V, Y and F must not be NULL");
  const size_t bytes = 3 * (size_t)c->N * tsize(c);
  if (int rc = stage_in(c, c->d.Y_in, Y, bytes, 0)) return rc;
  if (int rc = stage_in(c, c->d.F_in, F, bytes, 1)) return rc;
  int rc = c->dtype == FFCM_F32
               ? run_pipeline<float>(c, (float*)c->d.V_out, (const float*)c->d.Y_in, (const float*)c->d.F_in)
               : run_pipeline<double>(c, (double*)c->d.V_out, (const double*)c->d.Y_in, (const double*)c->d.F_in);
  if (rc) return rc;
  return stage_out(c, V, c->d.V_out, bytes, 2);
}

int ffcm_set_positions_device(ffcm_ctx* c, const void* Y, int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "Y must be 3xN for the config's N");
  if (!Y) return fail(c, FFCM_ERR_ARGUMENT, "Y must not be NULL");
  return c->dtype == FFCM_F32 ? run_positions<float>(c, (const float*)Y) : run_positions<double>(c, (const double*)Y);
}

int ffcm_set_positions(ffcm_ctx* c, const void* Y, int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "Y must be 3xN for the config's N");
  if (!Y) return fail(c, FFCM_ERR_ARGUMENT, "Y must not be NULL");
  if (int rc = stage_in(c, c->d.Y_in, Y, 3 * (size_t)c->N * tsize(c), 0)) return rc;
  int rc = c->dtype == FFCM_F32 ? run_positions<float>(c, (const float*)c->d.Y_in)
                                : run_positions<double>(c, (const double*)c->d.Y_in);
  if (rc) return rc;
  return sync(c);
}

int ffcm_apply_forces_device(ffcm_ctx* c, void* V, const void* F, int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "V and F must each be 3xN for the config's N");
  if (!V || !F) return fail(c, FFCM_ERR_ARGUMENT, "V and F must not be NULL");
  if (!c->positions_valid)
    return fail(c, FFCM_ERR_ARGUMENT, "no cached positions: call ffcm_set_positions first (a mobility apply, "
                                      "pullback or per-step call in between replaces the cell list)");
  return c->dtype == FFCM_F32 ? run_pipeline<float>(c, (float*)V, nullptr, (const float*)F)
                              : run_pipeline<double>(c, (double*)V, nullptr, (const double*)F);
}

int ffcm_apply_forces(ffcm_ctx* c, void* V, const void* F, int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "V and F must each be 3xN for the config's N");
  if (!V || !F) return fail(c, FFCM_ERR_ARGUMENT, "V and F must not be NULL");
  if (!c->positions_valid)
    return fail(c, FFCM_ERR_ARGUMENT, "no cached positions: call ffcm_set_positions first (a mobility apply, "
                                      "pullback or per-step call in between replaces the cell list)");
  const size_t bytes = 3 * (size_t)c->N * tsize(c);
  if (int rc = stage_in(c, c->d.F_in, F, bytes, 1)) return rc;
  int rc = c->dtype == FFCM_F32 ? run_pipeline<float>(c, (float*)c->d.V_out, nullptr, (const float*)c->d.F_in)
                                : run_pipeline<double>(c, (double*)c->d.V_out, nullptr, (const double*)c->d.F_in);
  if (rc) return rc;
  return stage_out(c, V, c->d.V_out, bytes, 2);
}

int ffcm_mobility_vjp_device(ffcm_ctx* c, void* dF, void* dY, const void* Y, const void* F, const void* G,
                             int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "dF, dY, Y, F and G must each be 3xN for the config's N");
  if (!dF || !dY || !Y || !F || !G) return fail(c, FFCM_ERR_ARGUMENT, "dF, dY, Y, F and G must not be NULL");
  if (c->dtype == FFCM_F32)
    return run_vjp<float>(c, (float*)dF, (float*)dY, (const float*)Y, (const float*)F, (const float*)G);
  return run_vjp<double>(c, (double*)dF, (double*)dY, (const double*)Y, (const double*)F, (const double*)G);
}

int ffcm_mobility_vjp(ffcm_ctx* c, void* dF, void* dY, const void* Y, const void* F, const void* G, int64_t N) {
  if (int rc = check_ready(c)) return rc;
  if (N != c->N) return fail(c, FFCM_ERR_DIMENSION, "dF, dY, Y, F and G must each be 3xN for the config's N");
  if (!dF || !dY || !Y || !F || !G) return fail(c, FFCM_ERR_ARGUMENT, "dF, dY, Y, F and G must not be NULL");
  const size_t bytes = 3 * (size_t)c->N * tsize(c);
  FFCM_HIP(c, hipMemcpyAsync(c->d.Y_in, Y, bytes, hipMemcpyHostToDevice, c->stream));
  FFCM_HIP(c, hipMemcpyAsync(c->d.F_in, F, bytes, hipMemcpyHostToDevice, c->stream));
  FFCM_HIP(c, hipMemcpyAsync(c->d.G_in, G, bytes, hipMemcpyHostToDevice, c->stream));
  int rc = c->dtype == FFCM_F32 ? run_vjp<float>(c, (float*)c->d.V_out, (float*)c->d.dY_out, (const float*)c->d.Y_in,
                                                 (const float*)c->d.F_in, (const float*)c->d.G_in)
                                : run_vjp<double>(c, (double*)c->d.V_out, (double*)c->d.dY_out,
                                                  (const double*)c->d.Y_in, (const double*)c->d.F_in,
                                                  (const double*)c->d.G_in);
  if (rc) return rc;
  FFCM_HIP(c, hipMemcpyAsync(dF, c->d.V_out, bytes, hipMemcpyDeviceToHost, c->stream));
  FFCM_HIP(c, hipMemcpyAsync(dY, c->d.dY_out, bytes, hipMemcpyDeviceToHost, c->stream));
  return sync(c);
}

#define DISPATCH(c, call_f, call_d) ((c)->dtype == FFCM_F32 ? (call_f) : (call_d))

int ffcm_wrap_positions(ffcm_ctx* c) {
  if (int rc = check_ready(c)) return rc;
  c->positions_valid = false;
  FFCM_HIP(c, DISPATCH(c, launch_wrap<float>(c, (const float*)c->d.Y_in),
                       launch_wrap<double>(c, (const double*)c->d.Y_in)));
  return sync(c);
}

int ffcm_assign_cells(ffcm_ctx* c) {
  if (int rc = check_ready(c)) return rc;
  c->positions_valid = false;
  FFCM_HIP(c, DISPATCH(c, launch_assign_cells<float>(c), launch_assign_cells<double>(c)));
  return sync(c);
}

int ffcm_sort_particles_by_cell(ffcm_ctx* c) {
  if (int rc = check_ready(c)) return rc;
  c->positions_valid = false;
  FFCM_HIP(c, launch_sort_cells(c));
  FFCM_HIP(c, DISPATCH(c, launch_gather<float>(c, (const float*)c->d.F_in),
                       launch_gather<double>(c, (const double*)c->d.F_in)));
  return sync(c);
}

int ffcm_spread_forces(ffcm_ctx* c) {
  if (int rc = check_ready(c)) return rc;
  c->tiles_valid = false;
  FFCM_HIP(c, DISPATCH(c, launch_spread<float>(c), launch_spread<double>(c)));
  return sync(c);
}

int ffcm_stokes_solve(ffcm_ctx* c) {
  if (int rc = check_ready(c)) return rc;
  FFCM_HIP(c, DISPATCH(c, launch_fft_stokes<float>(c), launch_fft_stokes<double>(c)));
  return sync(c);
}

int ffcm_stokes_forward_project(ffcm_ctx* c) {
  if (int rc = check_ready(c)) return rc;
  FFCM_HIP(c, DISPATCH(c, launch_fft_forward_project<float>(c), launch_fft_forward_project<double>(c)));
  return sync(c);
}

int ffcm_stokes_backward(ffcm_ctx* c) {
  if (int rc = check_ready(c)) return rc;
  FFCM_HIP(c, DISPATCH(c, launch_fft_inverse<float>(c), launch_fft_inverse<double>(c)));
  return sync(c);
}

int ffcm_interpolate_velocities(ffcm_ctx* c) {
  if (int rc = check_ready(c)) return rc;
  FFCM_HIP(c, DISPATCH(c, launch_interpolate<float>(c, (float*)c->d.V_out),
                       launch_interpolate<double>(c, (double*)c->d.V_out)));
  return sync(c);
}

int ffcm_correct_velocities(ffcm_ctx* c) {
  if (int rc = check_ready(c)) return rc;
  FFCM_HIP(c, DISPATCH(c, launch_correct<float>(c, (float*)c->d.V_out),
                       launch_correct<double>(c, (double*)c->d.V_out)));
  return sync(c);
}

int ffcm_buffer_bytes(const ffcm_ctx* c, int id, size_t* bytes) {
  if (!c || !bytes) return FFCM_ERR_ARGUMENT;
  const size_t ts = tsize(c);
  const size_t N = (size_t)c->N;
  switch (id) {
    case FFCM_BUF_Y_INPUT:
    case FFCM_BUF_F_INPUT:
    case FFCM_BUF_V_OUTPUT:
    case FFCM_BUF_Y_WRAPPED:
    case FFCM_BUF_Y_SORTED:
    case FFCM_BUF_F_SORTED: *bytes = 3 * N * ts; return FFCM_OK;
    case FFCM_BUF_CELL_HASH:
    case FFCM_BUF_ORIGINAL_INDEX: *bytes = N * 4; return FFCM_OK;
    case FFCM_BUF_CELL_START:
    case FFCM_BUF_CELL_END: *bytes = (size_t)c->C * 4; return FFCM_OK;
    case FFCM_BUF_FORCE_DENSITY:
    case FFCM_BUF_FLUID_VELOCITY: *bytes = 3 * (size_t)c->M3 * ts; return FFCM_OK;
    case FFCM_BUF_FLUID_HAT: *bytes = 3 * (size_t)c->Mh * c->host.M[1] * c->host.M[2] * 2 * ts; return FFCM_OK;
    case FFCM_BUF_NEIGHBOR_MAP: *bytes = 13 * (size_t)c->C * 4; return FFCM_OK;
    case FFCM_BUF_COUNTING_SORT_SCRATCH: *bytes = (size_t)c->C * 4; return FFCM_OK;
    default: return FFCM_ERR_ARGUMENT;
  }
}

}  // extern "C"

namespace {

// Host-side reshuffles between the reference layouts and the device layouts.
template <typename T>
void soa_to_interleaved(const T* src, T* dst, size_t npts, int comps_inner) {
  for (size_t i = 0; i < npts; ++i)
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < comps_inner; ++k) dst[(i * 3 + c) * comps_inner + k] = src[(c * npts + i) * comps_inner + k];
}
template <typename T>
void interleaved_to_soa(const T* src, T* dst, size_t npts, int comps_inner) {
  for (size_t i = 0; i < npts; ++i)
    for (int c = 0; c < 3; ++c)
      for (int k = 0; k < comps_inner; ++k) dst[(c * npts + i) * comps_inner + k] = src[(i * 3 + c) * comps_inner + k];
}

int copy_buffer(ffcm_ctx* c, int id, void* host, size_t bytes, bool to_device) {
  size_t want = 0;
  if (ffcm_buffer_bytes(c, id, &want) != FFCM_OK) return fail(c, FFCM_ERR_ARGUMENT, "unknown buffer id");
  if (bytes != want) return fail(c, FFCM_ERR_DIMENSION, "byte count does not match the buffer size");
  if (!host) return fail(c, FFCM_ERR_ARGUMENT, "host pointer must not be NULL");
  FFCM_HIP(c, hipStreamSynchronize(c->stream));
  const size_t ts = tsize(c);
  const size_t N = (size_t)c->N;
  auto raw = [&](void* dev) -> int {
    if (to_device)
      FFCM_HIP(c, hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice));
    else
      FFCM_HIP(c, hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
    return FFCM_OK;
  };
  auto plus_one = [&](int* dev, size_t n, int delta) -> int {
    std::vector<int32_t> tmp(n);
    if (to_device) {
      const int32_t* h = (const int32_t*)host;
      for (size_t i = 0; i < n; ++i) tmp[i] = h[i] - delta;
      FFCM_HIP(c, hipMemcpy(dev, tmp.data(), n * 4, hipMemcpyHostToDevice));
    } else {
      FFCM_HIP(c, hipMemcpy(tmp.data(), dev, n * 4, hipMemcpyDeviceToHost));
      int32_t* h = (int32_t*)host;
      for (size_t i = 0; i < n; ++i) h[i] = tmp[i] + delta;
    }
    return FFCM_OK;
  };
  auto vec4 = [&](void* dev) -> int {
    std::vector<unsigned char> tmp(4 * N * ts, 0);
    if (to_device) {
      for (size_t i = 0; i < N; ++i) std::memcpy(&tmp[4 * i * ts], (const unsigned char*)host + 3 * i * ts, 3 * ts);
      FFCM_HIP(c, hipMemcpy(dev, tmp.data(), tmp.size(), hipMemcpyHostToDevice));
    } else {
      FFCM_HIP(c, hipMemcpy(tmp.data(), dev, tmp.size(), hipMemcpyDeviceToHost));
      for (size_t i = 0; i < N; ++i) std::memcpy((unsigned char*)host + 3 * i * ts, &tmp[4 * i * ts], 3 * ts);
    }
    return FFCM_OK;
  };
  auto grid = [&](void* dev, size_t npts, int inner) -> int {
    std::vector<unsigned char> tmp(bytes);
    if (to_device) {
      if (ts == 4)
        soa_to_interleaved((const float*)host, (float*)tmp.data(), npts, inner);
      else
        soa_to_interleaved((const double*)host, (double*)tmp.data(), npts, inner);
      FFCM_HIP(c, hipMemcpy(dev, tmp.data(), bytes, hipMemcpyHostToDevice));
    } else {
      FFCM_HIP(c, hipMemcpy(tmp.data(), dev, bytes, hipMemcpyDeviceToHost));
      if (ts == 4)
        interleaved_to_soa((const float*)tmp.data(), (float*)host, npts, inner);
      else
        interleaved_to_soa((const double*)tmp.data(), (double*)host, npts, inner);
    }
    return FFCM_OK;
  };
  int rc = FFCM_OK;
  switch (id) {
    case FFCM_BUF_Y_INPUT: rc = raw(c->d.Y_in); break;
    case FFCM_BUF_F_INPUT: rc = raw(c->d.F_in); break;
    case FFCM_BUF_V_OUTPUT: rc = raw(c->d.V_out); break;
    case FFCM_BUF_Y_WRAPPED: rc = raw(c->d.Y_wrapped); break;
    case FFCM_BUF_CELL_HASH: rc = raw(c->d.hash); break;
    case FFCM_BUF_CELL_END: rc = raw(c->d.cell_end); break;
    case FFCM_BUF_COUNTING_SORT_SCRATCH: rc = raw(c->d.sort_scratch); break;
    case FFCM_BUF_Y_SORTED:
      rc = vec4(c->d.Y_sorted);
      if (to_device) c->tiles_valid = false;
      break;
    case FFCM_BUF_F_SORTED: rc = vec4(c->d.F_sorted); break;
    case FFCM_BUF_ORIGINAL_INDEX: rc = plus_one(c->d.perm, N, 1); break;
    case FFCM_BUF_CELL_START: rc = plus_one(c->d.cell_start, (size_t)c->C, 1); break;
    case FFCM_BUF_FORCE_DENSITY: rc = grid(c->d.force, (size_t)c->M3, 1); break;
    case FFCM_BUF_FLUID_VELOCITY: rc = grid(c->d.velocity, (size_t)c->M3, 1); break;
    case FFCM_BUF_FLUID_HAT: {
      // host: 3 SoA fields (Mh, My, Mz) complex; device: rows of Mhp modes x 3 components
      const size_t Mh = (size_t)c->Mh, Mhp = (size_t)c->Mhp, rows = (size_t)c->host.M[1] * c->host.M[2];
      std::vector<unsigned char> tmp(3 * Mhp * rows * 2 * ts, 0);
      if (!to_device) FFCM_HIP(c, hipMemcpy(tmp.data(), c->d.spectrum, tmp.size(), hipMemcpyDeviceToHost));
      unsigned char* h = (unsigned char*)host;
      for (size_t r = 0; r < rows; ++r)
        for (size_t kx = 0; kx < Mh; ++kx)
          for (int comp = 0; comp < 3; ++comp) {
            unsigned char* dv = &tmp[((r * Mhp + kx) * 3 + comp) * 2 * ts];
            unsigned char* hv = h + ((comp * rows + r) * Mh + kx) * 2 * ts;
            if (to_device)
              std::memcpy(dv, hv, 2 * ts);
            else
              std::memcpy(hv, dv, 2 * ts);
          }
      if (to_device) FFCM_HIP(c, hipMemcpy(c->d.spectrum, tmp.data(), tmp.size(), hipMemcpyHostToDevice));
      break;
    }
    case FFCM_BUF_NEIGHBOR_MAP: {
      if (to_device) return fail(c, FFCM_ERR_ARGUMENT, "neighbor_map is read only");
      if (c->d.neighbor_map) {  // the caller's map, copied at ffcm_assemble
        FFCM_HIP(c, hipMemcpy(host, c->d.neighbor_map, bytes, hipMemcpyDeviceToHost));
        break;
      }
      std::vector<int32_t> m = neighbor_map(c->host.num_cells);
      std::memcpy(host, m.data(), bytes);
      break;
    }
    default: return fail(c, FFCM_ERR_ARGUMENT, "unknown buffer id");
  }
  return rc;
}

}  // namespace

extern "C" {

int ffcm_set_buffer(ffcm_ctx* c, int id, const void* host, size_t bytes) {
  if (int rc = check_ctx(c)) return rc;
  c->positions_valid = false;
  return copy_buffer(c, id, const_cast<void*>(host), bytes, true);
}

int ffcm_get_buffer(ffcm_ctx* c, int id, void* host, size_t bytes) {
  if (int rc = check_ctx(c)) return rc;
  return copy_buffer(c, id, host, bytes, false);
}

int ffcm_get_stream(ffcm_ctx* c, void** stream) {
  if (!c || !stream) return FFCM_ERR_ARGUMENT;
  *stream = (void*)c->stream;
  return FFCM_OK;
}

int ffcm_synchronize(ffcm_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  return sync(c);
}

int ffcm_get_info(const ffcm_ctx* c, int32_t* tile, int32_t* n_tiles, int32_t* fft_shape, int32_t* interp_kernel) {
  if (!c) return FFCM_ERR_ARGUMENT;
  if (interp_kernel) *interp_kernel = interp_kernel_kind(c);
  if (tile)
    for (int a = 0; a < 3; ++a) tile[a] = c->tiling.T[a];
  if (n_tiles) *n_tiles = c->tiling.ntiles;
  if (fft_shape) {
    const FftPass* ps[4] = {&c->pass_xf, &c->pass_y, &c->pass_z, &c->pass_xi};
    for (int i = 0; i < 4; ++i) {
      fft_shape[3 * i + 0] = ps[i]->width;
      fft_shape[3 * i + 1] = ps[i]->threads;
      fft_shape[3 * i + 2] = ps[i]->lds_bytes;
    }
  }
  return FFCM_OK;
}

int ffcm_set_timing(ffcm_ctx* c, int enable) {
  if (int rc = check_ctx(c)) return rc;
  Timing* t = timing_of(c);
  if (enable && !t) {
    t = new Timing();
    std::memset(t, 0, sizeof(*t));
    for (int i = 0; i < kTimingSlots; ++i)
      for (int s = 0; s <= kStageCount; ++s) FFCM_HIP(c, hipEventCreate(&t->ev[i][s]));
    c->events[0] = t;
  }
  if (t) {
    FFCM_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < kTimingSlots; ++i) t->pending[i] = 0;
    for (int s = 0; s < kStageCount; ++s) t->ms[s] = 0;
    t->calls = 0;
    t->head = 0;
  }
  c->timing = enable != 0;
  return FFCM_OK;
}

int ffcm_get_timing(ffcm_ctx* c, double* ms_per_stage, int n_stages, int64_t* calls) {
  if (int rc = check_ctx(c)) return rc;
  Timing* t = timing_of(c);
  if (!t) return fail(c, FFCM_ERR_ARGUMENT, "timing was never enabled");
  for (int i = 0; i < kTimingSlots; ++i) timing_collect(c, i);
  for (int s = 0; s < n_stages && s < kStageCount; ++s) ms_per_stage[s] = t->ms[s];
  if (calls) *calls = t->calls;
  return FFCM_OK;
}

}  // extern "C"

namespace ffcm {
int call_begin(ffcm_ctx* c) { return check_ctx(c); }
int call_fail(ffcm_ctx* c, int code, const std::string& msg) { return fail(c, code, msg); }
int call_sync(ffcm_ctx* c) { return sync(c); }
std::vector<int32_t> host_neighbor_map(const int32_t nc[3]) { return neighbor_map(nc); }
}  // namespace ffcm
