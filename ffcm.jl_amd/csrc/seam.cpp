// The reference's own seams, call for call: _assemble_gpu_buffers with its nine arguments, and
// one entry point per L1 barrier taking the scalars that barrier receives, so that the
// unmodified reference drives this library through a package extension alone.
//
//   _assemble_gpu_buffers(T, N, num_grid_points, M_G, num_cells_total, k_x, k_y, k_z,
//                         neighbor_map_host)                     src/config.jl:495-502
//   wrap_positions!(Y_wrapped, Y_staged, L)                      src/cell_list.jl:22-28
//   _assign_cells_kernel!(cell_hash, Y, inv_cell_size, num_cells)  src/cell_list.jl:193-238
//   _build_cell_list_kernel!(original_index, cell_start, cell_end, counting_sort_scratch,
//                            cell_hash)                          src/cell_list.jl:119-127
//   _gather_particles_kernel!(Y_sorted, F_sorted, Y, F, original_index)  src/cell_list.jl:129-176
//   _spread_forces_kernel!(force_density, Y_sorted, F_sorted, σ, Σ, h, inv_h, num_grid_points,
//                          M_G, stencil...)                      src/spread_forces.jl:17-33
//   mul!(f̂_c, forward plan, f_c), c = x, y, z                    src/stokes_solve.jl:24-26
//   _apply_inverse_stokes_kernel!(f̂x, f̂y, f̂z, k_x, k_y, k_z, μ, inv_M)  src/stokes_solve.jl:28-36
//   mul!(u_c, inverse plan, f̂_c)                                 src/stokes_solve.jl:38-40
//   _interpolate_velocities_kernel!(V, fluid_velocity, Y_sorted, original_index, σ, Σ, h, inv_h,
//                                   num_grid_points, M_G, stencil...)  src/interpolate.jl:23-42
//   _correct_velocities_kernel!(V, Y_sorted, F_sorted, cell_start, cell_end, original_index,
//                               neighbor_map, num_cells, inv_cell_size, L, σ, Σ, a, μ, R_c)
//                                                                src/correct_velocities.jl:154-173
//
// _assemble_gpu_buffers receives no box, kernel or fluid scalars; a context built here learns
// them from the barrier calls (each stores the scalars it is given) or all at once from
// ffcm_set_scalars (what a fused mobility! method, which holds the config, calls). The
// transforms run the three components together: the forward call for component x runs the
// forward transforms of all three, and the calls for y and z that follow it in the reference's
// order find them done (likewise for the inverse).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ffcm_internal.hpp"

using namespace ffcm;

namespace {

constexpr int kFwdState = 16, kInvState = 32;

int set_error(ffcm_ctx* c, int code, const std::string& msg) { return call_fail(c, code, msg); }

// Rebuilds the kernels' scalars after the host parameters changed (a captured apply holds the
// old ones in its kernel arguments and is dropped).
void refresh(ffcm_ctx* c) {
  drop_graph(c);
  if (c->dtype == FFCM_F32)
    fill_scalars<float>(&c->host, (Scalars<float>*)c->scalars);
  else
    fill_scalars<double>(&c->host, (Scalars<double>*)c->scalars);
  (void)upload_pair_table(c);
}

int check_grid(ffcm_ctx* c, const int32_t* M, int32_t M_G) {
  if (!M) return set_error(c, FFCM_ERR_ARGUMENT, "num_grid_points must not be NULL");
  for (int a = 0; a < 3; ++a)
    if (M[a] != c->host.M[a]) return set_error(c, FFCM_ERR_ARGUMENT, "num_grid_points differ from the assembled grid");
  if (M_G != c->host.M_G) return set_error(c, FFCM_ERR_ARGUMENT, "M_G differs from the assembled stencil");
  return FFCM_OK;
}

int check_cells(ffcm_ctx* c, const int32_t* nc) {
  if (!nc) return set_error(c, FFCM_ERR_ARGUMENT, "num_cells must not be NULL");
  for (int a = 0; a < 3; ++a)
    if (nc[a] != c->host.num_cells[a]) return set_error(c, FFCM_ERR_ARGUMENT, "num_cells differ from the neighbor map");
  return FFCM_OK;
}

int launched(ffcm_ctx* c, hipError_t e, const char* what) {
  if (e != hipSuccess) return set_error(c, FFCM_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  return call_sync(c);
}

template <typename F32, typename F64>
hipError_t typed(ffcm_ctx* c, F32 f, F64 d) {
  return c->dtype == FFCM_F32 ? f() : d();
}

}  // namespace

extern "C" {

int ffcm_assemble(ffcm_ctx** out, int32_t dtype, int64_t N, const int32_t* num_grid_points, int32_t M_G,
                  int64_t num_cells_total, const void* k_x, const void* k_y, const void* k_z,
                  const int32_t* neighbor_map_host) {
  if (!out) return set_error(nullptr, FFCM_ERR_ARGUMENT, "out must not be NULL");
  *out = nullptr;
  if (!num_grid_points || !k_x || !k_y || !k_z || !neighbor_map_host)
    return set_error(nullptr, FFCM_ERR_ARGUMENT, "num_grid_points, k_x, k_y, k_z and neighbor_map must be provided");
  if (num_cells_total < 27 || num_cells_total >= (int64_t(1) << 30))
    return set_error(nullptr, FFCM_ERR_ARGUMENT, "num_cells_total must be at least 27 (3 cells per axis)");
  // the cell grid is the one the half-shell map was built for (_build_neighbor_map,
  // src/cell_list.jl:258-289): cell 0's neighbours (0, 1, 0) and (0, 0, 1) are m_x and m_x m_y
  const int32_t* nm = neighbor_map_host;
  const int64_t mx = nm[2], mxy = nm[12];
  if (mx < 3 || mxy < 9 || mxy % mx != 0 || num_cells_total % mxy != 0)
    return set_error(nullptr, FFCM_ERR_ARGUMENT, "neighbor_map is not the half-shell map of a cell grid");
  int32_t nc[3] = {(int32_t)mx, (int32_t)(mxy / mx), (int32_t)(num_cells_total / mxy)};
  if (nc[1] < 3 || nc[2] < 3)
    return set_error(nullptr, FFCM_ERR_ARGUMENT, "neighbor_map is not the half-shell map of a cell grid");
  const std::vector<int32_t> want = host_neighbor_map(nc);
  if (std::memcmp(want.data(), nm, want.size() * sizeof(int32_t)) != 0)
    return set_error(nullptr, FFCM_ERR_ARGUMENT, "neighbor_map is not the half-shell map of a cell grid");
  // provisional box and kernel scalars (h = 1, L = M, unit radius and viscosity) that pass the
  // parameter checks; the barriers or ffcm_set_scalars replace them before any kernel reads
  // them
  ffcm_params p;
  std::memset(&p, 0, sizeof(p));
  p.dtype = dtype;
  p.N = N;
  p.M_G = M_G;
  for (int a = 0; a < 3; ++a) {
    p.M[a] = num_grid_points[a];
    p.num_cells[a] = nc[a];
    p.L[a] = num_grid_points[a];
    p.cell_size[a] = p.L[a] / nc[a];
    p.inv_cell_size[a] = 1.0 / p.cell_size[a];
  }
  p.R_c = 1.0;
  p.a = 1.0;
  p.sigma = 1.0 / std::sqrt(3.141592653589793238462643383279502884);
  p.Sigma = 2.0 * p.sigma;
  p.h = 1.0;
  p.inv_h = 1.0;
  p.mu = 1.0;
  p.k_x = k_x;
  p.k_y = k_y;
  p.k_z = k_z;
  ffcm_ctx* c = nullptr;
  if (int rc = create_context(&c, &p, 1, true)) return rc;
  const size_t bytes = 13 * (size_t)num_cells_total * sizeof(int32_t);
  hipError_t e = hipMalloc((void**)&c->d.neighbor_map, bytes);
  if (e == hipSuccess) e = hipMemcpy(c->d.neighbor_map, nm, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(c->d.sort_scratch, 0, (size_t)num_cells_total * sizeof(int32_t));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    const std::string msg = std::string("neighbor map upload: ") + hipGetErrorString(e);
    ffcm_destroy(c);
    return set_error(nullptr, FFCM_ERR_MEMORY, msg);
  }
  c->scalar_groups = 0;
  *out = c;
  return FFCM_OK;
}

int ffcm_set_scalars(ffcm_ctx* c, const ffcm_params* p) {
  if (int rc = call_begin(c)) return rc;
  std::string msg;
  if (int rc = validate_params(p, &msg)) return set_error(c, rc, msg);
  if (p->dtype != c->dtype || p->N != c->N) return set_error(c, FFCM_ERR_ARGUMENT, "dtype and N must be the assembled ones");
  if (int rc = check_grid(c, p->M, p->M_G)) return rc;
  if (int rc = check_cells(c, p->num_cells)) return rc;
  const void* keep[3] = {c->host.k_x, c->host.k_y, c->host.k_z};
  c->host = *p;
  c->host.k_x = keep[0];
  c->host.k_y = keep[1];
  c->host.k_z = keep[2];
  refresh(c);
  c->scalar_groups = kScalarsAll;
  c->fft_state = 0;
  return FFCM_OK;
}

int ffcm_wrap_positions_kernel(ffcm_ctx* c, int dst, int src, const double* L) {
  if (int rc = call_begin(c)) return rc;
  if (dst != FFCM_BUF_Y_WRAPPED || src != FFCM_BUF_Y_INPUT)
    return set_error(c, FFCM_ERR_ARGUMENT, "wrap_positions! folds Y_input into Y_wrapped");
  if (!L || !(L[0] > 0) || !(L[1] > 0) || !(L[2] > 0)) return set_error(c, FFCM_ERR_ARGUMENT, "L components must be positive");
  for (int a = 0; a < 3; ++a) c->host.L[a] = L[a];
  refresh(c);
  c->scalar_groups |= kScalarsL;
  c->fft_state = 0;
  c->positions_valid = false;
  return launched(c, typed(c, [&] { return launch_wrap<float>(c, (const float*)c->d.Y_in); },
                           [&] { return launch_wrap<double>(c, (const double*)c->d.Y_in); }),
                  "wrap_positions!");
}

int ffcm_assign_cells_kernel(ffcm_ctx* c, int cell_hash, int Y, const double* inv_cell_size, const int32_t* num_cells) {
  if (int rc = call_begin(c)) return rc;
  if (cell_hash != FFCM_BUF_CELL_HASH || Y != FFCM_BUF_Y_WRAPPED)
    return set_error(c, FFCM_ERR_ARGUMENT, "_assign_cells_kernel! hashes Y_wrapped into cell_hash");
  if (int rc = check_cells(c, num_cells)) return rc;
  if (!inv_cell_size) return set_error(c, FFCM_ERR_ARGUMENT, "inv_cell_size must not be NULL");
  for (int a = 0; a < 3; ++a) {
    c->host.inv_cell_size[a] = inv_cell_size[a];
    c->host.cell_size[a] = 1.0 / inv_cell_size[a];
  }
  refresh(c);
  c->scalar_groups |= kScalarsCells;
  c->fft_state = 0;
  c->positions_valid = false;
  return launched(c, typed(c, [&] { return launch_assign_cells<float>(c); }, [&] { return launch_assign_cells<double>(c); }),
                  "_assign_cells_kernel!");
}

int ffcm_build_cell_list_kernel(ffcm_ctx* c) {
  if (int rc = call_begin(c)) return rc;
  c->fft_state = 0;
  c->positions_valid = false;
  return launched(c, launch_sort_cells(c, true), "_build_cell_list_kernel!");
}

int ffcm_gather_particles_kernel(ffcm_ctx* c, int Y, int F) {
  if (int rc = call_begin(c)) return rc;
  if (Y != FFCM_BUF_Y_WRAPPED || F != FFCM_BUF_F_INPUT)
    return set_error(c, FFCM_ERR_ARGUMENT, "_gather_particles_kernel! gathers Y_wrapped and F_input");
  c->fft_state = 0;
  c->positions_valid = false;
  return launched(c, typed(c, [&] { return launch_gather<float>(c, (const float*)c->d.F_in); },
                           [&] { return launch_gather<double>(c, (const double*)c->d.F_in); }),
                  "_gather_particles_kernel!");
}

namespace {
int set_kernel_scalars(ffcm_ctx* c, double sigma, double Sigma, double h, double inv_h, const int32_t* M, int32_t M_G) {
  if (int rc = check_grid(c, M, M_G)) return rc;
  if (!(sigma > 0) || !(Sigma >= sigma) || !(h > 0) || !(inv_h > 0))
    return set_error(c, FFCM_ERR_ARGUMENT, "kernel widths and grid spacing must be positive, with Sigma >= sigma");
  c->host.sigma = sigma;
  c->host.Sigma = Sigma;
  c->host.h = h;
  c->host.inv_h = inv_h;
  refresh(c);
  c->scalar_groups |= kScalarsKernel;
  c->fft_state = 0;
  return FFCM_OK;
}
}  // namespace

int ffcm_spread_forces_kernel(ffcm_ctx* c, double sigma, double Sigma, double h, double inv_h,
                              const int32_t* num_grid_points, int32_t M_G) {
  if (int rc = call_begin(c)) return rc;
  if (int rc = set_kernel_scalars(c, sigma, Sigma, h, inv_h, num_grid_points, M_G)) return rc;
  c->tiles_valid = false;  // the binning uses h
  return launched(c, typed(c, [&] { return launch_spread<float>(c); }, [&] { return launch_spread<double>(c); }),
                  "_spread_forces_kernel!");
}

int ffcm_forward_transform(ffcm_ctx* c, int component) {
  if (int rc = call_begin(c)) return rc;
  if (component < 0 || component > 2) return set_error(c, FFCM_ERR_ARGUMENT, "component must be 0, 1 or 2");
  const bool done = component > 0 && c->fft_state == kFwdState + component - 1;
  c->fft_state = kFwdState + component;
  if (done) return FFCM_OK;
  return launched(c, typed(c, [&] { return launch_fft_steps<float>(c, kFftForward); },
                           [&] { return launch_fft_steps<double>(c, kFftForward); }),
                  "mul!(fluid_hat, forward_fourier_transform, force_density)");
}

int ffcm_apply_inverse_stokes_kernel(ffcm_ctx* c, double mu, double inv_M) {
  if (int rc = call_begin(c)) return rc;
  if (!(mu > 0)) return set_error(c, FFCM_ERR_ARGUMENT, "viscosity must be positive");
  const double Mtot = (double)c->host.M[0] * c->host.M[1] * c->host.M[2];
  const bool exact = c->dtype == FFCM_F32 ? (float)inv_M == 1.0f / (float)Mtot : inv_M == 1.0 / Mtot;
  if (!exact) return set_error(c, FFCM_ERR_ARGUMENT, "inv_M must be 1 / prod(num_grid_points) in T");
  c->host.mu = mu;
  refresh(c);
  c->scalar_groups |= kScalarsMu;
  c->fft_state = 0;
  return launched(c, typed(c, [&] { return launch_fft_steps<float>(c, kFftProject); },
                           [&] { return launch_fft_steps<double>(c, kFftProject); }),
                  "_apply_inverse_stokes_kernel!");
}

int ffcm_inverse_transform(ffcm_ctx* c, int component) {
  if (int rc = call_begin(c)) return rc;
  if (component < 0 || component > 2) return set_error(c, FFCM_ERR_ARGUMENT, "component must be 0, 1 or 2");
  const bool done = component > 0 && c->fft_state == kInvState + component - 1;
  c->fft_state = kInvState + component;
  if (done) return FFCM_OK;
  return launched(c, typed(c, [&] { return launch_fft_steps<float>(c, kFftInverse); },
                           [&] { return launch_fft_steps<double>(c, kFftInverse); }),
                  "mul!(fluid_velocity, inverse_fourier_transform, fluid_hat)");
}

int ffcm_interpolate_velocities_kernel(ffcm_ctx* c, int V, double sigma, double Sigma, double h, double inv_h,
                                       const int32_t* num_grid_points, int32_t M_G) {
  if (int rc = call_begin(c)) return rc;
  if (V != FFCM_BUF_V_OUTPUT) return set_error(c, FFCM_ERR_ARGUMENT, "_interpolate_velocities_kernel! writes V_output");
  if (int rc = set_kernel_scalars(c, sigma, Sigma, h, inv_h, num_grid_points, M_G)) return rc;
  return launched(c, typed(c, [&] { return launch_interpolate<float>(c, (float*)c->d.V_out); },
                           [&] { return launch_interpolate<double>(c, (double*)c->d.V_out); }),
                  "_interpolate_velocities_kernel!");
}

int ffcm_correct_velocities_kernel(ffcm_ctx* c, int V, const int32_t* num_cells, const double* inv_cell_size,
                                   const double* L, double sigma, double Sigma, double a, double mu, double R_c) {
  if (int rc = call_begin(c)) return rc;
  if (V != FFCM_BUF_V_OUTPUT) return set_error(c, FFCM_ERR_ARGUMENT, "_correct_velocities_kernel! adds to V_output");
  if (int rc = check_cells(c, num_cells)) return rc;
  if (!inv_cell_size || !L) return set_error(c, FFCM_ERR_ARGUMENT, "inv_cell_size and L must not be NULL");
  if (!(R_c > 0) || !(mu > 0) || a != 1.0 || !(sigma > 0) || !(Sigma >= sigma))
    return set_error(c, FFCM_ERR_ARGUMENT, "invalid correction scalars (R_c, mu, sigma, Sigma; only a = 1 is supported)");
  for (int i = 0; i < 3; ++i) {
    c->host.L[i] = L[i];
    c->host.inv_cell_size[i] = inv_cell_size[i];
    c->host.cell_size[i] = 1.0 / inv_cell_size[i];
  }
  c->host.sigma = sigma;
  c->host.Sigma = Sigma;
  c->host.a = a;
  c->host.mu = mu;
  c->host.R_c = R_c;
  refresh(c);
  c->scalar_groups |= kScalarsL | kScalarsCells | kScalarsMu | kScalarsPair;
  c->fft_state = 0;
  return launched(c, typed(c, [&] { return launch_correct<float>(c, (float*)c->d.V_out); },
                           [&] { return launch_correct<double>(c, (double*)c->d.V_out); }),
                  "_correct_velocities_kernel!");
}

}  // extern "C"
