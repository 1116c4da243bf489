// Adjoint (vector-Jacobian product) of the mobility apply: for L = <G, M(Y) F>,
//   dL/dF = M(Y) G                      (M is symmetric: src/mobility.jl:157-162)
//   dL/dY = grid part + pair part
// This file holds the two kernels that have no forward counterpart.
//
// Grid part. The stencil weight of particle n at grid point g is
//   W = (a0 + a2 r^2) gx gy gz,  x_i = g_i h - Y_i,  g_i = inv_norm exp(-x_i^2 / (2 Sigma^2))
// (src/stencil.jl:80-112, src/spread_forces.jl:115-123), so
//   dW/dY_i = x_i gx gy gz ((a0 + a2 r^2) / Sigma^2 - 2 a2).
// With u_F and u_G the solved grids of F and G (the same spread and solve),
//   dL/dY_n = h^3 sum_g dW_n/dY (g) [G_n . u_F(g) + F_n . u_G(g)].
// The anchor j = round(Y / h) is piecewise constant, so inside each piece the support is fixed
// and the formula is the exact derivative.
//
// Pair part. With T(x) = A(r) I + B(r) x x^T (even in x) and x = minimage(Y_n - Y_m),
//   dL/dY_n = sum_{m != n, r < R_c} grad_x [G_n^T T F_m + G_m^T T F_n],
//   grad_x(a^T T b) = (A' (a.b) + B' (x.a)(x.b)) x / r + B ((x.b) a + (x.a) b),
// and the pair part of dL/dF_n is delta G_n + sum_m T G_m (the forward correction applied to G).
//
// Design: the grid part is one wave per particle, visited in the spread's tile order like the
// interpolation, reading both solved grids at each stencil point (6 values). The pair part is
// four lanes per slot over the 27 neighbour cells, like the forward correction; A, B, A', B'
// are evaluated once per pair and feed both dL/dF and dL/dY.
#include <cstdlib>

#include "pair_terms.hpp"
#include "stencil.hpp"

namespace ffcm {

namespace {

constexpr int kGradThreads = 256;
constexpr int kXcds = 8;

template <typename T, int MGC>
__global__ void __launch_bounds__(kGradThreads) interp_grad_kernel(
    const typename Vec<T>::v4* __restrict__ Ys, const typename Vec<T>::v4* __restrict__ Fs,
    const typename Vec<T>::v4* __restrict__ Gs, const int* __restrict__ perm, const int* __restrict__ order,
    const T* __restrict__ uF, const T* __restrict__ uG, T* __restrict__ dY, Scalars<T> s) {
  const int MG = MGC > 0 ? MGC : s.MG;
  const int lane = threadIdx.x & 63;
  const int waves_per_block = blockDim.x >> 6;
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  const int per_xcd = gridDim.x / kXcds;  // gridDim.x is a multiple of kXcds
  const int lblock = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  const T two_a2 = T(2) * s.a2;
  for (int i = lblock * waves_per_block + (threadIdx.x >> 6); i < s.N; i += gridDim.x * waves_per_block) {
    const int slot = order[i];
    const typename Vec<T>::v4 y = Ys[slot];
    const typename Vec<T>::v4 f = Fs[slot];
    const typename Vec<T>::v4 gq = Gs[slot];
    const AxisOrigin ax = axis_origin(y.x, s.inv_h, s.half, Mx);
    const AxisOrigin ay = axis_origin(y.y, s.inv_h, s.half, My);
    const AxisOrigin az = axis_origin(y.z, s.inv_h, s.half, Mz);
    AxisEntry<T> ex{T(0), T(0)}, ey{T(0), T(0)}, ez{T(0), T(0)};
    T sx = T(0), sy = T(0), sz = T(0);  // signed offsets x_i = g_i h - Y_i
    if (lane < MG) {
      ex = axis_entry(ax.lo, lane, y.x, s.h, s.inv_norm, s.inv_2S2);
      ey = axis_entry(ay.lo, lane, y.y, s.h, s.inv_norm, s.inv_2S2);
      ez = axis_entry(az.lo, lane, y.z, s.h, s.inv_norm, s.inv_2S2);
      sx = (T)(ax.lo + lane) * s.h - y.x;
      sy = (T)(ay.lo + lane) * s.h - y.y;
      sz = (T)(az.lo + lane) * s.h - y.z;
    }
    T d0 = T(0), d1 = T(0), d2 = T(0);
    if constexpr (MGC > 0 && MGC * MGC <= 64) {
      // as in the interpolation: every z-load of a lane in flight at once, z entries from
      // scalar reads of the owning lane
      const bool valid = lane < MGC * MGC;
      const int ky = valid ? lane / MGC : 0;
      const int kx = valid ? lane - ky * MGC : 0;
      const int ix = wrap_add(ax.low, kx, Mx);
      const int iy = wrap_add(ay.low, ky, My);
      // this lane's own x and y entries (no cross-lane permute)
      const AxisEntry<T> exk = axis_entry(ax.lo, kx, y.x, s.h, s.inv_norm, s.inv_2S2);
      const AxisEntry<T> eyk = axis_entry(ay.lo, ky, y.y, s.h, s.inv_norm, s.inv_2S2);
      const T gx = exk.g, r2x = exk.r2, xx = (T)(ax.lo + kx) * s.h - y.x;
      const T gy = eyk.g, r2y = eyk.r2, xy = (T)(ay.lo + ky) * s.h - y.y;
      const size_t rowxy = ((size_t)iy * Mx + ix) * 3;
      const size_t plane = (size_t)My * Mx * 3;
      T t[MGC];
#pragma unroll
      for (int kz = 0; kz < MGC; ++kz) {
        const size_t o = (size_t)wrap_add(az.low, kz, Mz) * plane + rowxy;
        const T* a = uF + o;
        const T* b = uG + o;
        t[kz] = gq.x * a[0] + gq.y * a[1] + gq.z * a[2] + f.x * b[0] + f.y * b[1] + f.z * b[2];
      }
#pragma unroll
      for (int kz = 0; kz < MGC; ++kz) {
        const T gz = lane_bcast(ez.g, kz), r2z = lane_bcast(ez.r2, kz), xz = lane_bcast(sz, kz);
        const T c = (s.a0 + s.a2 * (r2x + (r2y + r2z))) * s.inv_S2 - two_a2;
        const T tw = c * (gx * (gy * gz)) * t[kz];
        if (valid) {
          d0 += xx * tw;
          d1 += xy * tw;
          d2 += xz * tw;
        }
      }
    } else
    for (int p0 = 0; p0 < MG * MG; p0 += 64) {
      const int p = p0 + lane;
      const bool valid = p < MG * MG;
      const int ky = valid ? p / MG : 0;
      const int kx = valid ? p - ky * MG : 0;
      const int ix = wrap_add(ax.low, kx, Mx);
      const int iy = wrap_add(ay.low, ky, My);
      const T gx = wave_bcast(ex.g, kx), r2x = wave_bcast(ex.r2, kx), xx = wave_bcast(sx, kx);
      const T gy = wave_bcast(ey.g, ky), r2y = wave_bcast(ey.r2, ky), xy = wave_bcast(sy, ky);
      const size_t rowxy = ((size_t)iy * Mx + ix) * 3;
      for (int kz = 0; kz < MG; ++kz) {
        const int iz = wrap_add(az.low, kz, Mz);
        const T gz = wave_bcast(ez.g, kz), r2z = wave_bcast(ez.r2, kz), xz = wave_bcast(sz, kz);
        if (valid) {
          const T c = (s.a0 + s.a2 * (r2x + (r2y + r2z))) * s.inv_S2 - two_a2;
          const T w = c * (gx * (gy * gz));
          const size_t o = (size_t)iz * My * Mx * 3 + rowxy;
          const T* a = uF + o;
          const T* b = uG + o;
          const T t = w * (gq.x * a[0] + gq.y * a[1] + gq.z * a[2] + f.x * b[0] + f.y * b[1] + f.z * b[2]);
          d0 += xx * t;
          d1 += xy * t;
          d2 += xz * t;
        }
      }
    }
    d0 = wave_total(d0);
    d1 = wave_total(d1);
    d2 = wave_total(d2);
    if (lane == 0) {
      const int n = perm[slot];
      dY[3 * n + 0] = s.h3 * d0;
      dY[3 * n + 1] = s.h3 * d1;
      dY[3 * n + 2] = s.h3 * d2;
    }
  }
}

// Four lanes per slot, lane q taking cells q, q + 4, ... (as the forward correction).
constexpr int kVjpLanes = 4;

template <typename T>
__global__ void __launch_bounds__(256) correct_vjp_kernel(const typename Vec<T>::v4* __restrict__ Ys,
                                                          const typename Vec<T>::v4* __restrict__ Fs,
                                                          const typename Vec<T>::v4* __restrict__ Gs,
                                                          const int* __restrict__ cell_start,
                                                          const int* __restrict__ cell_end,
                                                          const int* __restrict__ perm, T* __restrict__ dF,
                                                          T* __restrict__ dY, Scalars<T> s,
                                                          const typename Vec<T>::v4* __restrict__ tab, T tab_scale) {
  __shared__ int list[kPairBatch * 256];
  const int total = s.N * kVjpLanes;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int slot = t / kVjpLanes, q = t % kVjpLanes;
    const typename Vec<T>::v4 yn = Ys[slot];
    const typename Vec<T>::v4 fn = Fs[slot];
    const typename Vec<T>::v4 gn = Gs[slot];
    const T d = q == 0 ? s.delta : T(0);
    T f0 = d * gn.x, f1 = d * gn.y, f2 = d * gn.z;  // dL/dF
    T y0 = T(0), y1 = T(0), y2 = T(0);              // dL/dY
    for_each_pair<T>(slot, yn, s, cell_start, cell_end, Ys, list, [&](int m, T x0, T x1, T x2, T rr) {
      // A, B and A'/r, B'/r: Float32 from the context's table (its second half holds the
      // derivative pieces), else from the formulas
      T A, B, dAr, dBr;
      if constexpr (sizeof(T) == 4) {
        if (tab) {
          pair_scalars_table(rr, tab, tab_scale, &A, &B);
          pair_scalars_table(rr, tab + 2 * kPairTableIntervals, tab_scale, &dAr, &dBr);
        }
      }
      if (sizeof(T) != 4 || !tab) {
        const T r = tsqrt(rr);
        T dA, dB;
        correction_scalars_d(r, s, &A, &B, &dA, &dB);
        dAr = dA / r;
        dBr = dB / r;
      }
      const typename Vec<T>::v4 fm = Fs[m];
      const typename Vec<T>::v4 gm = Gs[m];
      const T xgm = x0 * gm.x + x1 * gm.y + x2 * gm.z;
      const T xgn = x0 * gn.x + x1 * gn.y + x2 * gn.z;
      const T xfm = x0 * fm.x + x1 * fm.y + x2 * fm.z;
      const T xfn = x0 * fn.x + x1 * fn.y + x2 * fn.z;
      const T bg = B * xgm;
      f0 += A * gm.x + bg * x0;
      f1 += A * gm.y + bg * x1;
      f2 += A * gm.z + bg * x2;
      const T ab = (gn.x * fm.x + gn.y * fm.y + gn.z * fm.z) + (gm.x * fn.x + gm.y * fn.y + gm.z * fn.z);
      const T rad = dAr * ab + dBr * (xgn * xfm + xgm * xfn);
      y0 += rad * x0 + B * (xfm * gn.x + xgn * fm.x + xfn * gm.x + xgm * fn.x);
      y1 += rad * x1 + B * (xfm * gn.y + xgn * fm.y + xfn * gm.y + xgm * fn.y);
      y2 += rad * x2 + B * (xfm * gn.z + xgn * fm.z + xfn * gm.z + xgm * fn.z);
    }, q, kVjpLanes);
    f0 = lane_group_sum<kVjpLanes>(f0);
    f1 = lane_group_sum<kVjpLanes>(f1);
    f2 = lane_group_sum<kVjpLanes>(f2);
    y0 = lane_group_sum<kVjpLanes>(y0);
    y1 = lane_group_sum<kVjpLanes>(y1);
    y2 = lane_group_sum<kVjpLanes>(y2);
    if (q == 0) {
      const int n = perm[slot];
      dF[3 * n + 0] += f0;
      dF[3 * n + 1] += f1;
      dF[3 * n + 2] += f2;
      dY[3 * n + 0] += y0;
      dY[3 * n + 1] += y1;
      dY[3 * n + 2] += y2;
    }
  }
}

}  // namespace

template <typename T>
hipError_t launch_interp_grad(ffcm_ctx* c, T* dY, const T* uF, const T* uG) {
  if (!c->tiles_valid) {  // visiting order: the spread's tile lists
    hipError_t e = launch_bin_tiles<T>(c);
    if (e != hipSuccess) return e;
  }
  if constexpr (sizeof(T) == 4) {
    hipError_t e = hipSuccess;
    if (!std::getenv("FFCM_GRAD_WAVE") && launch_interp_grad_ring(c, dY, uF, uG, &e)) return e;
  }
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const int waves = kGradThreads / 64;
  int64_t blocks = (c->N + waves - 1) / waves;
  if (blocks > (1 << 20)) blocks = 1 << 20;
  if (blocks < 1) blocks = 1;
  blocks = (blocks + kXcds - 1) / kXcds * kXcds;
  using V4 = typename Vec<T>::v4;
  if (s.MG == 8)
    interp_grad_kernel<T, 8><<<(unsigned)blocks, kGradThreads, 0, c->stream>>>(
        (const V4*)c->d.Y_sorted, (const V4*)c->d.F_sorted, (const V4*)c->d.G_sorted, c->d.perm, c->d.tile_list, uF,
        uG, dY, s);
  else
    interp_grad_kernel<T, 0><<<(unsigned)blocks, kGradThreads, 0, c->stream>>>(
        (const V4*)c->d.Y_sorted, (const V4*)c->d.F_sorted, (const V4*)c->d.G_sorted, c->d.perm, c->d.tile_list, uF,
        uG, dY, s);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_correct_vjp(ffcm_ctx* c, T* dF, T* dY) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  int64_t blocks = (c->N * kVjpLanes + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > (1 << 20)) blocks = 1 << 20;
  using V4 = typename Vec<T>::v4;
  correct_vjp_kernel<T><<<(unsigned)blocks, 256, 0, c->stream>>>(
      (const V4*)c->d.Y_sorted, (const V4*)c->d.F_sorted, (const V4*)c->d.G_sorted, c->d.cell_start, c->d.cell_end,
      c->d.perm, dF, dY, s, (const V4*)c->d.pair_table, T(kPairTableIntervals) / s.Rc2);
  return hipGetLastError();
}

template hipError_t launch_interp_grad<float>(ffcm_ctx*, float*, const float*, const float*);
template hipError_t launch_interp_grad<double>(ffcm_ctx*, double*, const double*, const double*);
template hipError_t launch_correct_vjp<float>(ffcm_ctx*, float*, float*);
template hipError_t launch_correct_vjp<double>(ffcm_ctx*, double*, double*);

}  // namespace ffcm
