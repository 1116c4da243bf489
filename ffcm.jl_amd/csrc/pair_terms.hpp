// Pair-correction terms shared by the correction kernel and its adjoint.
//
// Reference: _correction_scalars (src/correct_velocities.jl:71-104), _min_image (:121-127) and
// the cell lookup of _assign_cells_kernel! (src/cell_list.jl:229-234). The pair tensor is
// T(x) = A(r) I + B(r) x x^T with
//   A = E (1/(8 pi mu r) + s2/(4 pi mu r^3)) - 2 s4/(mu r^2) Ds
//       + (2 S4/(mu r^2) + d/mu (1 + 2 S2/r^2) - dq (2 - r^2/(2 S2))/(2 mu S2)) DS
//   B = E (1/(8 pi mu r^3) - 3 s2/(4 pi mu r^5)) + 6 s4/(mu r^4) Ds
//       + (-6 S4/(mu r^4) - d/(mu r^2) (1 + 6 S2/r^2) - dq/(4 mu S4)) DS
// where E = erf(r/2 sigma) - erf(r/2 Sigma), Ds = (4 pi s2)^(-3/2) exp(-r^2/(4 s2)) (likewise
// DS with S2), d = s2 - S2 and dq = d^2/4.
#pragma once

#include "ffcm_internal.hpp"

namespace ffcm {

// The coefficient quotients of A(r), B(r) that do not depend on r, evaluated once per thread.
template <typename T>
struct PairConsts {
  T i2sig, i2Sig;   // 1 / (2 sigma), 1 / (2 Sigma): erf arguments
  T i4s2, i4S2;     // 1 / (4 s2), 1 / (4 S2): exponents
  T Ds_pref, DS_pref;
  T a1, a2;         // A: 1 / (8 pi mu), s2 / (4 pi mu)
  T a3;             // A: 2 s4 / mu
  T a4, a5, a6, a7; // A: 2 S4 / mu, dd / mu, 2 S2, dq / (2 mu S2)
  T i2S2;           // 1 / (2 S2)
  T b2;             // B: 3 s2 / (4 pi mu)
  T b3, b4, b6, b7; // B: 6 s4 / mu, 6 S4 / mu, 6 S2, dq / (4 mu S4)
};

template <typename T>
__device__ __forceinline__ PairConsts<T> pair_consts(const Scalars<T>& s) {
  PairConsts<T> k;
  k.i2sig = T(1) / s.two_sigma;
  k.i2Sig = T(1) / s.two_Sigma;
  k.i4s2 = T(1) / s.four_s2;
  k.i4S2 = T(1) / s.four_S2;
  k.Ds_pref = s.Ds_pref;
  k.DS_pref = s.DS_pref;
  k.a1 = T(1) / s.eightpimu;
  k.a2 = s.s2 / s.fourpimu;
  k.a3 = T(2) * s.s4 / s.mu;
  k.a4 = T(2) * s.S4 / s.mu;
  k.a5 = s.dd / s.mu;
  k.a6 = T(2) * s.S2;
  k.a7 = s.dq / s.two_mu_S2;
  k.i2S2 = T(1) / s.two_S2;
  k.b2 = T(3) * s.s2 / s.fourpimu;
  k.b3 = T(6) * s.s4 / s.mu;
  k.b4 = T(6) * s.S4 / s.mu;
  k.b6 = T(6) * s.S2;
  k.b7 = s.dq / s.four_mu_S4;
  return k;
}

// A(r), B(r): the reference's terms with one division (1/r, whose powers replace the divisions
// by r^n) and the r-independent quotients from PairConsts. The grouping of the reference is kept;
// the rounding differs from its operation order in the last bits, for about half the arithmetic
// per pair.
template <typename T>
__device__ __forceinline__ void correction_scalars(T r, const PairConsts<T>& k, T* A, T* B) {
  const T ir = T(1) / r;
  const T ir2 = ir * ir;
  const T ir3 = ir2 * ir;
  const T ir4 = ir2 * ir2;
  const T ir5 = ir4 * ir;
  const T r2 = r * r;
  const T erf_diff = erf(r * k.i2sig) - erf(r * k.i2Sig);
  const T Ds = k.Ds_pref * exp(-r2 * k.i4s2);
  const T DS = k.DS_pref * exp(-r2 * k.i4S2);
  *A = erf_diff * (k.a1 * ir + k.a2 * ir3) - k.a3 * ir2 * Ds +
       (k.a4 * ir2 + k.a5 * (T(1) + k.a6 * ir2) - k.a7 * (T(2) - r2 * k.i2S2)) * DS;
  *B = erf_diff * (k.a1 * ir3 - k.b2 * ir5) + k.b3 * ir4 * Ds +
       (-k.b4 * ir4 - k.a5 * ir2 * (T(1) + k.b6 * ir2) - k.b7) * DS;
}

// A(r), B(r) from the context's table of cubic pieces in s = r^2 (kPairTableIntervals over
// [0, R_c^2], interpolating the double-precision scalars at Chebyshev points): two 16-byte loads
// and two Horner evaluations in place of the erf/exp evaluation. Against the double-precision
// values the pieces are within about 1e-7 of A and B relative to their largest magnitude, where
// the Float32 evaluation of the formulas loses up to a few 1e-6 to cancellation between terms.
template <typename V4>
__device__ __forceinline__ void pair_scalars_table(float rr, const V4* __restrict__ tab, float scale, float* A,
                                                   float* B) {
  const float x = rr * scale;
  const int i = min((int)x, kPairTableIntervals - 1);
  const float t = x - (float)i;
  const V4 a = tab[2 * i], b = tab[2 * i + 1];
  *A = ((a.w * t + a.z) * t + a.y) * t + a.x;
  *B = ((b.w * t + b.z) * t + b.y) * t + b.x;
}

// A, B and their radial derivatives A'(r), B'(r). Each of A and B is a sum of three
// products (coefficient x radial function); the derivative is the product rule term by term
// with E' = (exp(-r^2/4 s2)/sigma - exp(-r^2/4 S2)/Sigma)/sqrt(pi), Ds' = -r Ds/(2 s2) and
// DS' = -r DS/(2 S2).
template <typename T>
__device__ __forceinline__ void correction_scalars_d(T r, const Scalars<T>& s, T* A, T* B, T* dA, T* dB) {
  const T ir = T(1) / r;
  const T ir2 = ir * ir;
  const T ir3 = ir2 * ir;
  const T ir4 = ir2 * ir2;
  const T ir5 = ir4 * ir;
  const T ir6 = ir3 * ir3;
  const T r2 = r * r;
  const T imu = T(1) / s.mu;
  const T es = exp(-r2 / s.four_s2);
  const T eS = exp(-r2 / s.four_S2);
  const T E = erf(r / s.two_sigma) - erf(r / s.two_Sigma);
  const T dE = s.inv_sqrtpi * (es / s.sigma - eS / s.Sigma);
  const T Ds = s.Ds_pref * es;
  const T DS = s.DS_pref * eS;
  const T dDs = -r * Ds / s.two_s2;
  const T dDS = -r * DS / s.two_S2;
  const T i8 = T(1) / s.eightpimu, i4 = T(1) / s.fourpimu;
  // A
  const T cA1 = i8 * ir + s.s2 * i4 * ir3;
  const T dcA1 = -i8 * ir2 - T(3) * s.s2 * i4 * ir4;
  const T cA2 = -T(2) * s.s4 * imu * ir2;
  const T dcA2 = T(4) * s.s4 * imu * ir3;
  const T cA3 = T(2) * s.S4 * imu * ir2 + s.dd * imu * (T(1) + T(2) * s.S2 * ir2) -
                s.dq * (T(2) - r2 / s.two_S2) / s.two_mu_S2;
  const T dcA3 = -T(4) * s.S4 * imu * ir3 - T(4) * s.dd * s.S2 * imu * ir3 + s.dq * r / (s.S2 * s.two_mu_S2);
  *A = E * cA1 + cA2 * Ds + cA3 * DS;
  *dA = dE * cA1 + E * dcA1 + dcA2 * Ds + cA2 * dDs + dcA3 * DS + cA3 * dDS;
  // B
  const T cB1 = i8 * ir3 - T(3) * s.s2 * i4 * ir5;
  const T dcB1 = -T(3) * i8 * ir4 + T(15) * s.s2 * i4 * ir6;
  const T cB2 = T(6) * s.s4 * imu * ir4;
  const T dcB2 = -T(24) * s.s4 * imu * ir5;
  const T cB3 = -T(6) * s.S4 * imu * ir4 - s.dd * imu * ir2 * (T(1) + T(6) * s.S2 * ir2) - s.dq / s.four_mu_S4;
  const T dcB3 = T(24) * s.S4 * imu * ir5 + s.dd * imu * (T(2) * ir3 + T(24) * s.S2 * ir5);
  *B = E * cB1 + cB2 * Ds + cB3 * DS;
  *dB = dE * cB1 + E * dcB1 + dcB2 * Ds + cB2 * dDs + dcB3 * DS + cB3 * dDS;
}

// floor(Int32, y * inv_cell_size) clamped to [0, m - 1] (src/cell_list.jl:229-234).
template <typename T>
__device__ __forceinline__ int cell_axis(T y, T inv_cs, int m) {
  int c = (int)floor(y * inv_cs);
  return c > m - 1 ? m - 1 : (c < 0 ? 0 : c);
}

// x - L round(x / L), ties to even (src/correct_velocities.jl:121-127). Neighbour-cell
// candidates are mostly close (|x| <= L/4), where x / L rounds to at most 1/4 in magnitude,
// round() gives 0 and the expression returns x: the division is only evaluated for pairs
// across the periodic boundary.
template <typename T>
__device__ __forceinline__ T min_image(T x, T L) {
  if (fabs(x) <= T(0.25) * L) return x;
  return x - L * rint(x / L);
}

__device__ __forceinline__ int wrap_cell(int c, int m) { return c < 0 ? c + m : (c >= m ? c - m : c); }

// sqrt in the working precision (an unqualified call inside a lambda can resolve to the
// double overload).
template <typename T>
__device__ __forceinline__ T tsqrt(T x) {
  if constexpr (sizeof(T) == 4)
    return sqrtf(x);
  else
    return ::sqrt(x);
}

// In-range pairs found per scan batch and thread (LDS list of kPairBatch x block threads).
constexpr int kPairBatch = 16;
// Candidates of one cell whose positions are loaded together during the scan.
constexpr int kScanGroup = 2;

// Calls fn(m, x0, x1, x2, rr) for every slot m != slot with rr = |minimage(Y_slot - Y_m)|^2 <
// R_c^2, in the reference's order: the 27 periodic neighbour cells in (dz, dy, dx) order
// (src/correct_velocities.jl:240-270), slots in range order. With k0, kstep the walk covers
// only cells k0, k0 + kstep, ... of that order (the share of one of kstep lanes serving the
// same slot). Two phases per batch: a cheap scan of the candidates records the in-range ones
// in the thread's LDS list (`list`, stride blockDim.x), then fn runs over the list. A wave thus
// runs the expensive pair work as many times as its lane with the most in-range pairs, not its
// lane with the most candidates. The scan is latency-bound on its dependent loads (cell bounds,
// then positions), so the positions of kScanGroup candidates of a cell are loaded together
// (at 2^20 particles on 512^3: 0.447 ms one at a time, 0.361 ms two, 0.370 ms four, 0.468 ms
// eight), and the bounds of the lane's next cell are loaded when a cell is opened.
template <typename T, typename Fn>
__device__ __forceinline__ void for_each_pair(int slot, const typename Vec<T>::v4& yn, const Scalars<T>& s,
                                              const int* __restrict__ cell_start, const int* __restrict__ cell_end,
                                              const typename Vec<T>::v4* __restrict__ Ys, int* list, Fn&& fn,
                                              int k0 = 0, int kstep = 1) {
  using V4 = typename Vec<T>::v4;
  const int mx = s.nc[0], my = s.nc[1], mz = s.nc[2];
  const int cx = cell_axis(yn.x, s.inv_cs[0], mx);
  const int cy = cell_axis(yn.y, s.inv_cs[1], my);
  const int cz = cell_axis(yn.z, s.inv_cs[2], mz);
  const int stride = (int)blockDim.x;
  int* const mine = list + threadIdx.x;
  int* const last = mine + (kPairBatch - kScanGroup) * stride;  // a batch takes no group past it
  // position m by a 32-bit byte offset (N <= 2^26), which the loads take as a register offset
  // from the uniform base
  auto pos = [&](int m) {
    return *(const V4*)((const unsigned char*)Ys + (unsigned)m * (unsigned)sizeof(V4));
  };
  auto bounds = [&](int k, int& b, int& e) {
    const int nz = wrap_cell(cz + k / 9 - 1, mz);
    const int ny = wrap_cell(cy + (k / 3) % 3 - 1, my);
    const int nx = wrap_cell(cx + k % 3 - 1, mx);
    const int cn = nx + (ny + nz * my) * mx;
    b = cell_start[cn];
    e = cell_end[cn];
  };
  int cell = k0, m = 0, e = 0, nm = 0, ne = 0;
  bounds(k0, m, e);  // k0 < 27
  if (k0 + kstep < 27) bounds(k0 + kstep, nm, ne);
  bool done = false;
  for (;;) {
    int* w = mine;  // next free list entry
    while (!done && w <= last) {
      if (m < e) {
        V4 yc[kScanGroup];
#pragma unroll
        for (int k = 0; k < kScanGroup; ++k) yc[k] = pos(min(m + k, e - 1));
#pragma unroll
        for (int k = 0; k < kScanGroup; ++k) {
          const int mk = m + k;
          if (mk < e && mk != slot) {
            const T x0 = min_image(yn.x - yc[k].x, s.L[0]);
            const T x1 = min_image(yn.y - yc[k].y, s.L[1]);
            const T x2 = min_image(yn.z - yc[k].z, s.L[2]);
            if (x0 * x0 + x1 * x1 + x2 * x2 < s.Rc2) {
              *w = mk;
              w += stride;
            }
          }
        }
        m += kScanGroup;
      } else if ((cell += kstep) >= 27) {
        done = true;
      } else {
        m = nm;
        e = ne;
        if (cell + kstep < 27) bounds(cell + kstep, nm, ne);
      }
    }
    for (const int* q = mine; q < w; q += stride) {
      const int mm = *q;
      const V4 ym = pos(mm);
      const T x0 = min_image(yn.x - ym.x, s.L[0]);
      const T x1 = min_image(yn.y - ym.y, s.L[1]);
      const T x2 = min_image(yn.z - ym.z, s.L[2]);
      fn(mm, x0, x1, x2, x0 * x0 + x1 * x1 + x2 * x2);
    }
    if (done) break;
  }
}

}  // namespace ffcm
