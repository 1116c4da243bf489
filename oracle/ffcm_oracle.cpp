// CPU restatement of ForceCouplingMethod.jl's CPU (`gpu_acceleration = false`) mobility path.
//
// TEST INFRASTRUCTURE ONLY. This file is the parity oracle for the HIP backend: only the test
// suite, the smoke check and the benchmark's CPU-baseline leg load it, and only as the checker
// (never as the thing measured or shipped).
//
// Every function follows the reference loop for loop and operation for operation, in the
// working precision T (float or double), so that a float build reproduces the reference's
// Float32 CPU path up to libm ulps and FFT summation order. Citations are relative to the
// ForceCouplingMethod.jl v0.1.0 source tree.
//
// Parity pins (what checks each function; see tests/test_oracle_*.py):
//   wrap / hash / cell list / gather: the reference's known-answer tests
//     (test/accuracy/test_wrap_positions.jl, test_assign_cells_kernel.jl,
//     test_sort_particles_by_cell.jl) restated as Python fixtures, plus brute force.
//   stencil + spread: the closed-form modified kernel (test_spread_forces.jl:7-52) and the
//     anchor known answers (test_spread_forces.jl:54-93).
//   FFT: numpy.fft.rfftn / irfftn in float64 on random grids, including odd sizes.
//   Stokes: the analytic single mode (test_stokes_solve.jl:208-247), divergence-free modes.
//   interpolation: the adjoint identity J = h^3 S^T (test_interpolate_velocities.jl:8-63).
//   correction: the independent full-tensor oracle (test_correct_velocities.jl:21-68).
//   end to end: the reciprocal-lattice self-mobility (test_single_sphere_mobility.jl).
//   adjoint (mobility_vjp, correction_scalars_deriv; no reference counterpart): float64
//     central differences of mobility() and of correction_scalars(), and dF = M G through
//     the symmetry <G, M F> = <F, M G>.
//   No function is left "parity unpinned"; erf/exp last-ulp agreement with Julia's libm is
//   not pinned (values are, by float64 mathematics).
//
// Threads (oracle_set_threads): the transforms can run over several threads, one line per
// thread at a time, as the reference's fft_threads option runs FFTW's (src/config.jl:173-176);
// the spread, interpolation and correction can too, for the large error-oracle runs. Either
// way every output is computed with the same operations in the same order as the serial
// loops (each transform line on its own; spread contributions to a grid point in particle
// order, each thread owning a range of z planes; one particle per thread in interpolation and
// correction), so the results are bit-identical for any thread count.
#include <omp.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>

extern "C" {
// Mirrors the scalars of FFCMConfig (src/config.jl:196-215). Every scalar was computed by the
// host in T and is passed here as a double, so the conversion back to T is exact.
typedef struct oracle_params {
  int32_t dtype;  // 0 = float, 1 = double
  int32_t pad0;
  int64_t N;
  int32_t M[3];
  int32_t M_G;
  int32_t num_cells[3];
  int32_t pad1;
  double L[3];
  double R_c, a, sigma, Sigma, h, inv_h, mu;
  double cell_size[3];
  double inv_cell_size[3];
  const void *k_x, *k_y, *k_z;  // host wavevectors in T: lengths M_x/2+1, M_y, M_z
} oracle_params;
}

namespace {

int g_fft_threads = 1;       // transforms
int g_particle_threads = 1;  // spread, interpolation, correction

template <typename T>
struct P {
  int64_t N;
  int32_t M[3], MG, nc[3];
  T L[3], Rc, a, sigma, Sigma, h, inv_h, mu, inv_cs[3];
  const T *kx, *ky, *kz;
  explicit P(const oracle_params* p) {
    N = p->N;
    for (int i = 0; i < 3; ++i) {
      M[i] = p->M[i];
      nc[i] = p->num_cells[i];
      L[i] = (T)p->L[i];
      inv_cs[i] = (T)p->inv_cell_size[i];
    }
    MG = p->M_G;
    Rc = (T)p->R_c;
    a = (T)p->a;
    sigma = (T)p->sigma;
    Sigma = (T)p->Sigma;
    h = (T)p->h;
    inv_h = (T)p->inv_h;
    mu = (T)p->mu;
    kx = (const T*)p->k_x;
    ky = (const T*)p->k_y;
    kz = (const T*)p->k_z;
  }
};

template <typename T>
inline T pi_T() { return (T)3.141592653589793238462643383279502884; }

// Julia `mod(x::T, y::T)` for floats: r = rem(x, y) (exact fmod); r == 0 -> copysign(0, y);
// r and y of different sign -> r + y; else r.
template <typename T>
inline T jl_mod(T x, T y) {
  T r = std::fmod(x, y);
  if (r == 0) return std::copysign(T(0), y);
  if ((r > 0) != (y > 0)) return r + y;
  return r;
}

inline int32_t jl_mod_int(int32_t x, int32_t m) {
  int32_t r = x % m;
  return r < 0 ? r + m : r;
}

// Julia `round(Int32, x)`: round to nearest, ties to even (default FE_TONEAREST mode).
template <typename T>
inline int32_t jl_round_i32(T x) { return (int32_t)std::nearbyint(x); }

template <typename T>
inline int32_t jl_floor_i32(T x) { return (int32_t)std::floor(x); }

// ---------------------------------------------------------------- step 1: wrap + hash
// wrap_positions! (src/cell_list.jl:22-28): destination .= mod.(source, L).
template <typename T>
void wrap(const P<T>& p, const T* Y, T* Yw) {
  for (int64_t n = 0; n < p.N; ++n)
    for (int i = 0; i < 3; ++i) Yw[3 * n + i] = jl_mod(Y[3 * n + i], p.L[i]);
}

// _assign_cells_kernel! (src/cell_list.jl:222-238).
template <typename T>
void assign_cells(const P<T>& p, const T* Yw, int32_t* hash) {
  for (int64_t n = 0; n < p.N; ++n) {
    int32_t c[3];
    for (int i = 0; i < 3; ++i) {
      c[i] = jl_floor_i32(Yw[3 * n + i] * p.inv_cs[i]);
      if (c[i] > p.nc[i] - 1) c[i] = p.nc[i] - 1;
    }
    hash[n] = c[0] + (c[1] + c[2] * p.nc[1]) * p.nc[0];
  }
}

// ---------------------------------------------------------------- step 2: counting sort
// _build_cell_list_kernel! (src/cell_list.jl:64-93). Outputs are 1-based like the reference.
void build_cell_list(int64_t N, int64_t total, const int32_t* hash, int32_t* orig,
                     int32_t* cell_start, int32_t* cell_end) {
  std::vector<int32_t> scratch(total, 0);
  for (int64_t n = 0; n < N; ++n) scratch[hash[n]] += 1;
  int32_t acc = 1;
  for (int64_t c = 0; c < total; ++c) {
    cell_start[c] = acc;
    acc += scratch[c];
    cell_end[c] = acc - 1;
    scratch[c] = cell_start[c];
  }
  for (int64_t n = 0; n < N; ++n) {
    int32_t c = hash[n];
    orig[scratch[c] - 1] = (int32_t)(n + 1);
    scratch[c] += 1;
  }
}

// _gather_particles_kernel! (src/cell_list.jl:159-176); orig is 1-based.
template <typename T>
void gather(int64_t N, const int32_t* orig, const T* Y, const T* F, T* Ys, T* Fs) {
  for (int64_t s = 0; s < N; ++s) {
    int64_t n = orig[s] - 1;
    for (int i = 0; i < 3; ++i) {
      Ys[3 * s + i] = Y[3 * n + i];
      Fs[3 * s + i] = F[3 * n + i];
    }
  }
}

// ---------------------------------------------------------------- stencil
// _modified_kernel_coefficients (src/stencil.jl:22-31).
template <typename T>
void kernel_coefficients(T sigma, T Sigma, T* a0, T* a2, T* inv_norm, T* inv_2S2) {
  T S2 = Sigma * Sigma;
  T S4 = S2 * S2;
  T d = sigma * sigma - S2;
  *a0 = T(1) - T(3) * d / (T(2) * S2);
  *a2 = d / (T(2) * S4);
  *inv_norm = T(1) / std::sqrt(T(2) * pi_T<T>() * S2);
  *inv_2S2 = T(1) / (T(2) * S2);
}

// _fill_particle_stencil! (src/stencil.jl:80-112). idx is 0-based here (reference: +1).
template <typename T>
void fill_stencil(const P<T>& p, const T* y, T inv_norm, T inv_2S2, T* gau, T* r2, int32_t* idx) {
  const int32_t half = p.MG / 2;
  for (int i = 0; i < 3; ++i) {
    int32_t j = jl_round_i32(y[i] * p.inv_h);
    for (int32_t k = 0; k < p.MG; ++k) {
      int32_t g = j + (k - half);
      T x = (T)g * p.h - y[i];
      T x2 = x * x;
      gau[i * p.MG + k] = inv_norm * std::exp(-x2 * inv_2S2);
      r2[i * p.MG + k] = x2;
      idx[i * p.MG + k] = jl_mod_int(g, p.M[i]);
    }
  }
}

// ---------------------------------------------------------------- step 3: spread
// _spread_forces_kernel! (src/spread_forces.jl:73-130). grid = 3 SoA fields (x fastest).
template <typename T>
void spread(const P<T>& p, const T* Ys, const T* Fs, T* grid) {
  const int64_t M3 = (int64_t)p.M[0] * p.M[1] * p.M[2];
  std::memset(grid, 0, sizeof(T) * 3 * M3);
  T a0, a2, inv_norm, inv_2S2;
  kernel_coefficients(p.sigma, p.Sigma, &a0, &a2, &inv_norm, &inv_2S2);
  const int MG = p.MG;
  T* fx = grid;
  T* fy = grid + M3;
  T* fz = grid + 2 * M3;
  const int nth = std::max(1, std::min(g_particle_threads, (int)p.M[2]));
  // thread t adds the contributions to grid planes [z0, z1), particle after particle: the
  // order of the additions at every grid point is the serial loop's
#pragma omp parallel num_threads(nth)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const int z0 = (int)((int64_t)p.M[2] * t / nt), z1 = (int)((int64_t)p.M[2] * (t + 1) / nt);
    std::vector<T> gau(3 * MG), r2(3 * MG);
    std::vector<int32_t> idx(3 * MG);
    for (int64_t s = 0; s < p.N; ++s) {
      if (nt > 1) {  // skip particles whose stencil misses the planes (anchor test, no exp)
        const int32_t j = jl_round_i32(Ys[3 * s + 2] * p.inv_h) - MG / 2;
        bool hit = false;
        for (int k = 0; k < MG && !hit; ++k) {
          const int32_t iz = jl_mod_int(j + k, p.M[2]);
          hit = iz >= z0 && iz < z1;
        }
        if (!hit) continue;
      }
      const T F0 = Fs[3 * s], F1 = Fs[3 * s + 1], F2 = Fs[3 * s + 2];
      fill_stencil(p, Ys + 3 * s, inv_norm, inv_2S2, gau.data(), r2.data(), idx.data());
      for (int kz = 0; kz < MG; ++kz) {
        int32_t iz = idx[2 * MG + kz];
        if (iz < z0 || iz >= z1) continue;
        T gz = gau[2 * MG + kz];
        T r2z = r2[2 * MG + kz];
        for (int ky = 0; ky < MG; ++ky) {
          int32_t iy = idx[MG + ky];
          T gyz = gau[MG + ky] * gz;
          T r2yz = r2[MG + ky] + r2z;
          for (int kx = 0; kx < MG; ++kx) {
            int32_t ix = idx[kx];
            T w = (a0 + a2 * (r2[kx] + r2yz)) * gau[kx] * gyz;
            int64_t g = ix + (int64_t)p.M[0] * (iy + (int64_t)p.M[1] * iz);
            fx[g] += w * F0;
            fy[g] += w * F1;
            fz[g] += w * F2;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------- FFT (FFTW conventions)
// Mixed-radix recursive Cooley-Tukey on std::complex<T>, twiddles from a table of
// exp(sign*2*pi*i*m/n) computed in double and rounded to T. Forward sign -1, unnormalised,
// matching FFTW's plan_rfft / plan_brfft conventions used at src/config.jl:418-423.
template <typename T>
struct Fft1 {
  int n = 0;
  std::vector<int> factors;
  std::vector<std::complex<T>> tw_fwd, tw_bwd;  // exp(-/+ 2 pi i m / n)
  std::vector<std::complex<T>> scratch;
  explicit Fft1(int n_) : n(n_) {
    int r = n;
    for (int f : {4, 2, 3, 5}) {
      while (r % f == 0) {
        factors.push_back(f);
        r /= f;
      }
    }
    for (int f = 7; r > 1; f += 2) {
      while (r % f == 0) {
        factors.push_back(f);
        r /= f;
      }
    }
    factors.push_back(1);
    tw_fwd.resize(n);
    tw_bwd.resize(n);
    for (int m = 0; m < n; ++m) {
      double ang = 2.0 * 3.141592653589793238462643383279502884 * (double)m / (double)n;
      tw_fwd[m] = std::complex<T>((T)std::cos(ang), (T)-std::sin(ang));
      tw_bwd[m] = std::complex<T>((T)std::cos(ang), (T)std::sin(ang));
    }
    scratch.resize(n);
  }

  void rec(const std::complex<T>* in, std::complex<T>* out, int len, long istride, const int* fac,
           int twstride, const std::complex<T>* tw) const {
    const int p = fac[0];
    const int m = len / p;
    if (m == 1) {
      // Direct DFT of length p.
      for (int r = 0; r < p; ++r) {
        std::complex<T> acc(0, 0);
        for (int q = 0; q < p; ++q) acc += in[q * istride] * tw[((long)q * r * (n / p)) % n];
        out[r] = acc;
      }
      return;
    }
    for (int q = 0; q < p; ++q) rec(in + q * istride, out + q * m, m, istride * p, fac + 1, twstride * p, tw);
    std::complex<T> t[64];
    std::vector<std::complex<T>> tv;
    std::complex<T>* tp = t;
    if (p > 64) {
      tv.resize(p);
      tp = tv.data();
    }
    for (int k = 0; k < m; ++k) {
      for (int q = 0; q < p; ++q) {
        std::complex<T> v = out[q * m + k];
        tp[q] = (q == 0) ? v : v * tw[((long)q * k * twstride) % n];
      }
      for (int r = 0; r < p; ++r) {
        std::complex<T> acc = tp[0];
        for (int q = 1; q < p; ++q) acc += tp[q] * tw[((long)q * r * (n / p)) % n];
        out[k + r * m] = acc;
      }
    }
  }

  // In-place transform of a strided line x[0], x[stride], ...; sign -1 forward, +1 backward.
  void line(std::complex<T>* x, long stride, int sign, std::complex<T>* tmp_in, std::complex<T>* tmp_out) const {
    for (int i = 0; i < n; ++i) tmp_in[i] = x[i * stride];
    rec(tmp_in, tmp_out, n, 1, factors.data(), 1, sign < 0 ? tw_fwd.data() : tw_bwd.data());
    for (int i = 0; i < n; ++i) x[i * stride] = tmp_out[i];
  }
};

// One pass of 1-D transforms over the lines x[o * ostride + k + i * stride], i < n, for
// o < nouter and k < ninner. Consecutive k are gathered together in runs of up to 16, which
// keeps the strided passes on whole cache lines; threads take whole runs.
template <typename T>
void pass(const Fft1<T>& f, std::complex<T>* data, long nouter, long ostride, long ninner, long stride, int sign) {
  const int n = f.n;
  const int run = ninner < 16 ? (int)ninner : 16;
  const long per = (ninner + run - 1) / run;
#pragma omp parallel num_threads(g_fft_threads)
  {
    std::vector<std::complex<T>> blk((size_t)run * n), a(n), b(n);
#pragma omp for schedule(static)
    for (long r = 0; r < nouter * per; ++r) {
      const long o = r / per, k0 = (r - o * per) * run;
      const int cnt = (int)std::min<long>(run, ninner - k0);
      std::complex<T>* x0 = data + o * ostride + k0;
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < cnt; ++j) blk[(size_t)j * n + i] = x0[i * stride + j];
      for (int j = 0; j < cnt; ++j) f.line(&blk[(size_t)j * n], 1, sign, a.data(), b.data());
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < cnt; ++j) x0[i * stride + j] = blk[(size_t)j * n + i];
    }
  }
}

// r2c of one real field (Mx, My, Mz), x fastest, into (Mx/2+1, My, Mz) complex; unnormalised.
template <typename T>
void rfft3(const int32_t* M, const T* in, std::complex<T>* out) {
  const int Mx = M[0], My = M[1], Mz = M[2], Mh = Mx / 2 + 1;
  Fft1<T> fx(Mx), fy(My), fz(Mz);
  const long rows = (long)My * Mz;
#pragma omp parallel num_threads(g_fft_threads)
  {
    std::vector<std::complex<T>> a(Mx), b(Mx), row(Mx);
#pragma omp for schedule(static)
    for (long r = 0; r < rows; ++r) {
      for (int i = 0; i < Mx; ++i) row[i] = std::complex<T>(in[r * Mx + i], 0);
      fx.line(row.data(), 1, -1, a.data(), b.data());
      for (int k = 0; k < Mh; ++k) out[r * Mh + k] = row[k];
    }
  }
  pass(fy, out, Mz, (long)My * Mh, Mh, Mh, -1);             // y lines (k, z)
  pass(fz, out, 1, 0, (long)Mh * My, (long)My * Mh, -1);     // z lines (k, y)
}

// Unnormalised c2r (FFTW brfft): complex passes along z and y, then c2r along x that ignores
// the imaginary parts of the x DC (and, for even Mx, Nyquist) bins. `in` is overwritten.
template <typename T>
void brfft3(const int32_t* M, std::complex<T>* in, T* out) {
  const int Mx = M[0], My = M[1], Mz = M[2], Mh = Mx / 2 + 1;
  Fft1<T> fx(Mx), fy(My), fz(Mz);
  pass(fz, in, 1, 0, (long)Mh * My, (long)My * Mh, +1);      // z lines (k, y)
  pass(fy, in, Mz, (long)My * Mh, Mh, Mh, +1);              // y lines (k, z)
  const long rows = (long)My * Mz;
#pragma omp parallel num_threads(g_fft_threads)
  {
    std::vector<std::complex<T>> a(Mx), b(Mx), row(Mx);
#pragma omp for schedule(static)
    for (long r = 0; r < rows; ++r) {
      const std::complex<T>* X = in + r * Mh;
      for (int k = 0; k < Mx; ++k) {
        if (k < Mh)
          row[k] = X[k];
        else
          row[k] = std::conj(X[Mx - k]);
      }
      row[0] = std::complex<T>(X[0].real(), 0);
      if (Mx % 2 == 0) row[Mx / 2] = std::complex<T>(X[Mx / 2].real(), 0);
      fx.line(row.data(), 1, +1, a.data(), b.data());
      for (int i = 0; i < Mx; ++i) out[r * Mx + i] = row[i].real();
    }
  }
}

// _apply_inverse_stokes_kernel! (src/stokes_solve.jl:70-110), in place on 3 components.
template <typename T>
void project(const P<T>& p, std::complex<T>* fx, std::complex<T>* fy, std::complex<T>* fz, T inv_M) {
  const int Mh = p.M[0] / 2 + 1, My = p.M[1], Mz = p.M[2];
  const T inv_mu_M = inv_M / p.mu;
  for (int iz = 0; iz < Mz; ++iz) {
    T kz = p.kz[iz];
    for (int iy = 0; iy < My; ++iy) {
      T ky = p.ky[iy];
      for (int ix = 0; ix < Mh; ++ix) {
        T kx = p.kx[ix];
        long g = ix + (long)Mh * (iy + (long)My * iz);
        T k2 = kx * kx + ky * ky + kz * kz;
        if (k2 == 0) {
          fx[g] = fy[g] = fz[g] = std::complex<T>(0, 0);
          continue;
        }
        std::complex<T> f0 = fx[g], f1 = fy[g], f2 = fz[g];
        T inv_k2 = T(1) / k2;
        // dot(k, f) with real k: kx*f0 + ky*f1 + kz*f2
        std::complex<T> kdotf = kx * f0 + ky * f1 + kz * f2;
        std::complex<T> c = kdotf * inv_k2;
        T s = inv_mu_M * inv_k2;
        fx[g] = s * (f0 - kx * c);
        fy[g] = s * (f1 - ky * c);
        fz[g] = s * (f2 - kz * c);
      }
    }
  }
}

// stokes_solve! (src/stokes_solve.jl:21-42). fgrid, ugrid: 3 SoA real fields; fhat: 3 SoA
// half-complex fields (optional output of the projected spectrum).
template <typename T>
void stokes(const P<T>& p, const T* fgrid, T* ugrid, std::complex<T>* fhat_out) {
  const long M3 = (long)p.M[0] * p.M[1] * p.M[2];
  const long H = (long)(p.M[0] / 2 + 1) * p.M[1] * p.M[2];
  std::vector<std::complex<T>> own;
  std::complex<T>* fh = fhat_out;
  if (!fh) {
    own.resize(3 * H);
    fh = own.data();
  }
  for (int c = 0; c < 3; ++c) rfft3<T>(p.M, fgrid + c * M3, fh + c * H);
  const long Mtot = (long)p.M[0] * (long)p.M[1] * (long)p.M[2];
  const T inv_M = T(1) / (T)Mtot;
  project(p, fh, fh + H, fh + 2 * H, inv_M);
  std::vector<std::complex<T>> tmp(H);
  for (int c = 0; c < 3; ++c) {
    std::memcpy(tmp.data(), fh + c * H, sizeof(std::complex<T>) * H);
    brfft3<T>(p.M, tmp.data(), ugrid + c * M3);
  }
}

// ---------------------------------------------------------------- step 5: interpolation
// _interpolate_velocities_kernel! (src/interpolate.jl:82-136). orig is 1-based.
template <typename T>
void interp(const P<T>& p, const T* ugrid, const T* Ys, const int32_t* orig, T* V) {
  const int64_t M3 = (int64_t)p.M[0] * p.M[1] * p.M[2];
  T a0, a2, inv_norm, inv_2S2;
  kernel_coefficients(p.sigma, p.Sigma, &a0, &a2, &inv_norm, &inv_2S2);
  const T h3 = p.h * p.h * p.h;
  const int MG = p.MG;
  std::vector<T> gau(3 * MG), r2(3 * MG);
  std::vector<int32_t> idx(3 * MG);
  const T* ux = ugrid;
  const T* uy = ugrid + M3;
  const T* uz = ugrid + 2 * M3;
#pragma omp parallel for schedule(static) num_threads(g_particle_threads) firstprivate(gau, r2, idx)
  for (int64_t s = 0; s < p.N; ++s) {
    fill_stencil(p, Ys + 3 * s, inv_norm, inv_2S2, gau.data(), r2.data(), idx.data());
    T v0 = 0, v1 = 0, v2 = 0;
    for (int kz = 0; kz < MG; ++kz) {
      int32_t iz = idx[2 * MG + kz];
      T gz = gau[2 * MG + kz];
      T r2z = r2[2 * MG + kz];
      for (int ky = 0; ky < MG; ++ky) {
        int32_t iy = idx[MG + ky];
        T gyz = gau[MG + ky] * gz;
        T r2yz = r2[MG + ky] + r2z;
        for (int kx = 0; kx < MG; ++kx) {
          int32_t ix = idx[kx];
          T w = (a0 + a2 * (r2[kx] + r2yz)) * gau[kx] * gyz;
          int64_t g = ix + (int64_t)p.M[0] * (iy + (int64_t)p.M[1] * iz);
          v0 += w * ux[g];
          v1 += w * uy[g];
          v2 += w * uz[g];
        }
      }
    }
    int64_t n = orig[s] - 1;
    V[3 * n] = h3 * v0;
    V[3 * n + 1] = h3 * v1;
    V[3 * n + 2] = h3 * v2;
  }
}

// ---------------------------------------------------------------- step 6: correction
// _self_correction (src/correct_velocities.jl:21-27).
template <typename T>
T self_correction(T sigma, T Sigma, T a, T mu) {
  const T pi = pi_T<T>();
  T d = sigma * sigma - Sigma * Sigma;
  T Ssp = Sigma * std::sqrt(pi);
  // Julia `Σ^5` on Float32 evaluates the power in Float64 and rounds once; on Float64 it is
  // a compensated power by squaring. std::pow in double reproduces both to the last ulp
  // (float) or within one ulp (double).
  T S5 = (T)std::pow((double)Sigma, 5.0);
  return T(1) / (T(6) * pi * mu * a) - T(1) / (T(6) * pi * mu * Ssp) +
         d / (T(12) * mu * (Ssp * Ssp * Ssp)) - (d * d) / (T(32) * mu * S5 * std::sqrt(pi * pi * pi));
}

// _correction_scalars (src/correct_velocities.jl:71-104).
template <typename T>
void correction_scalars(T r, T sigma, T Sigma, T mu, T* A, T* B) {
  const T pi = pi_T<T>();
  T r2 = r * r;
  T r3 = r * r * r;
  T r4 = r2 * r2;
  T r5 = r4 * r;
  T s2 = sigma * sigma;
  T s4 = s2 * s2;
  T S2 = Sigma * Sigma;
  T S4 = S2 * S2;
  T d = s2 - S2;
  T dq = d * d / T(4);
  T fourpimu = T(4) * pi * mu;
  T eightpimu = T(8) * pi * mu;
  T erf_diff = std::erf(r / (T(2) * sigma)) - std::erf(r / (T(2) * Sigma));
  T Ds = std::pow(T(4) * pi * s2, -T(3) / T(2)) * std::exp(-r2 / (T(4) * s2));
  T DS = std::pow(T(4) * pi * S2, -T(3) / T(2)) * std::exp(-r2 / (T(4) * S2));
  *A = erf_diff * (T(1) / (eightpimu * r) + s2 / (fourpimu * r3)) - T(2) * s4 / (mu * r2) * Ds +
       (T(2) * S4 / (mu * r2) + d / mu * (T(1) + T(2) * S2 / r2) -
        dq * (T(2) - r2 / (T(2) * S2)) / (T(2) * mu * S2)) *
           DS;
  *B = erf_diff * (T(1) / (eightpimu * r3) - T(3) * s2 / (fourpimu * r5)) + T(6) * s4 / (mu * r4) * Ds +
       (-T(6) * S4 / (mu * r4) - d / (mu * r2) * (T(1) + T(6) * S2 / r2) - dq / (T(4) * mu * S4)) * DS;
}

// _correct_velocities_kernel! CPU method (src/correct_velocities.jl:216-286). start/end/orig
// are 1-based like the reference. V is updated in place (+=).
template <typename T>
void correct(const P<T>& p, const T* Ys, const T* Fs, const int32_t* cstart, const int32_t* cend,
             const int32_t* orig, T* V) {
  const T delta = self_correction(p.sigma, p.Sigma, p.a, p.mu);
  const T Rc2 = p.Rc * p.Rc;
  const int32_t mx = p.nc[0], my = p.nc[1], mz = p.nc[2];
  // each particle's sum is its own (one particle per thread at a time, cells in the serial
  // order inside a thread's share)
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_particle_threads)
  for (int32_t cz = 0; cz < mz; ++cz)
    for (int32_t cy = 0; cy < my; ++cy)
      for (int32_t cx = 0; cx < mx; ++cx) {
        int32_t c = cx + (cy + cz * my) * mx;
        for (int32_t s = cstart[c]; s <= cend[c]; ++s) {
          const T* Yn = Ys + 3 * (int64_t)(s - 1);
          const T* Fn = Fs + 3 * (int64_t)(s - 1);
          T v0 = delta * Fn[0], v1 = delta * Fn[1], v2 = delta * Fn[2];
          for (int32_t dz = -1; dz <= 1; ++dz) {
            int32_t nz = jl_mod_int(cz + dz, mz);
            for (int32_t dy = -1; dy <= 1; ++dy) {
              int32_t ny = jl_mod_int(cy + dy, my);
              for (int32_t dx = -1; dx <= 1; ++dx) {
                int32_t nx = jl_mod_int(cx + dx, mx);
                int32_t cn = nx + (ny + nz * my) * mx;
                for (int32_t s2 = cstart[cn]; s2 <= cend[cn]; ++s2) {
                  if (s2 == s) continue;
                  const T* Ym = Ys + 3 * (int64_t)(s2 - 1);
                  T x[3];
                  for (int i = 0; i < 3; ++i) {
                    T d = Yn[i] - Ym[i];
                    x[i] = d - p.L[i] * std::nearbyint(d / p.L[i]);
                  }
                  T rr = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
                  if (!(rr < Rc2)) continue;
                  T A, B;
                  correction_scalars(std::sqrt(rr), p.sigma, p.Sigma, p.mu, &A, &B);
                  const T* Fm = Fs + 3 * (int64_t)(s2 - 1);
                  T xdotF = x[0] * Fm[0] + x[1] * Fm[1] + x[2] * Fm[2];
                  T bx = B * xdotF;
                  v0 += A * Fm[0] + bx * x[0];
                  v1 += A * Fm[1] + bx * x[1];
                  v2 += A * Fm[2] + bx * x[2];
                }
              }
            }
          }
          int64_t n = orig[s - 1] - 1;
          V[3 * n] += v0;
          V[3 * n + 1] += v1;
          V[3 * n + 2] += v2;
        }
      }
}

// mobility! (src/mobility.jl:23-45) with the CPU staging identity (src/mobility.jl:55-56).
template <typename T>
int mobility(const P<T>& p, const T* Y, const T* F, T* V) {
  const int64_t N = p.N;
  const int64_t C = (int64_t)p.nc[0] * p.nc[1] * p.nc[2];
  const int64_t M3 = (int64_t)p.M[0] * p.M[1] * p.M[2];
  std::vector<T> Yw(3 * N), Ys(3 * N), Fs(3 * N);
  std::vector<int32_t> hash(N), orig(N), cs(C), ce(C);
  wrap(p, Y, Yw.data());
  assign_cells(p, Yw.data(), hash.data());
  build_cell_list(N, C, hash.data(), orig.data(), cs.data(), ce.data());
  gather(N, orig.data(), Yw.data(), F, Ys.data(), Fs.data());
  std::vector<T> fgrid(3 * M3), ugrid(3 * M3);
  spread(p, Ys.data(), Fs.data(), fgrid.data());
  stokes(p, fgrid.data(), ugrid.data(), (std::complex<T>*)nullptr);
  fgrid.clear();
  fgrid.shrink_to_fit();
  interp(p, ugrid.data(), Ys.data(), orig.data(), V);
  correct(p, Ys.data(), Fs.data(), cs.data(), ce.data(), orig.data(), V);
  return 0;
}

// ---------------------------------------------------------------- adjoint (rrule pullback)
// Not a reference function: the derivative of mobility! for L = <G, M(Y) F>. Pinned by float64
// central differences of mobility() above (tests/test_oracle.py), not by retyped formulas.

// r-derivatives of correction_scalars: A = E a1 + Ds a2 + DS a3 and B = E b1 + Ds b2 + DS b3,
// differentiated factor by factor (E' from erf' = 2/sqrt(pi) exp(-z^2); D' = -r D / (2 w^2)).
template <typename T>
void correction_scalars_deriv(T r, T sigma, T Sigma, T mu, T* dA, T* dB) {
  const T pi = pi_T<T>();
  const T s2 = sigma * sigma, S2 = Sigma * Sigma;
  const T s4 = s2 * s2, S4 = S2 * S2;
  const T d = s2 - S2, dq = d * d / T(4);
  const T r2 = r * r, r3 = r2 * r, r4 = r2 * r2, r5 = r4 * r, r6 = r3 * r3;
  const T gs = std::exp(-r2 / (T(4) * s2)), gS = std::exp(-r2 / (T(4) * S2));
  const T E = std::erf(r / (T(2) * sigma)) - std::erf(r / (T(2) * Sigma));
  const T Ep = (gs / sigma - gS / Sigma) / std::sqrt(pi);
  const T Ds = std::pow(T(4) * pi * s2, -T(3) / T(2)) * gs;
  const T DS = std::pow(T(4) * pi * S2, -T(3) / T(2)) * gS;
  const T Dsp = -r / (T(2) * s2) * Ds;
  const T DSp = -r / (T(2) * S2) * DS;
  const T c8 = T(8) * pi * mu, c4 = T(4) * pi * mu;
  const T a1 = T(1) / (c8 * r) + s2 / (c4 * r3);
  const T a1p = -T(1) / (c8 * r2) - T(3) * s2 / (c4 * r4);
  const T a2 = -T(2) * s4 / (mu * r2);
  const T a2p = T(4) * s4 / (mu * r3);
  const T a3 = T(2) * S4 / (mu * r2) + d / mu * (T(1) + T(2) * S2 / r2) - dq * (T(2) - r2 / (T(2) * S2)) / (T(2) * mu * S2);
  const T a3p = -T(4) * S4 / (mu * r3) - T(4) * d * S2 / (mu * r3) + dq * r / (T(2) * mu * S2 * S2);
  *dA = Ep * a1 + E * a1p + Dsp * a2 + Ds * a2p + DSp * a3 + DS * a3p;
  const T b1 = T(1) / (c8 * r3) - T(3) * s2 / (c4 * r5);
  const T b1p = -T(3) / (c8 * r4) + T(15) * s2 / (c4 * r6);
  const T b2 = T(6) * s4 / (mu * r4);
  const T b2p = -T(24) * s4 / (mu * r5);
  const T b3 = -T(6) * S4 / (mu * r4) - d / (mu * r2) * (T(1) + T(6) * S2 / r2) - dq / (T(4) * mu * S4);
  const T b3p = T(24) * S4 / (mu * r5) + d / mu * (T(2) / r3 + T(24) * S2 / r5);
  *dB = Ep * b1 + E * b1p + Dsp * b2 + Ds * b2p + DSp * b3 + DS * b3p;
}

// dF = M(Y) G (M symmetric); dY = h^3 sum_g dW/dY [G . u_F + F . u_G] + pair terms, with
// dW/dY_i = x_i gx gy gz ((a0 + a2 r^2) / Sigma^2 - 2 a2) and, per pair within R_c,
// grad_x[G_n T F_m + G_m T F_n] for T = A I + B x x^T. Outputs are (N, 3) in original order.
template <typename T>
int mobility_vjp(const P<T>& p, const T* Y, const T* F, const T* G, T* dF, T* dY) {
  const int64_t N = p.N;
  const int64_t C = (int64_t)p.nc[0] * p.nc[1] * p.nc[2];
  const int64_t M3 = (int64_t)p.M[0] * p.M[1] * p.M[2];
  mobility(p, Y, G, dF);
  std::vector<T> Yw(3 * N), Ys(3 * N), Fs(3 * N), Gs(3 * N);
  std::vector<int32_t> hash(N), orig(N), cs(C), ce(C);
  wrap(p, Y, Yw.data());
  assign_cells(p, Yw.data(), hash.data());
  build_cell_list(N, C, hash.data(), orig.data(), cs.data(), ce.data());
  gather(N, orig.data(), Yw.data(), F, Ys.data(), Fs.data());
  gather(N, orig.data(), Yw.data(), G, Ys.data(), Gs.data());
  std::vector<T> fgrid(3 * M3), uF(3 * M3), uG(3 * M3);
  spread(p, Ys.data(), Fs.data(), fgrid.data());
  stokes(p, fgrid.data(), uF.data(), (std::complex<T>*)nullptr);
  spread(p, Ys.data(), Gs.data(), fgrid.data());
  stokes(p, fgrid.data(), uG.data(), (std::complex<T>*)nullptr);
  T a0, a2, inv_norm, inv_2S2;
  kernel_coefficients(p.sigma, p.Sigma, &a0, &a2, &inv_norm, &inv_2S2);
  const T inv_S2 = T(1) / (p.Sigma * p.Sigma);
  const T h3 = p.h * p.h * p.h;
  const int MG = p.MG;
  std::vector<T> gau(3 * MG), r2(3 * MG), xs(3 * MG);
  std::vector<int32_t> idx(3 * MG);
  for (int64_t s = 0; s < N; ++s) {
    const T* y = Ys.data() + 3 * s;
    const T* Fn = Fs.data() + 3 * s;
    const T* Gn = Gs.data() + 3 * s;
    fill_stencil(p, y, inv_norm, inv_2S2, gau.data(), r2.data(), idx.data());
    for (int i = 0; i < 3; ++i) {
      const int32_t j = jl_round_i32(y[i] * p.inv_h);
      for (int k = 0; k < MG; ++k) xs[i * MG + k] = (T)(j + (k - MG / 2)) * p.h - y[i];
    }
    T acc[3] = {0, 0, 0};
    for (int kz = 0; kz < MG; ++kz)
      for (int ky = 0; ky < MG; ++ky)
        for (int kx = 0; kx < MG; ++kx) {
          const T rr = r2[kx] + r2[MG + ky] + r2[2 * MG + kz];
          const T w = gau[kx] * gau[MG + ky] * gau[2 * MG + kz] * ((a0 + a2 * rr) * inv_S2 - T(2) * a2);
          const int64_t g = idx[kx] + (int64_t)p.M[0] * (idx[MG + ky] + (int64_t)p.M[1] * idx[2 * MG + kz]);
          T val = 0;
          for (int c = 0; c < 3; ++c) val += Gn[c] * uF[c * M3 + g] + Fn[c] * uG[c * M3 + g];
          acc[0] += xs[kx] * w * val;
          acc[1] += xs[MG + ky] * w * val;
          acc[2] += xs[2 * MG + kz] * w * val;
        }
    const int64_t n = orig[s] - 1;
    for (int c = 0; c < 3; ++c) dY[3 * n + c] = h3 * acc[c];
  }
  // pair part over the same 27-cell neighbourhood as correct()
  const T Rc2 = p.Rc * p.Rc;
  const int32_t mx = p.nc[0], my = p.nc[1], mz = p.nc[2];
  for (int32_t cz = 0; cz < mz; ++cz)
    for (int32_t cy = 0; cy < my; ++cy)
      for (int32_t cx = 0; cx < mx; ++cx) {
        const int32_t c = cx + (cy + cz * my) * mx;
        for (int32_t s = cs[c]; s <= ce[c]; ++s) {
          const T* Yn = Ys.data() + 3 * (int64_t)(s - 1);
          const T* Fn = Fs.data() + 3 * (int64_t)(s - 1);
          const T* Gn = Gs.data() + 3 * (int64_t)(s - 1);
          T acc[3] = {0, 0, 0};
          for (int32_t dz = -1; dz <= 1; ++dz)
            for (int32_t dy = -1; dy <= 1; ++dy)
              for (int32_t dx = -1; dx <= 1; ++dx) {
                const int32_t cn = jl_mod_int(cx + dx, mx) + (jl_mod_int(cy + dy, my) + jl_mod_int(cz + dz, mz) * my) * mx;
                for (int32_t s2 = cs[cn]; s2 <= ce[cn]; ++s2) {
                  if (s2 == s) continue;
                  const T* Ym = Ys.data() + 3 * (int64_t)(s2 - 1);
                  const T* Fm = Fs.data() + 3 * (int64_t)(s2 - 1);
                  const T* Gm = Gs.data() + 3 * (int64_t)(s2 - 1);
                  T x[3];
                  for (int i = 0; i < 3; ++i) {
                    T dd = Yn[i] - Ym[i];
                    x[i] = dd - p.L[i] * std::nearbyint(dd / p.L[i]);
                  }
                  const T rr = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
                  if (!(rr < Rc2)) continue;
                  const T r = std::sqrt(rr);
                  T A, B, dA, dB;
                  correction_scalars(r, p.sigma, p.Sigma, p.mu, &A, &B);
                  correction_scalars_deriv(r, p.sigma, p.Sigma, p.mu, &dA, &dB);
                  // two bilinear forms a^T T(x) b: (G_n, F_m) and (G_m, F_n)
                  const T* pa[2] = {Gn, Gm};
                  const T* pb[2] = {Fm, Fn};
                  for (int t = 0; t < 2; ++t) {
                    const T* a = pa[t];
                    const T* b = pb[t];
                    const T ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
                    const T xa = x[0] * a[0] + x[1] * a[1] + x[2] * a[2];
                    const T xb = x[0] * b[0] + x[1] * b[1] + x[2] * b[2];
                    for (int i = 0; i < 3; ++i)
                      acc[i] += (dA * ab + dB * xa * xb) * x[i] / r + B * (xb * a[i] + xa * b[i]);
                  }
                }
              }
          const int64_t n = orig[s - 1] - 1;
          for (int i = 0; i < 3; ++i) dY[3 * n + i] += acc[i];
        }
      }
  return 0;
}

}  // namespace

#define ORACLE_EXPORTS(T, SUF)                                                                        \
  int oracle_wrap_##SUF(const oracle_params* p, const T* Y, T* Yw) {                                  \
    wrap(P<T>(p), Y, Yw);                                                                             \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_assign_cells_##SUF(const oracle_params* p, const T* Yw, int32_t* hash) {                 \
    assign_cells(P<T>(p), Yw, hash);                                                                  \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_gather_##SUF(const oracle_params* p, const int32_t* orig, const T* Y, const T* F, T* Ys, \
                          T* Fs) {                                                                    \
    gather<T>(p->N, orig, Y, F, Ys, Fs);                                                              \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_spread_##SUF(const oracle_params* p, const T* Ys, const T* Fs, T* grid) {                \
    spread(P<T>(p), Ys, Fs, grid);                                                                    \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_rfft3_##SUF(const int32_t* M, const T* in, T* out_cplx) {                                \
    rfft3<T>(M, in, (std::complex<T>*)out_cplx);                                                      \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_brfft3_##SUF(const int32_t* M, T* in_cplx, T* out) {                                     \
    brfft3<T>(M, (std::complex<T>*)in_cplx, out);                                                     \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_stokes_##SUF(const oracle_params* p, const T* fgrid, T* ugrid, T* fhat) {                \
    stokes(P<T>(p), fgrid, ugrid, (std::complex<T>*)fhat);                                            \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_interp_##SUF(const oracle_params* p, const T* ugrid, const T* Ys, const int32_t* orig,   \
                          T* V) {                                                                     \
    interp(P<T>(p), ugrid, Ys, orig, V);                                                              \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_correct_##SUF(const oracle_params* p, const T* Ys, const T* Fs, const int32_t* cs,       \
                           const int32_t* ce, const int32_t* orig, T* V) {                            \
    correct(P<T>(p), Ys, Fs, cs, ce, orig, V);                                                        \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_mobility_##SUF(const oracle_params* p, const T* Y, const T* F, T* V) {                   \
    return mobility(P<T>(p), Y, F, V);                                                                \
  }                                                                                                   \
  int oracle_mobility_vjp_##SUF(const oracle_params* p, const T* Y, const T* F, const T* G, T* dF, T* dY) { \
    return mobility_vjp(P<T>(p), Y, F, G, dF, dY);                                                    \
  }                                                                                                   \
  int oracle_correction_scalars_deriv_##SUF(T r, T sigma, T Sigma, T mu, T* dA, T* dB) {              \
    correction_scalars_deriv(r, sigma, Sigma, mu, dA, dB);                                            \
    return 0;                                                                                         \
  }                                                                                                   \
  int oracle_correction_scalars_##SUF(T r, T sigma, T Sigma, T mu, T* A, T* B) {                      \
    correction_scalars(r, sigma, Sigma, mu, A, B);                                                    \
    return 0;                                                                                         \
  }                                                                                                   \
  T oracle_self_correction_##SUF(T sigma, T Sigma, T a, T mu) { return self_correction(sigma, Sigma, a, mu); } \
  int oracle_kernel_coefficients_##SUF(T sigma, T Sigma, T* out4) {                                   \
    kernel_coefficients(sigma, Sigma, out4, out4 + 1, out4 + 2, out4 + 3);                            \
    return 0;                                                                                         \
  }

extern "C" {
ORACLE_EXPORTS(float, f32)
ORACLE_EXPORTS(double, f64)

// Threads of the transforms and of the particle loops (1 and 1 by default; see the header).
int oracle_set_threads(int fft_threads, int particle_threads) {
  g_fft_threads = fft_threads < 1 ? 1 : fft_threads;
  g_particle_threads = particle_threads < 1 ? 1 : particle_threads;
  return 0;
}

int oracle_build_cell_list(int64_t N, int64_t total, const int32_t* hash, int32_t* orig, int32_t* cs,
                           int32_t* ce) {
  build_cell_list(N, total, hash, orig, cs, ce);
  return 0;
}
}
