// Step 4 of the mobility apply: the Stokes solve as a hand-written 3-D real FFT.
//
// Reference: stokes_solve! (src/stokes_solve.jl:21-42) with FFTW plan_rfft / plan_brfft
// (src/config.jl:418-423) and the projection _apply_inverse_stokes_kernel!
// (src/stokes_solve.jl:70-110). Conventions kept from the reference:
//   - forward transforms use exp(-2 pi i jk/n), unnormalised; the backward transform is the
//     unnormalised c2r and the 1/(Mx My Mz) factor lives in the projection scalar inv_M/mu;
//   - the projection is f <- (inv_M/mu * 1/k^2) (f - k (k.f)/k^2) with k = 0 set to exactly 0
//     and the wavevectors passed in verbatim (Nyquist entries are +pi/h);
//   - the c2r treats each x row as Hermitian, using the real parts of its DC and (even Mx)
//     Nyquist bins only, after complex inverse passes along z and y.
//
// Structure (five HBM passes over the spectrum):
//   x_fwd   real rows -> half-spectrum rows; two real rows share one complex transform
//   y       complex transforms along y (forward)
//   z       forward along z, projection of the three components, inverse along z, fused
//   y       complex transforms along y (inverse)
//   x_inv   half-spectrum rows -> real rows, two rows per complex transform
// Each block stages its pencils in LDS; the transforms are Stockham radix-2/3/4/5/7/8/11/13
// stages executed from registers with the LDS array as the exchange buffer.
#include <cmath>

#include "ffcm_internal.hpp"

namespace ffcm {

namespace {

constexpr int kElemMax = 16;  // complex values a thread holds per stage

template <typename T>
using C2 = typename Vec<T>::c2;
template <typename T>
struct C4 {
  C2<T> a, b;
};
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
};
template <>
struct Vec16<double> {
  using type = double2;
};
template <typename T>
constexpr int VE = 16 / (int)sizeof(T);  // elements per 16-byte vector

template <typename V>
__device__ __forceinline__ V cadd(V a, V b) {
  return V{a.x + b.x, a.y + b.y};
}
template <typename V>
__device__ __forceinline__ V csub(V a, V b) {
  return V{a.x - b.x, a.y - b.y};
}
template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
  return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename V, typename T>
__device__ __forceinline__ V cscale(V a, T s) {
  return V{a.x * s, a.y * s};
}
// z * (S i)
template <int S, typename V>
__device__ __forceinline__ V rot(V z) {
  if constexpr (S < 0)
    return V{z.y, -z.x};
  else
    return V{-z.y, z.x};
}
// Twiddle exp(S 2 pi i m / n) from the forward table exp(-2 pi i m / n).
template <int S, typename V>
__device__ __forceinline__ V twid(const V* tw, int m) {
  V w = tw[m];
  if constexpr (S > 0) w.y = -w.y;
  return w;
}

// In-register DFTs of size R with sign S.
template <int R, int S, typename T>
struct Dft {
  // Generic small prime: X_k = sum_r v_r W^{rk}, W = exp(S 2 pi i / R) = tw[(n/R) * 1].
  __device__ static void run(C2<T>* v, const C2<T>* tw, int n) {
    C2<T> out[R];
    const int step = n / R;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      C2<T> acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(v[r], twid<S>(tw, ((r * k) % R) * step)));
      out[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
};

template <int S, typename T>
struct Dft<2, S, T> {
  __device__ static void run(C2<T>* v, const C2<T>*, int) {
    C2<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <int S, typename T>
struct Dft<3, S, T> {
  __device__ static void run(C2<T>* v, const C2<T>*, int) {
    const T c = T(0.86602540378443864676372317075294);  // sin(2 pi / 3)
    C2<T> t1 = cadd(v[1], v[2]);
    C2<T> t2 = csub(v[0], cscale(t1, T(0.5)));
    C2<T> t3 = cscale(rot<S>(csub(v[1], v[2])), c);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  }
};

template <int S, typename T>
struct Dft<4, S, T> {
  __device__ static void run(C2<T>* v, const C2<T>*, int) {
    C2<T> t0 = cadd(v[0], v[2]);
    C2<T> t1 = csub(v[0], v[2]);
    C2<T> t2 = cadd(v[1], v[3]);
    C2<T> t3 = rot<S>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3);
    v[3] = csub(t1, t3);
  }
};

template <int S, typename T>
struct Dft<5, S, T> {
  __device__ static void run(C2<T>* v, const C2<T>*, int) {
    const T c1 = T(0.30901699437494742410229341718282);   // cos(2 pi / 5)
    const T c2 = T(-0.80901699437494742410229341718282);  // cos(4 pi / 5)
    const T s1 = T(0.95105651629515357211643933337938);   // sin(2 pi / 5)
    const T s2 = T(0.58778525229247312916870595463907);   // sin(4 pi / 5)
    C2<T> b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    C2<T> d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    C2<T> a0 = v[0];
    C2<T> p1 = cadd(a0, cadd(cscale(b1, c1), cscale(b2, c2)));
    C2<T> p2 = cadd(a0, cadd(cscale(b1, c2), cscale(b2, c1)));
    C2<T> q1 = rot<S>(cadd(cscale(d1, s1), cscale(d2, s2)));
    C2<T> q2 = rot<S>(csub(cscale(d1, s2), cscale(d2, s1)));
    v[0] = cadd(a0, cadd(b1, b2));
    v[1] = cadd(p1, q1);
    v[4] = csub(p1, q1);
    v[2] = cadd(p2, q2);
    v[3] = csub(p2, q2);
  }
};

template <int S, typename T>
struct Dft<8, S, T> {
  __device__ static void run(C2<T>* v, const C2<T>*, int) {
    const T r = T(0.70710678118654752440084436210485);
    C2<T> e[4] = {v[0], v[2], v[4], v[6]};
    C2<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, S, T>::run(e, nullptr, 0);
    Dft<4, S, T>::run(o, nullptr, 0);
    // o_k *= exp(S 2 pi i k / 8)
    o[1] = cscale(cadd(o[1], rot<S>(o[1])), r);
    o[2] = rot<S>(o[2]);
    o[3] = cscale(csub(rot<S>(o[3]), o[3]), r);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};

// Workgroup barrier that orders LDS only: global loads in flight (the next tile's prefetch)
// are not drained by it.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Pencil geometry of one block: element t of pencil p lives at buf[p * ps + t * es].
struct Geo {
  int np, tp;      // pencils, threads per pencil
  int my_p, my_t;  // this thread's pencil and lane within the pencil
  int ps, es;
};

// One Stockham stage (Ns = product of the earlier radices): butterfly j reads
// x[j + r n/R] * exp(S 2 pi i r (j mod Ns) / (Ns R)), applies the size-R DFT and writes
// x[(j - j mod Ns) R + j mod Ns + r Ns]. Reads and writes are separated by a block barrier.
template <int R, int S, typename T>
__device__ __forceinline__ void stage(C2<T>* buf, int n, int Ns, const Geo& g, const C2<T>* tw) {
  constexpr int K = (kElemMax / R) > 0 ? (kElemMax / R) : 1;
  const int nb = n / R;
  const int tstep = n / (Ns * R);
  C2<T> v[K][R];
  C2<T>* base = buf + g.my_p * g.ps;
  const bool active = g.my_p < g.np;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = g.my_t + k * g.tp;
    if (active && j < nb) {
      const int jm = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[k][r] = base[(j + r * nb) * g.es];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[k][r] = cmul(v[k][r], twid<S>(tw, r * jm * tstep));
      }
      Dft<R, S, T>::run(v[k], tw, n);
    }
  }
  lds_sync();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = g.my_t + k * g.tp;
    if (active && j < nb) {
      const int jm = j % Ns;
      const int o = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) base[(o + r * Ns) * g.es] = v[k][r];
    }
  }
  lds_sync();
}

// Full transform of every pencil of the block, in place in LDS. Ends with a barrier.
template <int S, typename T>
__device__ __noinline__ void lds_fft(C2<T>* buf, const FftAxis& ax, const Geo& g, const C2<T>* tw) {
  int Ns = 1;
  for (int st = 0; st < ax.nstages; ++st) {
    const int R = (int)((ax.packed >> (4 * st)) & 15ull);
    switch (R) {
      case 2: stage<2, S, T>(buf, ax.n, Ns, g, tw); break;
      case 3: stage<3, S, T>(buf, ax.n, Ns, g, tw); break;
      case 4: stage<4, S, T>(buf, ax.n, Ns, g, tw); break;
      case 5: stage<5, S, T>(buf, ax.n, Ns, g, tw); break;
      case 7: stage<7, S, T>(buf, ax.n, Ns, g, tw); break;
      case 8: stage<8, S, T>(buf, ax.n, Ns, g, tw); break;
      case 11: stage<11, S, T>(buf, ax.n, Ns, g, tw); break;
      default: stage<13, S, T>(buf, ax.n, Ns, g, tw); break;
    }
    Ns *= R;
  }
}


// Compile-time radix plans for the common sizes: the stage loop is unrolled with constant
// strides, so the kernels need no call and keep the next tile's prefetch in registers.
struct DynPlan {};
template <int... Rs>
struct Radices {};

template <int R, int S, typename T, int N, int Ns>
__device__ __forceinline__ void stage_ct(C2<T>* buf, const Geo& g, const C2<T>* tw) {
  constexpr int K = (kElemMax / R) > 0 ? (kElemMax / R) : 1;
  constexpr int nb = N / R;
  constexpr int tstep = N / (Ns * R);
  C2<T> v[K][R];
  C2<T>* base = buf + g.my_p * g.ps;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = g.my_t + k * g.tp;
    if (j < nb) {
      const int jm = j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[k][r] = base[(j + r * nb) * g.es];
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[k][r] = cmul(v[k][r], twid<S>(tw, r * jm * tstep));
      }
      Dft<R, S, T>::run(v[k], tw, N);
    }
  }
  lds_sync();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = g.my_t + k * g.tp;
    if (j < nb) {
      const int jm = j % Ns;
      const int o = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) base[(o + r * Ns) * g.es] = v[k][r];
    }
  }
  lds_sync();
}

template <int S, typename T, int N, int Ns, int R, int... Rest>
__device__ __forceinline__ void stages_ct(C2<T>* buf, const Geo& g, const C2<T>* tw) {
  stage_ct<R, S, T, N, Ns>(buf, g, tw);
  if constexpr (sizeof...(Rest) > 0) stages_ct<S, T, N, Ns * R, Rest...>(buf, g, tw);
}

template <int S, typename T, typename Plan>
struct FftRun;
template <int S, typename T>
struct FftRun<S, T, DynPlan> {
  __device__ static void run(C2<T>* buf, const FftAxis& ax, const Geo& g, const C2<T>* tw) {
    lds_fft<S, T>(buf, ax, g, tw);
  }
};
template <int S, typename T, int... Rs>
struct FftRun<S, T, Radices<Rs...>> {
  __device__ __forceinline__ static void run(C2<T>* buf, const FftAxis&, const Geo& g, const C2<T>* tw) {
    stages_ct<S, T, (Rs * ...), 1, Rs...>(buf, g, tw);
  }
};

template <typename V>
__device__ __forceinline__ void load_twiddles(V* tw, const V* __restrict__ g, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = g[i];
}

// All four pass kernels are persistent: block b transforms tiles b, b + gridDim, ... While
// tile t is transformed in LDS, the loads of tile t + gridDim are already in flight into
// registers (kEPT values per thread), so HBM latency overlaps the arithmetic.
constexpr int kEPT = 16;  // complex values of one tile per thread (n / tp <= 16)

// Advances (row, col) of a row-major walk with row length `len` by `step` elements.
__device__ __forceinline__ void walk(int& row, int& col, int step, int len) {
  col += step;
  while (col >= len) {
    col -= len;
    ++row;
  }
}

// ------------------------------------------------------------------------------ x forward
// Tile: one z plane, 2P consecutive y rows; rows (2p, 2p+1) form one complex pencil per
// component: z[x] = a[x] + i b[x]. Then A = (Z[k] + conj Z[-k]) / 2, B = (Z[k] - conj Z[-k]) / 2i.
template <typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_x_fwd_kernel(const T* __restrict__ grid, C2<T>* __restrict__ spec,
                                                        FftAxis ax, int My, int Mz, int P, int tp,
                                                        const C2<T>* __restrict__ tw_g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Mx = ax.n, Mh = Mx / 2 + 1;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + Mx;
  const int pairs = (My + 1) / 2;
  const int groups = (pairs + P - 1) / P;
  const int ntiles = Mz * groups;
  const int nthr = blockDim.x;
  const int rowlen = 3 * Mx;
  const int count = 2 * P * rowlen;  // = 2 nthr Mx / tp <= 2 kEPT nthr
  load_twiddles(tw, tw_g, Mx);
  Geo g;
  g.np = 3 * P;
  g.tp = tp;
  g.my_t = threadIdx.x % tp;
  g.my_p = threadIdx.x / tp;
  g.ps = Mx;
  g.es = 1;
  T pre[2 * kEPT];
  auto fetch = [&](int tile) {
    const int z = tile / groups;
    const int y0 = (tile - z * groups) * 2 * P;
    const int valid = min(2 * P, My - y0) * rowlen;
    const T* rows = grid + ((size_t)z * My + y0) * (size_t)rowlen;
#pragma unroll
    for (int u = 0; u < 2 * kEPT; ++u) {
      const int e = u * nthr + (int)threadIdx.x;
      if (e < count) pre[u] = e < valid ? rows[e] : T(0);
    }
  };
  int tile = blockIdx.x;
  if (tile < ntiles) fetch(tile);
  while (tile < ntiles) {
    {
      int rr = 0, f = threadIdx.x;
      walk(rr, f, 0, rowlen);
#pragma unroll
      for (int u = 0; u < 2 * kEPT; ++u) {
        if (u * nthr + (int)threadIdx.x < count) {
          const int x = f / 3;
          const int c = f - 3 * x;
          ((T*)&buf[(3 * (rr >> 1) + c) * Mx + x])[rr & 1] = pre[u];
        }
        walk(rr, f, nthr, rowlen);
      }
    }
    lds_sync();
    const int next = tile + gridDim.x;
    if (next < ntiles) fetch(next);
    FftRun<-1, T, Plan>::run(buf, ax, g, tw);
    const int z = tile / groups;
    const int y0 = (tile - z * groups) * 2 * P;
    for (int p = 0; p < P; ++p) {
      const int ya = y0 + 2 * p, yb = ya + 1;
      if (ya >= My) break;
      C2<T>* ra = spec + ((size_t)z * My + ya) * (size_t)Mh * 3;
      C2<T>* rb = spec + ((size_t)z * My + yb) * (size_t)Mh * 3;
      for (int e = threadIdx.x; e < 3 * Mh; e += nthr) {
        const int kx = e / 3;
        const int c = e - 3 * kx;
        const C2<T>* pen = buf + (3 * p + c) * Mx;
        const C2<T> Z = pen[kx];
        const C2<T> Zm = pen[kx == 0 ? 0 : Mx - kx];
        // conj(Zm) = (Zm.x, -Zm.y)
        ra[e] = C2<T>{(Z.x + Zm.x) * T(0.5), (Z.y - Zm.y) * T(0.5)};
        if (yb < My) rb[e] = C2<T>{(Z.y + Zm.y) * T(0.5), (Zm.x - Z.x) * T(0.5)};
      }
    }
    lds_sync();
    tile = next;
  }
}

// ------------------------------------------------------------------------------ y passes
// Tile: one z plane, W consecutive columns of the interleaved (kx, component) row of length
// Q = 3 Mh; pencils run along y with stride Q. LDS layout [y][W].
template <int S, typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_y_kernel(C2<T>* __restrict__ spec, FftAxis ax, int Q, int W, int lw,
                                                    int tp, int Mz, const C2<T>* __restrict__ tw_g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int My = ax.n;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + My;
  const int groups = (Q + W - 1) / W;
  const int ntiles = Mz * groups;
  const int nthr = blockDim.x;
  const int count = My * W;
  load_twiddles(tw, tw_g, My);
  Geo g;
  g.np = W;
  g.tp = tp;
  g.my_p = threadIdx.x & (W - 1);
  g.my_t = threadIdx.x >> lw;
  g.ps = 1;
  g.es = W;
  C2<T> pre[kEPT];
  auto fetch = [&](int tile) {
    const int z = tile / groups;
    const int q0 = (tile - z * groups) * W;
    const int wq = min(W, Q - q0);
    const C2<T>* plane = spec + (size_t)z * My * Q + q0;
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
      const int e = u * nthr + (int)threadIdx.x;
      if (e < count) {
        const int w = e & (W - 1);
        pre[u] = w < wq ? plane[(size_t)(e >> lw) * Q + w] : C2<T>{T(0), T(0)};
      }
    }
  };
  int tile = blockIdx.x;
  if (tile < ntiles) fetch(tile);
  while (tile < ntiles) {
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
      const int e = u * nthr + (int)threadIdx.x;
      if (e < count) buf[e] = pre[u];
    }
    lds_sync();
    const int next = tile + gridDim.x;
    if (next < ntiles) fetch(next);
    FftRun<S, T, Plan>::run(buf, ax, g, tw);
    const int z = tile / groups;
    const int q0 = (tile - z * groups) * W;
    const int wq = min(W, Q - q0);
    C2<T>* plane = spec + (size_t)z * My * Q + q0;
    for (int e = threadIdx.x; e < count; e += nthr) {
      const int w = e & (W - 1);
      if (w < wq) plane[(size_t)(e >> lw) * Q + w] = buf[e];
    }
    lds_sync();
    tile = next;
  }
}

// ------------------------------------------------------------------------------ z pass
// Tile: one ky, W consecutive kx columns with their three components (3W pencils along z,
// contiguous per kz). Forward transform, projection, inverse transform (mode bits 1, 2).
template <typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_z_project_kernel(C2<T>* __restrict__ spec, FftAxis ax, int My, int Mh,
                                                            int W, int lw, int tp, const C2<T>* __restrict__ tw_g,
                                                            const T* __restrict__ k_x, const T* __restrict__ k_y,
                                                            const T* __restrict__ k_z, T inv_mu_M, int mode) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Mz = ax.n;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + Mz;
  const int groups = (Mh + W - 1) / W;
  const int ntiles = My * groups;
  const int np = 3 * W;
  const int nthr = blockDim.x;
  const int count = Mz * np;
  const size_t zstride = (size_t)My * Mh * 3;
  load_twiddles(tw, tw_g, Mz);
  Geo g;
  g.np = np;
  g.tp = tp;
  g.my_p = threadIdx.x % np;
  g.my_t = threadIdx.x / np;
  g.ps = 1;
  g.es = np;
  C2<T> pre[kEPT];
  auto fetch = [&](int tile) {
    const int ky = tile / groups;
    const int kx0 = (tile - ky * groups) * W;
    const int nq = 3 * min(W, Mh - kx0);
    const C2<T>* col = spec + ((size_t)ky * Mh + kx0) * 3;
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
      const int e = u * nthr + (int)threadIdx.x;
      if (e < count) {
        const int kz = (e >> lw) / 3;
        const int q = e - kz * np;
        pre[u] = q < nq ? col[kz * zstride + q] : C2<T>{T(0), T(0)};
      }
    }
  };
  int tile = blockIdx.x;
  if (tile < ntiles) fetch(tile);
  while (tile < ntiles) {
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
      const int e = u * nthr + (int)threadIdx.x;
      if (e < count) buf[e] = pre[u];
    }
    lds_sync();
    const int next = tile + gridDim.x;
    if (next < ntiles) fetch(next);
    const int ky = tile / groups;
    const int kx0 = (tile - ky * groups) * W;
    const int wq = min(W, Mh - kx0);
    if (mode & 1) {
      FftRun<-1, T, Plan>::run(buf, ax, g, tw);
      const T kyv = k_y[ky];
      for (int e = threadIdx.x; e < Mz * W; e += nthr) {
        const int kz = e >> lw;
        const int w = e - (kz << lw);
        if (w >= wq) continue;
        const T kxv = k_x[kx0 + w];
        const T kzv = k_z[kz];
        C2<T>* f = buf + kz * np + 3 * w;
        const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
        if (k2 == T(0)) {
          f[0] = f[1] = f[2] = C2<T>{T(0), T(0)};
          continue;
        }
        const C2<T> f0 = f[0], f1 = f[1], f2 = f[2];
        const T inv_k2 = T(1) / k2;
        const C2<T> kf{kxv * f0.x + kyv * f1.x + kzv * f2.x, kxv * f0.y + kyv * f1.y + kzv * f2.y};
        const C2<T> cc{kf.x * inv_k2, kf.y * inv_k2};
        const T sc = inv_mu_M * inv_k2;
        f[0] = C2<T>{sc * (f0.x - kxv * cc.x), sc * (f0.y - kxv * cc.y)};
        f[1] = C2<T>{sc * (f1.x - kyv * cc.x), sc * (f1.y - kyv * cc.y)};
        f[2] = C2<T>{sc * (f2.x - kzv * cc.x), sc * (f2.y - kzv * cc.y)};
      }
      lds_sync();
    }
    if (mode & 2) FftRun<1, T, Plan>::run(buf, ax, g, tw);
    C2<T>* col = spec + ((size_t)ky * Mh + kx0) * 3;
    for (int e = threadIdx.x; e < count; e += nthr) {
      const int kz = (e >> lw) / 3;
      const int q = e - kz * np;
      if (q < 3 * wq) col[kz * zstride + q] = buf[e];
    }
    lds_sync();
    tile = next;
  }
}

// ------------------------------------------------------------------------------ x inverse
// Tile: one z plane, 2P consecutive y rows. Pencil per (row pair, component):
// z[x] = A[x] + i B[x] with A, B Hermitian-extended (imaginary parts of DC and Nyquist
// dropped); the inverse transform returns a + i b.
template <typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_x_inv_kernel(const C2<T>* __restrict__ spec, T* __restrict__ grid,
                                                        FftAxis ax, int My, int Mz, int P, int tp,
                                                        const C2<T>* __restrict__ tw_g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Mx = ax.n, Mh = Mx / 2 + 1;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + Mx;
  const int pairs = (My + 1) / 2;
  const int groups = (pairs + P - 1) / P;
  const int ntiles = Mz * groups;
  const int nthr = blockDim.x;
  const bool even = (Mx & 1) == 0;
  const int row = 3 * Mh;
  const int count = P * row;  // (A, B) pairs; = nthr Mh / tp <= (kEPT / 2 + 1) nthr
  constexpr int kAB = kEPT / 2 + 1;
  load_twiddles(tw, tw_g, Mx);
  Geo g;
  g.np = 3 * P;
  g.tp = tp;
  g.my_t = threadIdx.x % tp;
  g.my_p = threadIdx.x / tp;
  g.ps = Mx;
  g.es = 1;
  C4<T> pre[kAB];
  auto fetch = [&](int tile) {
    const int z = tile / groups;
    const int y0 = (tile - z * groups) * 2 * P;
    const C2<T>* base = spec + ((size_t)z * My + y0) * (size_t)row;
    int p = 0, i = threadIdx.x;
    walk(p, i, 0, row);
#pragma unroll
    for (int u = 0; u < kAB; ++u) {
      if (u * nthr + (int)threadIdx.x < count) {
        const int ya = y0 + 2 * p;
        pre[u].a = ya < My ? base[(size_t)(2 * p) * row + i] : C2<T>{T(0), T(0)};
        pre[u].b = ya + 1 < My ? base[(size_t)(2 * p + 1) * row + i] : C2<T>{T(0), T(0)};
      }
      walk(p, i, nthr, row);
    }
  };
  int tile = blockIdx.x;
  if (tile < ntiles) fetch(tile);
  while (tile < ntiles) {
    {
      int p = 0, i = threadIdx.x;
      walk(p, i, 0, row);
#pragma unroll
      for (int u = 0; u < kAB; ++u) {
        if (u * nthr + (int)threadIdx.x < count) {
          const int kx = i / 3;
          const int c = i - 3 * kx;
          C2<T> A = pre[u].a, B = pre[u].b;
          if (kx == 0 || (even && kx == Mx / 2)) {
            A.y = T(0);
            B.y = T(0);
          }
          C2<T>* pen = buf + (3 * p + c) * Mx;
          pen[kx] = C2<T>{A.x - B.y, A.y + B.x};
          if (kx >= 1 && Mx - kx >= Mh) pen[Mx - kx] = C2<T>{A.x + B.y, B.x - A.y};
        }
        walk(p, i, nthr, row);
      }
    }
    lds_sync();
    const int next = tile + gridDim.x;
    if (next < ntiles) fetch(next);
    FftRun<1, T, Plan>::run(buf, ax, g, tw);
    const int z = tile / groups;
    const int y0 = (tile - z * groups) * 2 * P;
    const int nrows = min(2 * P, My - y0);
    T* rows = grid + ((size_t)z * My + y0) * (size_t)Mx * 3;
    {
      int rr = 0, f = threadIdx.x;
      walk(rr, f, 0, 3 * Mx);
      for (int e = threadIdx.x; e < nrows * 3 * Mx; e += nthr) {
        const int x = f / 3;
        const int c = f - 3 * x;
        rows[e] = ((const T*)&buf[(3 * (rr >> 1) + c) * Mx + x])[rr & 1];
        walk(rr, f, nthr, 3 * Mx);
      }
    }
    lds_sync();
    tile = next;
  }
}

// Registers-per-stage bound: the smallest threads-per-pencil that keeps every stage within
// kElemMax values per thread.
int threads_per_pencil(const FftAxis& ax) {
  int tp = 1;
  for (;;) {
    bool ok = true;
    for (int s = 0; s < ax.nstages; ++s) {
      const int R = ax.radix[s];
      const int K = kElemMax / R > 0 ? kElemMax / R : 1;
      const int nb = ax.n / R;
      if ((nb + tp - 1) / tp > K) ok = false;
    }
    if (ok) return tp;
    ++tp;
  }
}


}  // namespace

bool plan_axis(int n, FftAxis* ax) {
  ax->n = n;
  ax->nstages = 0;
  int r = n;
  const int order[] = {8, 4, 2, 3, 5, 7, 11, 13};
  ax->packed = 0;
  for (int f : order) {
    while (r % f == 0 && ax->nstages < kMaxStages) {
      ax->packed |= (unsigned long long)f << (4 * ax->nstages);
      ax->radix[ax->nstages++] = f;
      r /= f;
    }
  }
  return r == 1;
}

namespace {
// Calls f(plan) with the compile-time radix plan of length n (Float32), else DynPlan.
template <typename T, typename F>
void with_plan(int n, F&& f) {
  if constexpr (sizeof(T) == 4) {
    switch (n) {
      case 64: f(Radices<8, 8>{}); return;
      case 128: f(Radices<8, 8, 2>{}); return;
      case 256: f(Radices<8, 8, 4>{}); return;
      case 512: f(Radices<8, 8, 8>{}); return;
      case 1024: f(Radices<8, 8, 8, 2>{}); return;
      default: break;
    }
  }
  f(DynPlan{});
}

int log2i(int w) {
  int l = 0;
  while ((1 << l) < w) ++l;
  return l;
}

// Persistent grid: as many blocks as can be resident at once (occupancy query), at most the
// number of tiles.
int resident_blocks(const void* kernel, int threads, int lds, int tiles, int cus) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  const long long b = (long long)per_cu * cus;
  return (int)(b < tiles ? b : tiles);
}
}  // namespace

template <typename T>
void plan_passes(ffcm_ctx* c, int lds_budget) {
  const int Mx = c->host.M[0], My = c->host.M[1], Mz = c->host.M[2];
  const int Mh = Mx / 2 + 1;
  const int cs = (int)sizeof(C2<T>);
  const int tpx = threads_per_pencil(c->fft[0]);
  const int tpy = threads_per_pencil(c->fft[1]);
  const int tpz = threads_per_pencil(c->fft[2]);
  // tiles of about 32-64 KiB: two or more resident blocks per CU
  const int lds_cap = lds_budget < 72 * 1024 ? lds_budget : 72 * 1024;
  int cus = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) cus = 256;
  (void)hipGetLastError();
  // x passes: P row pairs
  {
    const int pairs = (My + 1) / 2;
    int P = 256 / (3 * tpx);
    if (P < 1) P = 1;
    if (P > pairs) P = pairs;
    while (P > 1 && ((3 * P * Mx + Mx) * cs > lds_cap || 3 * P * tpx > 512)) --P;
    FftPass fp;
    fp.width = P;
    fp.log2w = 0;
    fp.pencils = 3 * P;
    fp.tp = tpx;
    fp.threads = 3 * P * tpx;
    fp.lds_bytes = (3 * P * Mx + Mx) * cs;
    fp.tiles = Mz * ((pairs + P - 1) / P);
    with_plan<T>(Mx, [&](auto plan) {
      using Pl = decltype(plan);
      fp.blocks = resident_blocks((const void*)fft_x_fwd_kernel<T, Pl>, fp.threads, fp.lds_bytes, fp.tiles, cus);
      c->pass_xf = fp;
      fp.blocks = resident_blocks((const void*)fft_x_inv_kernel<T, Pl>, fp.threads, fp.lds_bytes, fp.tiles, cus);
      c->pass_xi = fp;
    });
  }
  // y passes: W columns (power of two)
  {
    const int Q = 3 * Mh;
    int W = 16;
    while (W > 1 && ((W * My + My) * cs > lds_cap || W * tpy > 512)) W >>= 1;
    FftPass fp;
    fp.width = W;
    fp.log2w = log2i(W);
    fp.pencils = W;
    fp.tp = tpy;
    fp.threads = W * tpy;
    fp.lds_bytes = (W * My + My) * cs;
    fp.tiles = Mz * ((Q + W - 1) / W);
    with_plan<T>(My, [&](auto plan) {
      fp.blocks = resident_blocks((const void*)fft_y_kernel<-1, T, decltype(plan)>, fp.threads, fp.lds_bytes,
                                  fp.tiles, cus);
    });
    c->pass_y = fp;
  }
  // z pass: W kx columns x 3 components (W a power of two)
  {
    int W = 8;
    while (W > 1 && ((3 * W * Mz + Mz) * cs > lds_cap || 3 * W * tpz > 512)) W >>= 1;
    FftPass fp;
    fp.width = W;
    fp.log2w = log2i(W);
    fp.pencils = 3 * W;
    fp.tp = tpz;
    fp.threads = 3 * W * tpz;
    fp.lds_bytes = (3 * W * Mz + Mz) * cs;
    fp.tiles = My * ((Mh + W - 1) / W);
    with_plan<T>(Mz, [&](auto plan) {
      fp.blocks = resident_blocks((const void*)fft_z_project_kernel<T, decltype(plan)>, fp.threads, fp.lds_bytes,
                                  fp.tiles, cus);
    });
    c->pass_z = fp;
  }
}

namespace {
// mode bit 0: forward transforms + projection; bit 1: backward transforms.
template <typename T>
hipError_t run_fft(ffcm_ctx* c, int mode) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const int My = c->host.M[1];
  const int Mh = c->Mh;
  const FftPass& xf = c->pass_xf;
  const FftPass& py = c->pass_y;
  const FftPass& pz = c->pass_z;
  const FftPass& xi = c->pass_xi;
  C2<T>* spec = (C2<T>*)c->d.spectrum;
  const int Mz = c->host.M[2];
  const int Mx = c->host.M[0];
  if (mode & 1) {
    with_plan<T>(Mx, [&](auto plan) {
      fft_x_fwd_kernel<T, decltype(plan)><<<xf.blocks, xf.threads, xf.lds_bytes, c->stream>>>(
          (const T*)c->d.force, spec, c->fft[0], My, Mz, xf.width, xf.tp, (const C2<T>*)c->d.twiddle[0]);
    });
    with_plan<T>(My, [&](auto plan) {
      fft_y_kernel<-1, T, decltype(plan)><<<py.blocks, py.threads, py.lds_bytes, c->stream>>>(
          spec, c->fft[1], 3 * Mh, py.width, py.log2w, py.tp, Mz, (const C2<T>*)c->d.twiddle[1]);
    });
  }
  with_plan<T>(Mz, [&](auto plan) {
    fft_z_project_kernel<T, decltype(plan)><<<pz.blocks, pz.threads, pz.lds_bytes, c->stream>>>(
        spec, c->fft[2], My, Mh, pz.width, pz.log2w, pz.tp, (const C2<T>*)c->d.twiddle[2], (const T*)c->d.k_x,
        (const T*)c->d.k_y, (const T*)c->d.k_z, s.inv_mu_M, mode);
  });
  if (mode & 2) {
    with_plan<T>(My, [&](auto plan) {
      fft_y_kernel<1, T, decltype(plan)><<<py.blocks, py.threads, py.lds_bytes, c->stream>>>(
          spec, c->fft[1], 3 * Mh, py.width, py.log2w, py.tp, Mz, (const C2<T>*)c->d.twiddle[1]);
    });
    with_plan<T>(Mx, [&](auto plan) {
      fft_x_inv_kernel<T, decltype(plan)><<<xi.blocks, xi.threads, xi.lds_bytes, c->stream>>>(
          spec, (T*)c->d.velocity, c->fft[0], My, Mz, xi.width, xi.tp, (const C2<T>*)c->d.twiddle[0]);
    });
  }
  return hipGetLastError();
}
}  // namespace

// Full Stokes solve: force grid -> velocity grid.
template <typename T>
hipError_t launch_fft_stokes(ffcm_ctx* c) {
  return run_fft<T>(c, 3);
}

// Forward transforms + projection only; the spectrum is left projected, in the state the
// reference's backward transforms read (src/stokes_solve.jl:33-40).
template <typename T>
hipError_t launch_fft_forward_project(ffcm_ctx* c) {
  return run_fft<T>(c, 1);
}

// Backward transforms alone, from the spectrum buffer (src/stokes_solve.jl:38-40).
template <typename T>
hipError_t launch_fft_inverse(ffcm_ctx* c) {
  return run_fft<T>(c, 2);
}

template void plan_passes<float>(ffcm_ctx*, int);
template void plan_passes<double>(ffcm_ctx*, int);
template hipError_t launch_fft_stokes<float>(ffcm_ctx*);
template hipError_t launch_fft_stokes<double>(ffcm_ctx*);
template hipError_t launch_fft_forward_project<float>(ffcm_ctx*);
template hipError_t launch_fft_forward_project<double>(ffcm_ctx*);
template hipError_t launch_fft_inverse<float>(ffcm_ctx*);
template hipError_t launch_fft_inverse<double>(ffcm_ctx*);

}  // namespace ffcm
