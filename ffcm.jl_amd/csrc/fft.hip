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

constexpr int kElemMax = 32;  // complex values a thread holds per stage

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
  __syncthreads();
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
  __syncthreads();
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

// Copies `count` elements with U independent loads in flight per thread: all loads of a batch
// are issued before the first store, so a block keeps U * blockDim requests outstanding.
template <int U, typename V, typename Load, typename Store>
__device__ __forceinline__ void batched_copy(int count, Load ld, Store st) {
  for (int base = threadIdx.x; base < count; base += U * (int)blockDim.x) {
    V tmp[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = base + u * (int)blockDim.x;
      if (e < count) tmp[u] = ld(e);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e = base + u * (int)blockDim.x;
      if (e < count) st(e, tmp[u]);
    }
  }
}

template <typename V>
__device__ __forceinline__ void load_twiddles(V* tw, const V* __restrict__ g, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = g[i];
}

// ------------------------------------------------------------------------------ x forward
// Block: one z plane, 2P consecutive y rows; rows (2p, 2p+1) form one complex pencil per
// component: z[x] = a[x] + i b[x]. Then A = (Z[k] + conj Z[-k]) / 2, B = (Z[k] - conj Z[-k]) / 2i.
template <typename T>
__global__ void __launch_bounds__(512) fft_x_fwd_kernel(const T* __restrict__ grid, C2<T>* __restrict__ spec,
                                                         FftAxis ax, int My, int P, int tp,
                                                         const C2<T>* __restrict__ tw_g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Mx = ax.n, Mh = Mx / 2 + 1;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + Mx;
  const int pairs = (My + 1) / 2;
  const int groups = (pairs + P - 1) / P;
  const int z = blockIdx.x / groups;
  const int y0 = (blockIdx.x % groups) * 2 * P;
  const int nthr = blockDim.x;
  load_twiddles(tw, tw_g, Mx);
  {
    // rows y0 .. y0 + nrows - 1 are contiguous: 3 Mx values each
    const int nrows = min(2 * P, My - y0);
    const T* rows = grid + ((size_t)z * My + y0) * (size_t)Mx * 3;
    if ((3 * Mx) % VE<T> == 0) {
      using W4 = typename Vec16<T>::type;
      const int per_row = 3 * Mx / VE<T>;
      if (2 * P > nrows) {
        // zero the b half of the last pair when My is odd
        for (int e = threadIdx.x; e < 3 * Mx; e += nthr) {
          const int x = e / 3;
          const int c = e - 3 * x;
          ((T*)&buf[(3 * (nrows >> 1) + c) * Mx + x])[1] = T(0);
        }
      }
      batched_copy<8, W4>(
          nrows * per_row, [&](int e) { return ((const W4*)rows)[e]; },
          [&](int e, W4 v) {
            const int rr = e / per_row;
            const int f0 = (e - rr * per_row) * VE<T>;
            const int p3 = 3 * (rr >> 1);
            const int part = rr & 1;
            const T* vv = (const T*)&v;
#pragma unroll
            for (int j = 0; j < VE<T>; ++j) {
              const int f = f0 + j;
              const int x = f / 3;
              const int c = f - 3 * x;
              ((T*)&buf[(p3 + c) * Mx + x])[part] = vv[j];
            }
          });
    } else {
      batched_copy<16, T>(
          2 * P * 3 * Mx, [&](int e) { return e < nrows * 3 * Mx ? rows[e] : T(0); },
          [&](int e, T v) {
            const int rr = e / (3 * Mx);
            const int f = e - rr * 3 * Mx;
            const int x = f / 3;
            const int c = f - 3 * x;
            ((T*)&buf[(3 * (rr >> 1) + c) * Mx + x])[rr & 1] = v;
          });
    }
  }
  __syncthreads();
  Geo g;
  g.np = 3 * P;
  g.tp = tp;
  g.my_t = threadIdx.x % tp;
  g.my_p = threadIdx.x / tp;
  g.ps = Mx;
  g.es = 1;
  lds_fft<-1, T>(buf, ax, g, tw);
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
      C2<T> A{(Z.x + Zm.x) * T(0.5), (Z.y - Zm.y) * T(0.5)};
      C2<T> B{(Z.y + Zm.y) * T(0.5), (Zm.x - Z.x) * T(0.5)};
      ra[e] = A;
      if (yb < My) rb[e] = B;
    }
  }
}

// ------------------------------------------------------------------------------ y passes
// Block: one z plane, W consecutive columns of the interleaved (kx, component) row of
// length Q = 3 Mh; pencils run along y with stride Q. LDS layout [y][W].
template <int S, typename T>
__global__ void __launch_bounds__(512) fft_y_kernel(C2<T>* __restrict__ spec, FftAxis ax, int Q, int W, int tp,
                                                     const C2<T>* __restrict__ tw_g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int My = ax.n;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + My;
  const int groups = (Q + W - 1) / W;
  const int z = blockIdx.x / groups;
  const int q0 = (blockIdx.x % groups) * W;
  const int nthr = blockDim.x;
  load_twiddles(tw, tw_g, My);
  C2<T>* plane = spec + (size_t)z * My * Q + q0;
  const int wq = (Q - q0) < W ? (Q - q0) : W;
  batched_copy<16, C2<T>>(
      My * W,
      [&](int e) {
        const int y = e / W;
        const int w = e - y * W;
        return w < wq ? plane[(size_t)y * Q + w] : C2<T>{T(0), T(0)};
      },
      [&](int e, C2<T> v) { buf[e] = v; });
  __syncthreads();
  Geo g;
  g.np = W;
  g.tp = tp;
  g.my_p = threadIdx.x % W;
  g.my_t = threadIdx.x / W;
  g.ps = 1;
  g.es = W;
  lds_fft<S, T>(buf, ax, g, tw);
  batched_copy<16, C2<T>>(
      My * W, [&](int e) { return buf[e]; },
      [&](int e, C2<T> v) {
        const int y = e / W;
        const int w = e - y * W;
        if (w < wq) plane[(size_t)y * Q + w] = v;
      });
  (void)nthr;
}

// ------------------------------------------------------------------------------ z pass
// Block: one ky, W consecutive kx columns with their three components (3W pencils along z,
// contiguous per kz). Forward transform, projection, inverse transform.
template <typename T>
__global__ void __launch_bounds__(512) fft_z_project_kernel(C2<T>* __restrict__ spec, FftAxis ax, int My, int Mh,
                                                             int W, int tp, const C2<T>* __restrict__ tw_g,
                                                             const T* __restrict__ k_x, const T* __restrict__ k_y,
                                                             const T* __restrict__ k_z, T inv_mu_M,
                                                             int mode) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Mz = ax.n;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + Mz;
  const int groups = (Mh + W - 1) / W;
  const int ky = blockIdx.x / groups;
  const int kx0 = (blockIdx.x % groups) * W;
  const int wq = (Mh - kx0) < W ? (Mh - kx0) : W;
  const int np = 3 * W;
  const int nthr = blockDim.x;
  const size_t zstride = (size_t)My * Mh * 3;
  C2<T>* col = spec + ((size_t)ky * Mh + kx0) * 3;
  load_twiddles(tw, tw_g, Mz);
  batched_copy<16, C2<T>>(
      Mz * np,
      [&](int e) {
        const int kz = e / np;
        const int q = e - kz * np;
        return q < 3 * wq ? col[kz * zstride + q] : C2<T>{T(0), T(0)};
      },
      [&](int e, C2<T> v) { buf[e] = v; });
  __syncthreads();
  Geo g;
  g.np = np;
  g.tp = tp;
  g.my_p = threadIdx.x % np;
  g.my_t = threadIdx.x / np;
  g.ps = 1;
  g.es = np;
  if (mode & 1) lds_fft<-1, T>(buf, ax, g, tw);
  const T kyv = k_y[ky];
  for (int e = threadIdx.x; e < ((mode & 1) ? Mz * W : 0); e += nthr) {
    const int kz = e / W;
    const int w = e - kz * W;
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
    const T s = inv_mu_M * inv_k2;
    f[0] = C2<T>{s * (f0.x - kxv * cc.x), s * (f0.y - kxv * cc.y)};
    f[1] = C2<T>{s * (f1.x - kyv * cc.x), s * (f1.y - kyv * cc.y)};
    f[2] = C2<T>{s * (f2.x - kzv * cc.x), s * (f2.y - kzv * cc.y)};
  }
  __syncthreads();
  if (mode & 2) lds_fft<1, T>(buf, ax, g, tw);
  batched_copy<16, C2<T>>(
      Mz * np, [&](int e) { return buf[e]; },
      [&](int e, C2<T> v) {
        const int kz = e / np;
        const int q = e - kz * np;
        if (q < 3 * wq) col[kz * zstride + q] = v;
      });
}

// ------------------------------------------------------------------------------ x inverse
// Block: one z plane, 2P consecutive y rows. Pencil per (row pair, component):
// z[x] = A[x] + i B[x] with A, B Hermitian-extended (imaginary parts of DC and Nyquist
// dropped); the inverse transform returns a + i b.
template <typename T>
__global__ void __launch_bounds__(512) fft_x_inv_kernel(const C2<T>* __restrict__ spec, T* __restrict__ grid,
                                                         FftAxis ax, int My, int P, int tp,
                                                         const C2<T>* __restrict__ tw_g) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Mx = ax.n, Mh = Mx / 2 + 1;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + Mx;
  const int pairs = (My + 1) / 2;
  const int groups = (pairs + P - 1) / P;
  const int z = blockIdx.x / groups;
  const int y0 = (blockIdx.x % groups) * 2 * P;
  const int nthr = blockDim.x;
  const bool even = (Mx & 1) == 0;
  load_twiddles(tw, tw_g, Mx);
  {
    const int row = 3 * Mh;
    const C2<T>* base = spec + ((size_t)z * My + y0) * (size_t)row;
    batched_copy<8, C4<T>>(
        P * row,
        [&](int e) {
          const int p = e / row;
          const int i = e - p * row;
          const int ya = y0 + 2 * p;
          C4<T> ab;
          ab.a = ya < My ? base[(size_t)(2 * p) * row + i] : C2<T>{T(0), T(0)};
          ab.b = ya + 1 < My ? base[(size_t)(2 * p + 1) * row + i] : C2<T>{T(0), T(0)};
          return ab;
        },
        [&](int e, C4<T> ab) {
          const int p = e / row;
          const int i = e - p * row;
          const int kx = i / 3;
          const int c = i - 3 * kx;
          C2<T> A = ab.a, B = ab.b;
          if (kx == 0 || (even && kx == Mx / 2)) {
            A.y = T(0);
            B.y = T(0);
          }
          C2<T>* pen = buf + (3 * p + c) * Mx;
          pen[kx] = C2<T>{A.x - B.y, A.y + B.x};
          if (kx >= 1 && Mx - kx >= Mh) pen[Mx - kx] = C2<T>{A.x + B.y, B.x - A.y};
        });
  }
  __syncthreads();
  Geo g;
  g.np = 3 * P;
  g.tp = tp;
  g.my_t = threadIdx.x % tp;
  g.my_p = threadIdx.x / tp;
  g.ps = Mx;
  g.es = 1;
  lds_fft<1, T>(buf, ax, g, tw);
  {
    const int nrows = min(2 * P, My - y0);
    T* rows = grid + ((size_t)z * My + y0) * (size_t)Mx * 3;
    batched_copy<16, T>(
        nrows * 3 * Mx,
        [&](int e) {
          const int rr = e / (3 * Mx);
          const int f = e - rr * 3 * Mx;
          const int x = f / 3;
          const int c = f - 3 * x;
          return ((const T*)&buf[(3 * (rr >> 1) + c) * Mx + x])[rr & 1];
        },
        [&](int e, T v) { rows[e] = v; });
  }
  (void)nthr;
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

template <typename T>
void plan_passes(ffcm_ctx* c, int lds_budget) {
  const int Mx = c->host.M[0], My = c->host.M[1], Mz = c->host.M[2];
  const int Mh = Mx / 2 + 1;
  const int cs = (int)sizeof(C2<T>);
  const int tpx = threads_per_pencil(c->fft[0]);
  const int tpy = threads_per_pencil(c->fft[1]);
  const int tpz = threads_per_pencil(c->fft[2]);
  const int lds_cap = lds_budget;
  // x passes: P row pairs, ~256 threads
  {
    const int pairs = (My + 1) / 2;
    int P = 256 / (3 * tpx);
    if (P < 1) P = 1;
    if (P > pairs) P = pairs;
    while (P > 1 && (3 * P * Mx + Mx) * cs > lds_cap) --P;
    FftPass fp;
    fp.width = P;
    fp.pencils = 3 * P;
    fp.tp = tpx;
    fp.threads = 3 * P * tpx;
    fp.lds_bytes = (3 * P * Mx + Mx) * cs;
    fp.blocks = Mz * ((pairs + P - 1) / P);
    c->pass_xf = fp;
    c->pass_xi = fp;
  }
  // y passes: W columns (power of two)
  {
    const int Q = 3 * Mh;
    int W = 16;
    while (W > 1 && ((W * My + My) * cs > lds_cap || W * tpy > 1024)) W >>= 1;
    FftPass fp;
    fp.width = W;
    fp.pencils = W;
    fp.tp = tpy;
    fp.threads = W * tpy;
    fp.lds_bytes = (W * My + My) * cs;
    fp.blocks = Mz * ((Q + W - 1) / W);
    c->pass_y = fp;
  }
  // z pass: W kx columns x 3 components
  {
    int W = 8;
    while (W > 1 && ((3 * W * Mz + Mz) * cs > lds_cap || 3 * W * tpz > 1024)) W >>= 1;
    FftPass fp;
    fp.width = W;
    fp.pencils = 3 * W;
    fp.tp = tpz;
    fp.threads = 3 * W * tpz;
    fp.lds_bytes = (3 * W * Mz + Mz) * cs;
    fp.blocks = My * ((Mh + W - 1) / W);
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
  if (mode & 1) {
    fft_x_fwd_kernel<T><<<xf.blocks, xf.threads, xf.lds_bytes, c->stream>>>(
        (const T*)c->d.force, spec, c->fft[0], My, xf.width, xf.tp, (const C2<T>*)c->d.twiddle[0]);
    fft_y_kernel<-1, T><<<py.blocks, py.threads, py.lds_bytes, c->stream>>>(spec, c->fft[1], 3 * Mh, py.width,
                                                                           py.tp, (const C2<T>*)c->d.twiddle[1]);
  }
  fft_z_project_kernel<T><<<pz.blocks, pz.threads, pz.lds_bytes, c->stream>>>(
      spec, c->fft[2], My, Mh, pz.width, pz.tp, (const C2<T>*)c->d.twiddle[2], (const T*)c->d.k_x,
      (const T*)c->d.k_y, (const T*)c->d.k_z, s.inv_mu_M, mode);
  if (mode & 2) {
    fft_y_kernel<1, T><<<py.blocks, py.threads, py.lds_bytes, c->stream>>>(spec, c->fft[1], 3 * Mh, py.width, py.tp,
                                                                          (const C2<T>*)c->d.twiddle[1]);
    fft_x_inv_kernel<T><<<xi.blocks, xi.threads, xi.lds_bytes, c->stream>>>(
        spec, (T*)c->d.velocity, c->fft[0], My, xi.width, xi.tp, (const C2<T>*)c->d.twiddle[0]);
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
