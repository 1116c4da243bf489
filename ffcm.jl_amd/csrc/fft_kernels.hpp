// Device side of the hand-written 3-D real FFT (see fft.hip for the structure and the
// reference conventions): complex arithmetic, in-register DFTs, Stockham stages, the LDS
// pencil layout, the four persistent pass kernels and the kernel tables.
#pragma once
#include <algorithm>
#include <cmath>
#include <type_traits>

#include "ffcm_internal.hpp"
#include "group.hpp"
#include "stencil.hpp"

namespace ffcm {

namespace fftk {

constexpr int kElemMax = 8;  // complex values a thread holds per stage (one radix-8 butterfly)

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

// Waits until every vector memory operation of this wave has completed (s_waitcnt vmcnt(0);
// expcnt and lgkmcnt left at their maxima). Placed once before the prefetch registers are
// consumed: the compiler then knows no load into them is pending and inserts no wait of its
// own, neither at each conditional use nor before the next fetch overwrites them.
__device__ __forceinline__ void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Wave-level LDS ordering: a wave's LDS operations complete in issue order, so within one
// wave only compiler reordering has to be prevented.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// LDS holds the block's pencils pencil-major with padding: element t of pencil p is at
// buf[p * PS + t + t / 32], with PS = n + n / 32 rounded up to an odd number. Stage reads
// (consecutive t), stage writes (stride R) and the coalesced staging loops (consecutive p)
// then fall on distinct banks.
// The compile-time plans of lengths that are not powers of two use the unpadded layout
// (element t at p * PS + t, PS = n rounded up to an odd number): their first stage has an odd
// radix (see the radix order of those plans), so every stage's reads and writes are
// bank-conflict-free without padding, and each element of a butterfly is a compile-time offset
// from one base index.
template <bool Padded>
__host__ __device__ constexpr int lds_i(int t) {
  return Padded ? t + (t >> 5) : t;
}
__host__ __device__ constexpr int lds_t(int t) { return lds_i<true>(t); }
__host__ __device__ inline int pencil_stride(int n, bool padded = true) {
  const int ps = padded ? n + (n >> 5) : n;
  return ps | 1;
}

// Pencil geometry of one block. Threads are pencil-major (my_p = tid / tp): when tp divides
// 64 every wave owns whole pencils and the stages synchronise per wave.
struct Geo {
  int np, tp;      // pencils of the tile, threads per pencil
  int ng;          // pencils transformed at once (threads / tp); rounds cover np
  int my_p, my_t;  // this thread's first pencil and its lane within the pencil
  int ps;          // pencil stride (pencil_stride(n))
  bool wave_local;
};

__device__ __forceinline__ void stage_sync(bool wave_local) {
  if (wave_local)
    wave_sync();
  else
    lds_sync();
}

// One Stockham stage (Ns = product of the earlier radices): butterfly j reads
// x[j + r n/R] * exp(S 2 pi i r (j mod Ns) / (Ns R)), applies the size-R DFT and writes
// x[(j - j mod Ns) R + j mod Ns + r Ns]. Reads and writes are separated by a barrier.
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
      for (int r = 0; r < R; ++r) v[k][r] = base[lds_t(j + r * nb)];
      if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) v[k][r] = cmul(v[k][r], twid<S>(tw, r * jm * tstep));
      }
      Dft<R, S, T>::run(v[k], tw, n);
    }
  }
  stage_sync(g.wave_local);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = g.my_t + k * g.tp;
    if (active && j < nb) {
      const int jm = j % Ns;
      const int o = (j - jm) * R + jm;
#pragma unroll
      for (int r = 0; r < R; ++r) base[lds_t(o + r * Ns)] = v[k][r];
    }
  }
  stage_sync(g.wave_local);
}

// Full transform of every pencil of the block, in place in LDS.
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

template <int N>
struct PowerOfTwo {
  static constexpr bool value = N > 0 && (N & (N - 1)) == 0;
};

// Padded offsets of one stage that do not depend on the butterfly: lds_t(j + r nb) - lds_t(j)
// = lds_t(r nb) (reads) and lds_t(o + r Ns) - lds_t(o) = lds_t(r Ns) with
// o = (j - j mod Ns) R + j mod Ns (writes), for every butterfly j < N / R. Holds for power-of-two plans; then each element is a constant
// offset from one base index (immediate offsets of the LDS instructions).
template <int R, int N, int Ns, bool P>
struct StageOffsets {
  static constexpr bool reads_const() {
    constexpr int nb = N / R;
    for (int j = 0; j < nb; ++j)
      for (int r = 1; r < R; ++r)
        if (lds_i<P>(j + r * nb) - lds_i<P>(j) != lds_i<P>(r * nb)) return false;
    return true;
  }
  static constexpr bool writes_const() {
    constexpr int nb = N / R;
    for (int j = 0; j < nb; ++j) {
      const int jm = j % Ns;
      const int o = (j - jm) * R + jm;
      for (int r = 1; r < R; ++r)
        if (lds_i<P>(o + r * Ns) - lds_i<P>(o) != lds_i<P>(r * Ns)) return false;
    }
    return true;
  }
  static constexpr bool kReads = reads_const();
  static constexpr bool kWrites = writes_const();
};

// Complex values a thread of a compile-time plan may hold per stage: every stage's butterflies
// are spread over TP threads per pencil, ceil((N / R) / TP) butterflies each.
constexpr int kElemStatic = 12;

// Complex values of the next tile each thread of the x passes prefetches (the planner sizes
// tiles to fit; the y and z passes of some plans take more, ept_strided_of).
constexpr int kPrefetchSlots = 8;
// Threads of the wide z-pass kernel (a separate instance, only for the longest Float32 plan).
constexpr int kZWideThreads = 768;

// Threads per pencil of a compile-time plan: the smallest power of two (at most 64, so that
// pencils never straddle a wave and the stages synchronise per wave) with which no stage
// holds more than kElemStatic values per thread (one butterfly if R alone exceeds it) and a
// pencil's N values fit its threads' prefetch slots. Compared with one radix-8 butterfly per
// thread, the radix-3 and radix-5 stages of the 5-smooth lengths take two or three butterflies
// per thread, which halves their threads per pencil (216: 64 -> 32).
template <int N, int... Rs>
constexpr int plan_tp() {
  for (int p = 1; p < 64; p *= 2) {
    bool ok = N <= p * kPrefetchSlots;
    for (int R : {Rs...}) {
      const int K = (N / R + p - 1) / p;
      if (K > 1 && K * R > kElemStatic) ok = false;
    }
    if (ok) return p;
  }
  return 64;
}

// One Stockham stage of a compile-time plan with TP threads per pencil, for IL pencils per
// thread at once (the pencils of IL consecutive rounds, my_p + p0 + i * ng): their LDS reads
// are in flight together and each twiddle is read once for all of them. Pencils past the tile
// (p >= np) are masked.
template <int R, int S, typename T, int N, int Ns, int IL, int TP>
__device__ __forceinline__ void stage_ct_il(C2<T>* buf, const Geo& g, const C2<T>* tw, int p0) {
  constexpr int nb = N / R;
  constexpr int K = (nb + TP - 1) / TP;
  constexpr bool P = PowerOfTwo<N>::value;
  constexpr int tstep = N / (Ns * R);
  using Off = StageOffsets<R, N, Ns, P>;
  C2<T> v[IL][K][R];
  bool act[IL];
#pragma unroll
  for (int i = 0; i < IL; ++i) act[i] = g.my_p + p0 + i * g.ng < g.np;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = g.my_t + k * TP;
    if (j < nb) {
      const int jm = j % Ns;
#pragma unroll
      for (int i = 0; i < IL; ++i) {
        const C2<T>* base = buf + (g.my_p + p0 + i * g.ng) * g.ps;
        if (act[i]) {
          if constexpr (Off::kReads) {
            const C2<T>* b0 = base + lds_i<P>(j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][k][r] = b0[lds_i<P>(r * nb)];
          } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][k][r] = base[lds_i<P>(j + r * nb)];
          }
        }
      }
      if constexpr (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const C2<T> w = twid<S>(tw, r * jm * tstep);
#pragma unroll
          for (int i = 0; i < IL; ++i) v[i][k][r] = cmul(v[i][k][r], w);
        }
      }
#pragma unroll
      for (int i = 0; i < IL; ++i) Dft<R, S, T>::run(v[i][k], tw, N);
    }
  }
  wave_sync();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int j = g.my_t + k * TP;
    if (j < nb) {
      const int jm = j % Ns;
      const int o = (j - jm) * R + jm;
#pragma unroll
      for (int i = 0; i < IL; ++i) {
        C2<T>* base = buf + (g.my_p + p0 + i * g.ng) * g.ps;
        if (act[i]) {
          if constexpr (Off::kWrites) {
            C2<T>* b0 = base + lds_i<P>(o);
#pragma unroll
            for (int r = 0; r < R; ++r) b0[lds_i<P>(r * Ns)] = v[i][k][r];
          } else {
#pragma unroll
            for (int r = 0; r < R; ++r) base[lds_i<P>(o + r * Ns)] = v[i][k][r];
          }
        }
      }
    }
  }
  wave_sync();
}

template <int S, typename T, int N, int Ns, int IL, int TP, int R, int... Rest>
__device__ __forceinline__ void stages_ct_il(C2<T>* buf, const Geo& g, const C2<T>* tw, int p0) {
  stage_ct_il<R, S, T, N, Ns, IL, TP>(buf, g, tw, p0);
  if constexpr (sizeof...(Rest) > 0) stages_ct_il<S, T, N, Ns * R, IL, TP, Rest...>(buf, g, tw, p0);
}

// Pencils (rounds) interleaved per thread by the compile-time plans (2 measured no faster
// than 1 once the prefetch depth was cut to eight values).
constexpr int kInterleave = 1;

template <int S, typename T, typename Plan>
struct FftRun;
template <int S, typename T>
struct FftRun<S, T, DynPlan> {
  __device__ static void run(C2<T>* buf, const FftAxis& ax, const Geo& g, const C2<T>* tw) {
    Geo r = g;
    for (int p0 = 0; p0 < g.np; p0 += g.ng) {  // every thread runs every round (block barriers)
      r.my_p = g.my_p + p0;
      lds_fft<S, T>(buf, ax, r, tw);
    }
  }
};
// Static plans run with plan_tp threads per pencil (a power of two dividing 64; the host
// launches them so), so their stages sync per wave.
template <int S, typename T, int... Rs>
struct FftRun<S, T, Radices<Rs...>> {
  __device__ __forceinline__ static void run(C2<T>* buf, const FftAxis&, const Geo& g, const C2<T>* tw) {
    constexpr int N = (Rs * ...);
    for (int p0 = 0; p0 < g.np; p0 += kInterleave * g.ng)
      stages_ct_il<S, T, N, 1, kInterleave, plan_tp<N, Rs...>(), Rs...>(buf, g, tw, p0);
  }
};

// First and last radix of a compile-time plan (0 for the runtime plan).
template <typename Plan>
struct PlanEnds {
  static constexpr int first = 0, last = 0;
};
template <int R1, int... Rs>
struct PlanEnds<Radices<R1, Rs...>> {
  static constexpr int first = R1;
  static constexpr int last = [] {
    int r = R1;
    ((r = Rs), ...);
    return r;
  }();
};
// Does a thread that holds E values of a pencil run the plan's first and last stages in
// registers? They need whole butterflies (E a multiple of the radix); the strided passes do so
// where E holds at least two, the 16-value prefetch of the longest Float32 plans (the kernels
// of the other plans would carry the extra registers without being measured faster).
template <typename Plan, int E>
constexpr bool plan_ends_fit() {
  constexpr int f = PlanEnds<Plan>::first, l = PlanEnds<Plan>::last;
  return f > 0 && E % f == 0 && E % l == 0 && E >= 2 * f && E >= 2 * l;
}

template <int S, typename T, int N, int TP, int R1, int... Rest>
__device__ __forceinline__ void stages_after_first(C2<T>* buf, const Geo& g, const C2<T>* tw, int p0) {
  if constexpr (sizeof...(Rest) > 0) stages_ct_il<S, T, N, R1, kInterleave, TP, Rest...>(buf, g, tw, p0);
}
template <int S, typename T, int N, int Ns, int TP, int R, int... Rest>
__device__ __forceinline__ void stages_but_last(C2<T>* buf, const Geo& g, const C2<T>* tw, int p0) {
  if constexpr (sizeof...(Rest) > 0) {
    stage_ct_il<R, S, T, N, Ns, kInterleave, TP>(buf, g, tw, p0);
    stages_but_last<S, T, N, Ns * R, TP, Rest...>(buf, g, tw, p0);
  }
}
// The LDS stages of a compile-time plan whose first (after_first) or last (but_last) stage the
// caller runs in registers.
template <int S, typename T, typename Plan>
struct FftMiddle;
template <int S, typename T, int... Rs>
struct FftMiddle<S, T, Radices<Rs...>> {
  static constexpr int N = (Rs * ...);
  __device__ __forceinline__ static void after_first(C2<T>* buf, const Geo& g, const C2<T>* tw) {
    for (int p0 = 0; p0 < g.np; p0 += kInterleave * g.ng)
      stages_after_first<S, T, N, plan_tp<N, Rs...>(), Rs...>(buf, g, tw, p0);
  }
  __device__ __forceinline__ static void but_last(C2<T>* buf, const Geo& g, const C2<T>* tw) {
    for (int p0 = 0; p0 < g.np; p0 += kInterleave * g.ng)
      stages_but_last<S, T, N, 1, plan_tp<N, Rs...>(), Rs...>(buf, g, tw, p0);
  }
};

// First stage (sign S) of pencil `pen` from the values t = t0 + rpu u (u < E) a thread holds,
// when rpu E = N: butterfly j = t0 + rpu k takes u = k + (E / R) r and writes t = j R + r.
template <int S, typename T, typename Plan, bool Pad, int E>
__device__ __forceinline__ void first_stage_regs(const C2<T> (&in)[E], C2<T>* pen, int t0, int rpu,
                                                 const C2<T>* tw, int N) {
  constexpr int R = PlanEnds<Plan>::first, K = E / R;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    C2<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[k + K * r];
    Dft<R, S, T>::run(v, tw, N);
    const int j = t0 + rpu * k;
#pragma unroll
    for (int r = 0; r < R; ++r) pen[lds_i<Pad>(j * R + r)] = v[r];
  }
}
template <typename V>
__device__ __forceinline__ void load_twiddles(V* tw, const V* __restrict__ g, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) tw[i] = g[i];
}

struct DynPlan;
// All four pass kernels are persistent: block b transforms tiles b, b + gridDim, ... While
// tile t is transformed in LDS, the loads of tile t + gridDim are already in flight into
// registers (kEPT values per thread), so HBM latency overlaps the arithmetic.
// Eight prefetch slots keep the compile-time plans at about 70-125 VGPRs, four or more waves
// per SIMD: measured on 256^3, the FFT stage took 0.50 ms with 8, 0.63 ms with 16 and 0.67 ms
// with 32 (more prefetch, larger tiles, two waves per SIMD); at 512^3 8 and 16 were equal.
template <typename T, typename Plan>
constexpr int ept_of() {
  return kPrefetchSlots;
}

// Transform length of a compile-time plan (0 for the runtime plan).
template <typename Plan>
struct PlanLength {
  static constexpr int value = 0;
};
template <int... Rs>
struct PlanLength<Radices<Rs...>> {
  static constexpr int value = (Rs * ...);
};

// LDS layout of a plan: padded for the runtime plan and the power-of-two plans.
template <typename Plan>
struct PlanPadded {
  static constexpr bool value = true;
};
template <int... Rs>
struct PlanPadded<Radices<Rs...>> {
  static constexpr bool value = PowerOfTwo<(Rs * ...)>::value;
};

// The strided y and z passes of the longest Float32 plans (512) prefetch 16 values: with 8
// their column groups shrink to 64- and 48-byte row segments, which fetch partial lines. (For
// the 5-smooth lengths 96..240 at 16 right-hand sides, 16 slots made the y passes slower and
// the z pass no faster.)
template <typename T, typename Plan>
constexpr int ept_strided_of() {
  return (sizeof(T) == 4 && PlanLength<Plan>::value >= 512) ? 16 : kPrefetchSlots;
}

// Advances (row, col) of a row-major walk with row length `len` by `step` elements.
__device__ __forceinline__ void walk(int& row, int& col, int step, int len) {
  col += step;
  while (col >= len) {
    col -= len;
    ++row;
  }
}

// Arguments of the pass kernels.
template <typename T>
struct FftXArgs {  // x passes: real grid (read by the forward pass, written by the inverse) and spectrum
  T* grid;
  C2<T>* spec;
  FftAxis ax;
  int My, Mz, P, tp, Mhp;
  const C2<T>* tw;
};
template <typename T>
struct FftYArgs {
  C2<T>* spec;
  FftAxis ax;
  int Q, Qs, W, lw, tp, Mz;
  const C2<T>* tw;
};
template <typename T>
struct FftZArgs {
  C2<T>* spec;
  FftAxis ax;
  int My, Mh, Mhp, W, lw, tp;
  const C2<T>* tw;
  const T* k_x;
  const T* k_y;
  const T* k_z;
  T inv_mu_M;
  int mode, nb;
};

// ------------------------------------------------------------------------------ x forward
// Tile: one z plane, 2P consecutive y rows; rows (2p, 2p+1) form one complex pencil per
// component: z[x] = a[x] + i b[x]. Then A = (Z[k] + conj Z[-k]) / 2, B = (Z[k] - conj Z[-k]) / 2i.
// Each row is read in U1 = ceil(3 Mx / nthr) chunks of nthr values; iteration u of the fetch
// covers pair u / U1, chunk u % U1 (both wave-uniform), and holds the same element of both
// rows of the pair, so one 8-byte LDS store places it.
template <typename T, typename Plan>
__device__ __forceinline__ void fft_x_fwd_body(const Blk& blk, const FftXArgs<T>& a) {
  const T* __restrict__ grid = a.grid;
  C2<T>* __restrict__ spec = a.spec;
  const FftAxis& ax = a.ax;
  const int My = a.My, Mz = a.Mz, P = a.P, tp = a.tp, Mhp = a.Mhp;
  const C2<T>* __restrict__ tw_g = a.tw;
  constexpr int kEPT = ept_of<T, Plan>();
  constexpr bool kPad = PlanPadded<Plan>::value;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Mx = ax.n, Mh = Mx / 2 + 1;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + Mx;
  const int pairs = (My + 1) / 2;
  const int groups = (pairs + P - 1) / P;
  const int ntiles = Mz * groups;
  const int nthr = blockDim.x;
  const int rowlen = 3 * Mx;
  const int U1 = (rowlen + nthr - 1) / nthr;
  const int nu = P * U1;  // planner: <= kEPT
  load_twiddles(tw, tw_g, Mx);
  Geo g;
  g.np = 3 * P;
  g.tp = tp;
  g.ng = nthr / tp;
  g.my_t = threadIdx.x % tp;
  g.my_p = threadIdx.x / tp;
  g.ps = pencil_stride(Mx, kPad);
  g.wave_local = (64 % tp) == 0;
  const int PS = g.ps;
  C2<T> pre[kEPT];  // (row 2p value, row 2p+1 value)
  const unsigned voff = threadIdx.x;
  // The prefetch registers are written only by the loads themselves (addresses clamped into
  // the tile, out-of-range values masked when they are consumed): a select or zero fill of a
  // register with a load still in flight would make the compiler wait for all outstanding
  // memory operations before every load.
  // Slots [u0, u1) of a tile's fetch; tiles past the end read the last tile (loads issued
  // unconditionally, the last ones never used).
  auto fetch = [&](int tile, int u0, int u1) {
    tile = min(tile, ntiles - 1);
    const int z = tile / groups;
    const int y0 = (tile - z * groups) * 2 * P;
    const T* rows = grid + ((size_t)z * My + y0) * (size_t)rowlen;
    const int last = My - 1 - y0;
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
      if (u < nu && u >= u0 && u < u1) {
        const int p = u / U1;
        const int f0 = (u - p * U1) * nthr;  // wave-uniform chunk start, < rowlen
        const unsigned off = min(voff, (unsigned)(rowlen - 1 - f0));
        pre[u].x = (rows + ((size_t)min(2 * p, last) * rowlen + f0))[off];
        pre[u].y = (rows + ((size_t)min(2 * p + 1, last) * rowlen + f0))[off];
      }
    }
  };
  int tile = xcd_first_tile(blk.id, blk.n);
  if (tile < ntiles) fetch(tile, 0, kEPT);
  while (tile < ntiles) {
    const int y0n = (tile % groups) * 2 * P;
    vm_wait_all();
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
      if (u < nu) {
        const int p = u / U1;
        const int f = (u - p * U1) * nthr + (int)threadIdx.x;
        if (f < rowlen) {
          const int x = f / 3;
          const int c = f - 3 * x;
          C2<T> v = pre[u];
          if (y0n + 2 * p + 1 >= My) v.y = T(0);  // odd My: the last pair has one row
          buf[(3 * p + c) * PS + lds_i<kPad>(x)] = v;
        }
      }
    }
    lds_sync();
    const int next = tile + blk.n;
    // three quarters of the next tile's fetch before the transform, the rest after it, in
    // flight during the stores (512^3: 0.741 -> 0.729 ms; halves 0.737)
    fetch(next, 0, kEPT * 3 / 4);
    FftRun<-1, T, Plan>::run(buf, ax, g, tw);
    fetch(next, kEPT * 3 / 4, kEPT);
    lds_sync();
    const int z = tile / groups;
    const int y0 = (tile - z * groups) * 2 * P;
    for (int p = 0; p < P; ++p) {
      const int ya = y0 + 2 * p, yb = ya + 1;
      if (ya >= My) break;
      C2<T>* ra = spec + ((size_t)z * My + ya) * (size_t)Mhp * 3;
      C2<T>* rb = spec + ((size_t)z * My + yb) * (size_t)Mhp * 3;
      for (int e = threadIdx.x; e < 3 * Mhp; e += nthr) {
        const int kx = e / 3;
        const int c = e - 3 * kx;
        if (kx >= Mh) {  // row padding: written (as zeros) so every 128-byte line is whole
          ra[e] = C2<T>{T(0), T(0)};
          if (yb < My) rb[e] = C2<T>{T(0), T(0)};
          continue;
        }
        const C2<T>* pen = buf + (3 * p + c) * PS;
        const C2<T> Z = pen[lds_i<kPad>(kx)];
        const C2<T> Zm = pen[lds_i<kPad>(kx == 0 ? 0 : Mx - kx)];
        // conj(Zm) = (Zm.x, -Zm.y)
        ra[e] = C2<T>{(Z.x + Zm.x) * T(0.5), (Z.y - Zm.y) * T(0.5)};
        if (yb < My) rb[e] = C2<T>{(Z.y + Zm.y) * T(0.5), (Zm.x - Z.x) * T(0.5)};
      }
    }
    lds_sync();
    tile = next;
  }
  vm_wait_all();  // the unused loads issued during the last tile
}

// ------------------------------------------------------------------------------ y passes
// Tile: one z plane, W consecutive columns of the interleaved (kx, component) row of length
// Q = 3 Mh; pencils run along y with stride Q. LDS layout [y][W].
template <int S, typename T, typename Plan>
__device__ __forceinline__ void fft_y_body(const Blk& blk, const FftYArgs<T>& a) {
  C2<T>* __restrict__ spec = a.spec;
  const FftAxis& ax = a.ax;
  const int Q = a.Q, Qs = a.Qs, W = a.W, lw = a.lw, tp = a.tp, Mz = a.Mz;
  const C2<T>* __restrict__ tw_g = a.tw;
  constexpr int kEPT = ept_strided_of<T, Plan>();
  constexpr bool kPad = PlanPadded<Plan>::value;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int My = ax.n;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + My;
  const int groups = (Q + W - 1) / W;
  const int ntiles = Mz * groups;
  const int nthr = blockDim.x;
  load_twiddles(tw, tw_g, My);
  Geo g;
  g.np = W;
  g.tp = tp;
  g.ng = nthr / tp;
  g.my_t = threadIdx.x % tp;
  g.my_p = threadIdx.x / tp;
  g.ps = pencil_stride(My, kPad);
  g.wave_local = (64 % tp) == 0;
  const int PS = g.ps;
  C2<T> pre[kEPT];
  // e = u * nthr + tid covers row u * rpu + ty, column w (nthr is a multiple of W), so each
  // load is a wave-uniform base plus one per-thread 32-bit offset
  const int rpu = nthr >> lw;
  const int ty = threadIdx.x >> lw;
  const int wcol = threadIdx.x & (W - 1);
  // loads only write the prefetch registers: rows and columns past the end are clamped
  // (those pencils are transformed but never stored)
  auto fetch = [&](int tile) {
    const int z = tile / groups;
    const int q0 = (tile - z * groups) * W;
    const unsigned wc = (unsigned)min(wcol, Q - q0 - 1);
    const C2<T>* plane = spec + (size_t)z * My * Qs + q0;
    // one base and a 32-bit element offset per lane, advanced by rpu rows per slot; rows past
    // the end read the last row
    const unsigned last = (unsigned)(My - 1) * (unsigned)Qs + wc;
    const unsigned step = (unsigned)rpu * (unsigned)Qs;
    unsigned off = (unsigned)ty * (unsigned)Qs + wc;
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
      pre[u] = plane[u * rpu + ty < My ? off : last];
      off += step;
    }
  };
  // A tile's results wait in registers (the fetch's mapping: thread (ty, wcol) holds rows
  // ty + u rpu of column wcol) and are stored in the next iteration, before its fetch: the
  // wait for the loads at the top of an iteration then finds the stores long done, instead of
  // waiting for stores issued just before it.
  C2<T> out[kEPT];
  auto store = [&](int t) {
    const int z = t / groups;
    const int q0 = (t - z * groups) * W;
    if (wcol < min(W, Q - q0)) {
      C2<T>* dst = spec + (size_t)z * My * Qs + q0 + (size_t)ty * Qs + wcol;
      const size_t step = (size_t)rpu * Qs;
#pragma unroll
      for (int u = 0; u < kEPT; ++u)
        if (u * rpu + ty < My) dst[u * step] = out[u];
    }
  };
  // When rpu kEPT = My, the fetched values are whole butterflies of the first stage, which then
  // runs in registers (as in the z pass). Running the last stage into the held results as well
  // measured slower at 512^3 (703 / 690 us per pass against 687 / 677 us for the first alone and
  // 691 / 682 us for neither).
  constexpr bool kEnds = plan_ends_fit<Plan, kEPT>();
  bool in_regs = false;
  if constexpr (kEnds) in_regs = rpu * kEPT == My;
  int tile = xcd_first_tile(blk.id, blk.n);
  int held = -1;  // tile whose results are in out
  if (tile < ntiles) fetch(tile);
  while (tile < ntiles) {
    vm_wait_all();
    if constexpr (kEnds) {
      if (in_regs) first_stage_regs<S, T, Plan, kPad>(pre, buf + wcol * PS, ty, rpu, tw, My);
    }
    if (!in_regs) {
#pragma unroll
      for (int u = 0; u < kEPT; ++u) {
        const int y = u * rpu + ty;
        if (y < My) buf[wcol * PS + lds_i<kPad>(y)] = pre[u];
      }
    }
    lds_sync();
    if (held >= 0) store(held);
    const int next = tile + blk.n;
    if (next < ntiles) fetch(next);
    if constexpr (kEnds) {
      if (in_regs)
        FftMiddle<S, T, Plan>::after_first(buf, g, tw);
      else
        FftRun<S, T, Plan>::run(buf, ax, g, tw);
    } else {
      FftRun<S, T, Plan>::run(buf, ax, g, tw);
    }
    lds_sync();
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
      const int y = u * rpu + ty;
      if (y < My) out[u] = buf[wcol * PS + lds_i<kPad>(y)];
    }
    held = tile;
    lds_sync();
    tile = next;
  }
  if (held >= 0) store(held);
}

// ------------------------------------------------------------------------------ z pass
// Complex slots of the z pass's LDS that hold k_x (Mh values) and k_z (Mz values).
template <typename T>
__host__ __device__ inline int z_pass_k_slots(int Mh, int Mz) {
  const int cs = 2 * (int)sizeof(T);
  return ((Mh + Mz) * (int)sizeof(T) + cs - 1) / cs;
}

// ------------------------------------------------------------------------------ z pass
// Tile: one ky, W consecutive kx columns with their three components (3W pencils along z,
// contiguous per kz). Forward transform (mode bit 1), projection (bit 8), inverse transform
// (bit 2), in that order.
template <typename T, typename Plan>
__device__ __forceinline__ void fft_z_project_body(const Blk& blk, const FftZArgs<T>& a) {
  C2<T>* __restrict__ spec = a.spec;
  const FftAxis& ax = a.ax;
  const int My = a.My, Mh = a.Mh, Mhp = a.Mhp, W = a.W, lw = a.lw, tp = a.tp;
  const C2<T>* __restrict__ tw_g = a.tw;
  const T* __restrict__ k_x = a.k_x;
  const T* __restrict__ k_y = a.k_y;
  const T* __restrict__ k_z = a.k_z;
  const T inv_mu_M = a.inv_mu_M;
  const int mode = a.mode, nb = a.nb;
  constexpr int kEPT = ept_strided_of<T, Plan>();
  constexpr bool kPad = PlanPadded<Plan>::value;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Mz = ax.n;
  C2<T>* tw = (C2<T>*)smem_raw;
  // wavenumbers in LDS: global loads inside the projection would wait behind the prefetch
  T* kx_s = (T*)(tw + Mz);
  T* kz_s = kx_s + Mh;
  C2<T>* buf = tw + Mz + z_pass_k_slots<T>(Mh, Mz);
  const int groups = (Mh + W - 1) / W;
  const int per_grid = My * groups;  // tiles of one grid; nb grids follow each other
  const int ntiles = per_grid * nb;
  const int np = 3 * W;
  const int nthr = blockDim.x;
  const size_t zstride = (size_t)My * Mhp * 3;
  load_twiddles(tw, tw_g, Mz);
  Geo g;
  g.np = np;
  g.tp = tp;
  g.ng = nthr / tp;
  g.my_t = threadIdx.x % tp;
  g.my_p = threadIdx.x / tp;
  g.ps = pencil_stride(Mz, kPad);
  g.wave_local = (64 % tp) == 0;
  const int PS = g.ps;
  C2<T> pre[kEPT];
  // nthr is a multiple of np: e = u * nthr + tid covers kz = u * rpu + tz, pencil q
  const int rpu = nthr / np;
  const int tz = threadIdx.x / np;
  const int tq = threadIdx.x - tz * np;
  // loads only write the prefetch registers: planes and columns past the end are clamped
  // (those pencils are transformed but never stored), and so are tiles past the end (the
  // loads after the last tile are issued unconditionally and never used)
  // Slots [u0, u1) of a tile's fetch.
  auto fetch = [&](int tile, int u0, int u1) {
    tile = min(tile, ntiles - 1);
    const int b = tile / per_grid;
    const int t = tile - b * per_grid;
    const int ky = t / groups;
    const int kx0 = (t - ky * groups) * W;
    const unsigned q = (unsigned)min(tq, 3 * min(W, Mh - kx0) - 1);
    const C2<T>* col = spec + (size_t)b * Mz * zstride + ((size_t)ky * Mhp + kx0) * 3;
    // one base and a 32-bit element offset per lane (the planner checks it fits); rows past
    // the end read the last row
    const unsigned zs = (unsigned)zstride;
    const unsigned last = (unsigned)(Mz - 1) * zs + q;
#pragma unroll
    for (int u = 0; u < kEPT; ++u) {
      if (u >= u0 && u < u1) {
        const unsigned off = (unsigned)(tz + u * rpu) * zs + q;
        pre[u] = col[u * rpu + tz < Mz ? off : last];
      }
    }
  };
  // The next tile's fetch is spread over the tile's phases, a part after the staging, the
  // forward transform, the projection and the inverse transform's LDS stages, so that loads
  // stay in flight through the whole tile rather than in one burst (512^3: 0.990 ms with the
  // whole fetch after the projection, 0.98 after the staging, 0.87 spread in parts of 5, 4, 4
  // and 3 of 16 slots; 0.88 with 4 x 4 or 8, 4, 4).
  constexpr int kF1 = kEPT * 5 / 16, kF2 = kEPT * 9 / 16, kF3 = kEPT * 13 / 16;
  // A thread fetches the values kz = tz + rpu u (u < kEPT) of pencil tq. When rpu kEPT = Mz, these
  // are whole butterflies of the first stage (j = tz + rpu k, inputs j + (Mz / R) r at
  // u = k + (kEPT / R) r) and of the last (the same values, which it stores): the forward
  // transform's first stage then runs on the fetched registers and the inverse's last stage
  // feeds the stores, each saving a round trip of the tile through LDS.
  constexpr bool kEnds = plan_ends_fit<Plan, kEPT>();
  bool first_in_regs = false, last_in_regs = false;
  if constexpr (kEnds) {
    first_in_regs = (mode & 1) && rpu * kEPT == Mz;
    last_in_regs = (mode & 2) && rpu * kEPT == Mz;
  }
  // (made visible by the barrier that follows the first tile's staging)
  for (int i = threadIdx.x; i < Mh + Mz; i += blockDim.x) kx_s[i] = i < Mh ? k_x[i] : k_z[i - Mh];
  int tile = xcd_first_tile(blk.id, blk.n);
  if (tile < ntiles) fetch(tile, 0, kEPT);
  while (tile < ntiles) {
    vm_wait_all();
    if constexpr (kEnds) {
      if (first_in_regs) first_stage_regs<-1, T, Plan, kPad>(pre, buf + tq * PS, tz, rpu, tw, Mz);
    }
    if (!first_in_regs) {
#pragma unroll
      for (int u = 0; u < kEPT; ++u) {
        const int kz = u * rpu + tz;
        if (kz < Mz) buf[tq * PS + lds_i<kPad>(kz)] = pre[u];
      }
    }
    lds_sync();
    const int next = tile + blk.n;
    const int bt = tile / per_grid;
    const int ty = tile - bt * per_grid;
    const int ky = ty / groups;
    const int kx0 = (ty - ky * groups) * W;
    const int wq = min(W, Mh - kx0);
    // loaded before the fetch, so that waiting for it does not wait for the fetch
    const T kyv = k_y[ky];
    fetch(next, 0, kF1);
    if (mode & 1) {
      if constexpr (kEnds) {
        if (first_in_regs)
          FftMiddle<-1, T, Plan>::after_first(buf, g, tw);
        else
          FftRun<-1, T, Plan>::run(buf, ax, g, tw);
      } else {
        FftRun<-1, T, Plan>::run(buf, ax, g, tw);
      }
      lds_sync();
    }
    fetch(next, kF1, kF2);
    if (mode & 8) {
      // nthr is a multiple of 3W: thread t serves column w = t mod W at kz = t / W, t / W + nthr / W,
      // ... (the column's wavenumber, pencils and validity hoisted out of the kz loop)
      const int w = threadIdx.x & (W - 1);
      if (w < wq) {
        const T kxv = kx_s[kx0 + w];
        C2<T>* col = buf + 3 * w * PS;
#pragma unroll 2
        for (int kz = threadIdx.x >> lw; kz < Mz; kz += nthr >> lw) {
          const T kzv = kz_s[kz];
          C2<T>* f = col + lds_i<kPad>(kz);
          const T k2 = kxv * kxv + kyv * kyv + kzv * kzv;
          if (k2 == T(0)) {
            f[0] = f[PS] = f[2 * PS] = C2<T>{T(0), T(0)};
            continue;
          }
          const C2<T> f0 = f[0], f1 = f[PS], f2 = f[2 * PS];
          // Float32: the hardware reciprocal (within 1 ulp) in place of the correctly rounded
          // division sequence
          T inv_k2;
          if constexpr (sizeof(T) == 4)
            inv_k2 = __builtin_amdgcn_rcpf(k2);
          else
            inv_k2 = T(1) / k2;
          const C2<T> kf{kxv * f0.x + kyv * f1.x + kzv * f2.x, kxv * f0.y + kyv * f1.y + kzv * f2.y};
          const C2<T> cc{kf.x * inv_k2, kf.y * inv_k2};
          const T sc = inv_mu_M * inv_k2;
          f[0] = C2<T>{sc * (f0.x - kxv * cc.x), sc * (f0.y - kxv * cc.y)};
          f[PS] = C2<T>{sc * (f1.x - kyv * cc.x), sc * (f1.y - kyv * cc.y)};
          f[2 * PS] = C2<T>{sc * (f2.x - kzv * cc.x), sc * (f2.y - kzv * cc.y)};
        }
      }
      lds_sync();
    }
    fetch(next, kF2, kF3);
    if (mode & 2) {
      if constexpr (kEnds) {
        if (last_in_regs)
          FftMiddle<1, T, Plan>::but_last(buf, g, tw);
        else
          FftRun<1, T, Plan>::run(buf, ax, g, tw);
      } else {
        FftRun<1, T, Plan>::run(buf, ax, g, tw);
      }
      lds_sync();
    }
    fetch(next, kF3, kEPT);
    // the fetch's mapping: thread (tz, tq) stores kz = u * rpu + tz of pencil tq (stored at once:
    // holding the results in registers until the next tile's fetch, as the y pass does, made
    // the 4-column pass slower, 1.14 -> 1.65 ms at 512^3, and the 8-column one no faster)
    C2<T>* col = spec + (size_t)bt * Mz * zstride + ((size_t)ky * Mhp + kx0) * 3;
    if constexpr (kEnds) {
      if (last_in_regs && tq < 3 * wq) {
        // Stockham last stage (Ns = Mz / R): butterfly j reads and writes j + (Mz / R) r, with
        // twiddles exp(2 pi i r j / Mz)
        constexpr int R = PlanEnds<Plan>::last, K = kEPT / R;
        const int nb = Mz / R;
        const C2<T>* src = buf + tq * PS;
#pragma unroll
        for (int k = 0; k < K; ++k) {
          const int j = tz + rpu * k;
          C2<T> v[R];
#pragma unroll
          for (int r = 0; r < R; ++r) v[r] = src[lds_i<kPad>(j + nb * r)];
#pragma unroll
          for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<1>(tw, r * j));
          Dft<R, 1, T>::run(v, tw, Mz);
          C2<T>* dst = col + (size_t)j * zstride + tq;
#pragma unroll
          for (int r = 0; r < R; ++r) dst[(size_t)(nb * r) * zstride] = v[r];
        }
      }
    }
    if (!last_in_regs && tq < 3 * wq) {
      C2<T>* dst = col + (size_t)tz * zstride + tq;
      const size_t step = (size_t)rpu * zstride;
#pragma unroll 1
      for (int kz = tz; kz < Mz; kz += rpu, dst += step) *dst = buf[tq * PS + lds_i<kPad>(kz)];
    }
    lds_sync();
    tile = next;
  }
  vm_wait_all();  // the unused loads issued during the last tile
}

// ------------------------------------------------------------------------------ x inverse
// Tile: one z plane, 2P consecutive y rows. Pencil per (row pair, component):
// z[x] = A[x] + i B[x] with A, B Hermitian-extended (imaginary parts of DC and Nyquist
// dropped); the inverse transform returns a + i b. The half-spectrum rows are read in
// U1 = ceil(3 Mh / nthr) chunks per row pair (pair and chunk wave-uniform per iteration).
template <typename T, typename Plan>
__device__ __forceinline__ void fft_x_inv_body(const Blk& blk, const FftXArgs<T>& a) {
  const C2<T>* __restrict__ spec = a.spec;
  T* __restrict__ grid = a.grid;
  const FftAxis& ax = a.ax;
  const int My = a.My, Mz = a.Mz, P = a.P, tp = a.tp, Mhp = a.Mhp;
  const C2<T>* __restrict__ tw_g = a.tw;
  constexpr int kEPT = ept_of<T, Plan>();
  constexpr bool kPad = PlanPadded<Plan>::value;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int Mx = ax.n, Mh = Mx / 2 + 1;
  C2<T>* tw = (C2<T>*)smem_raw;
  C2<T>* buf = tw + Mx;
  const int pairs = (My + 1) / 2;
  const int groups = (pairs + P - 1) / P;
  const int ntiles = Mz * groups;
  const int nthr = blockDim.x;
  const bool even = (Mx & 1) == 0;
  const int row = 3 * Mh;     // valid entries of a half-spectrum row
  const int rstr = 3 * Mhp;   // row stride
  const int U1 = (row + nthr - 1) / nthr;
  const int nu = P * U1;  // planner: <= kAB
  constexpr int kAB = kEPT / 2 + 1;
  load_twiddles(tw, tw_g, Mx);
  Geo g;
  g.np = 3 * P;
  g.tp = tp;
  g.ng = nthr / tp;
  g.my_t = threadIdx.x % tp;
  g.my_p = threadIdx.x / tp;
  g.ps = pencil_stride(Mx, kPad);
  g.wave_local = (64 % tp) == 0;
  const int PS = g.ps;
  C4<T> pre[kAB];
  const unsigned voff = threadIdx.x;
  // loads only write the prefetch registers (clamped addresses; masking at consumption)
  // (slots [u0, u1) of a tile, tiles past the end reading the last one, as in the forward pass)
  auto fetch = [&](int tile, int u0, int u1) {
    tile = min(tile, ntiles - 1);
    const int z = tile / groups;
    const int y0 = (tile - z * groups) * 2 * P;
    const C2<T>* base = spec + ((size_t)z * My + y0) * (size_t)rstr;
    const int last = My - 1 - y0;
#pragma unroll
    for (int u = 0; u < kAB; ++u) {
      if (u < nu && u >= u0 && u < u1) {
        const int p = u / U1;
        const int i0 = (u - p * U1) * nthr;  // wave-uniform chunk start, < row
        const unsigned off = min(voff, (unsigned)(row - 1 - i0));
        pre[u].a = (base + ((size_t)min(2 * p, last) * rstr + i0))[off];
        pre[u].b = (base + ((size_t)min(2 * p + 1, last) * rstr + i0))[off];
      }
    }
  };
  int tile = xcd_first_tile(blk.id, blk.n);
  if (tile < ntiles) fetch(tile, 0, kAB);
  while (tile < ntiles) {
    const int y0n = (tile % groups) * 2 * P;
    vm_wait_all();
#pragma unroll
    for (int u = 0; u < kAB; ++u) {
      if (u < nu) {
        const int p = u / U1;
        const int i = (u - p * U1) * nthr + (int)threadIdx.x;
        if (i < row) {
          const int kx = i / 3;
          const int c = i - 3 * kx;
          C2<T> A = pre[u].a, B = pre[u].b;
          if (y0n + 2 * p + 1 >= My) B = C2<T>{T(0), T(0)};  // odd My: the last pair has one row
          if (kx == 0 || (even && kx == Mx / 2)) {
            A.y = T(0);
            B.y = T(0);
          }
          C2<T>* pen = buf + (3 * p + c) * PS;
          pen[lds_i<kPad>(kx)] = C2<T>{A.x - B.y, A.y + B.x};
          if (kx >= 1 && Mx - kx >= Mh) pen[lds_i<kPad>(Mx - kx)] = C2<T>{A.x + B.y, B.x - A.y};
        }
      }
    }
    lds_sync();
    const int next = tile + blk.n;
    // half of the next tile's fetch before the transform, half after it (512^3: 0.725 ->
    // 0.708 ms; three quarters before: 0.712)
    fetch(next, 0, kAB / 2);
    FftRun<1, T, Plan>::run(buf, ax, g, tw);
    fetch(next, kAB / 2, kAB);
    lds_sync();
    const int z = tile / groups;
    const int y0 = (tile - z * groups) * 2 * P;
    const int rowlen = 3 * Mx;
    T* rows = grid + ((size_t)z * My + y0) * (size_t)rowlen;
    for (int p = 0; p < P; ++p) {
      const int ya = y0 + 2 * p;
      if (ya >= My) break;
      const bool hasb = ya + 1 < My;
      T* ra = rows + (size_t)(2 * p) * rowlen;
      for (int f = threadIdx.x; f < rowlen; f += nthr) {
        const int x = f / 3;
        const int c = f - 3 * x;
        const C2<T> v = buf[(3 * p + c) * PS + lds_i<kPad>(x)];
        ra[f] = v.x;
        if (hasb) ra[rowlen + f] = v.y;
      }
    }
    lds_sync();
    tile = next;
  }
  vm_wait_all();  // the unused loads issued during the last tile
}

// Entry points of the four pass kernels: plain and grouped (many systems in one launch).
template <typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_x_fwd_kernel(FftXArgs<T> a) {
  fft_x_fwd_body<T, Plan>(plain_blk(), a);
}
template <typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_x_fwd_kernel_group(const GroupEntry<FftXArgs<T>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const FftXArgs<T>& a) { fft_x_fwd_body<T, Plan>(b, a); });
}
template <int S, typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_y_kernel(FftYArgs<T> a) {
  fft_y_body<S, T, Plan>(plain_blk(), a);
}
template <int S, typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_y_kernel_group(const GroupEntry<FftYArgs<T>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const FftYArgs<T>& a) { fft_y_body<S, T, Plan>(b, a); });
}
template <typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_z_project_kernel(FftZArgs<T> a) {
  fft_z_project_body<T, Plan>(plain_blk(), a);
}
template <typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_z_project_kernel_group(const GroupEntry<FftZArgs<T>>* __restrict__ g,
                                                                  int n) {
  group_run(g, n, [](const Blk& b, const FftZArgs<T>& a) { fft_z_project_body<T, Plan>(b, a); });
}
// The z pass of the longest Float32 plan with up to kZWideThreads threads: eight kx columns
// (24 pencils) per tile. Its own instances, so that the launch bound of every other kernel
// stays 512.
template <typename T, typename Plan>
__global__ void __launch_bounds__(kZWideThreads) fft_z_project_wide_kernel(FftZArgs<T> a) {
  fft_z_project_body<T, Plan>(plain_blk(), a);
}
template <typename T, typename Plan>
__global__ void __launch_bounds__(kZWideThreads)
    fft_z_project_wide_kernel_group(const GroupEntry<FftZArgs<T>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const FftZArgs<T>& a) { fft_z_project_body<T, Plan>(b, a); });
}
template <typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_x_inv_kernel(FftXArgs<T> a) {
  fft_x_inv_body<T, Plan>(plain_blk(), a);
}
template <typename T, typename Plan>
__global__ void __launch_bounds__(512) fft_x_inv_kernel_group(const GroupEntry<FftXArgs<T>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const FftXArgs<T>& a) { fft_x_inv_body<T, Plan>(b, a); });
}

// Registers-per-stage bound: the smallest threads-per-pencil that keeps every stage within
// kElemMax values per thread.
inline int threads_per_pencil(const FftAxis& ax) {
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


}  // namespace fftk
}  // namespace ffcm

namespace ffcm {
namespace fftk {

// The five pass kernels of one radix plan, launched through these pointers: the compile-time
// plans are instantiated in separate translation units (fft_plans_*.hip) and looked up by
// length at context creation.
template <typename T>
struct FftKernels {
  void (*x_fwd)(FftXArgs<T>);
  void (*y_fwd)(FftYArgs<T>);
  void (*y_inv)(FftYArgs<T>);
  void (*z)(FftZArgs<T>);
  void (*x_inv)(FftXArgs<T>);
  void (*x_fwd_g)(const GroupEntry<FftXArgs<T>>*, int);
  void (*y_fwd_g)(const GroupEntry<FftYArgs<T>>*, int);
  void (*y_inv_g)(const GroupEntry<FftYArgs<T>>*, int);
  void (*z_g)(const GroupEntry<FftZArgs<T>>*, int);
  void (*x_inv_g)(const GroupEntry<FftXArgs<T>>*, int);
  void (*z_wide)(FftZArgs<T>);  // wide z pass (make_kernels_wide), else null
  void (*z_wide_g)(const GroupEntry<FftZArgs<T>>*, int);
  int ept;          // prefetch slots per thread of the x passes (ept_of)
  int ept_strided;  // and of the y and z passes (ept_strided_of)
  bool is_static;   // compile-time plan
  int tp;           // threads per pencil of a compile-time plan (plan_tp), 0 for the runtime plan
  bool padded;      // LDS layout (PlanPadded)
};

template <typename Plan>
struct PlanTp {
  static constexpr int value = 0;
};
template <int... Rs>
struct PlanTp<Radices<Rs...>> {
  static constexpr int value = plan_tp<(Rs * ...), Rs...>();
};

template <typename T, typename Plan>
FftKernels<T> make_kernels() {
  FftKernels<T> k;
  k.x_fwd = fft_x_fwd_kernel<T, Plan>;
  k.y_fwd = fft_y_kernel<-1, T, Plan>;
  k.y_inv = fft_y_kernel<1, T, Plan>;
  k.z = fft_z_project_kernel<T, Plan>;
  k.x_inv = fft_x_inv_kernel<T, Plan>;
  k.x_fwd_g = fft_x_fwd_kernel_group<T, Plan>;
  k.y_fwd_g = fft_y_kernel_group<-1, T, Plan>;
  k.y_inv_g = fft_y_kernel_group<1, T, Plan>;
  k.z_g = fft_z_project_kernel_group<T, Plan>;
  k.x_inv_g = fft_x_inv_kernel_group<T, Plan>;
  k.z_wide = nullptr;
  k.z_wide_g = nullptr;
  k.ept = ept_of<T, Plan>();
  k.ept_strided = ept_strided_of<T, Plan>();
  k.is_static = !std::is_same_v<Plan, DynPlan>;
  k.tp = PlanTp<Plan>::value;
  k.padded = PlanPadded<Plan>::value;
  return k;
}

// make_kernels plus the wide z-pass kernels.
template <typename T, typename Plan>
FftKernels<T> make_kernels_wide() {
  FftKernels<T> k = make_kernels<T, Plan>();
  k.z_wide = fft_z_project_wide_kernel<T, Plan>;
  k.z_wide_g = fft_z_project_wide_kernel_group<T, Plan>;
  return k;
}

// Compile-time Float32 plans, one lookup per translation unit; false when n has none there.
bool static_kernels_part0(int n, FftKernels<float>* out);
bool static_kernels_part1(int n, FftKernels<float>* out);
bool static_kernels_part2(int n, FftKernels<float>* out);
bool static_kernels_part3(int n, FftKernels<float>* out);

}  // namespace fftk
}  // namespace ffcm
