// Step 6 of the mobility apply: the self term and the real-space pair correction.
//
// Reference: correct_velocities! / _correct_velocities_kernel! CPU method
// (src/correct_velocities.jl:154-286), _self_correction (:21-27), _correction_scalars
// (:71-104), _min_image (:121-127). For every slot: v = delta * F_s, then the 27 periodic
// neighbour cells in (dz, dy, dx) order, slots in range order, skipping s itself, adding
// A(r) F_m + B(r) (x . F_m) x for every pair with r^2 < R_c^2; finally V[:, n] += v.
//
// Design: LPP consecutive lanes per sorted slot (1 or 4) gather over its 27 cells, lane q taking
// cells q, q + LPP, ... of the reference's CPU visit order; their partial sums are added with
// row-level moves (no atomics, deterministic). Four lanes keep more of the walk's dependent
// loads in flight than one (one wave per SIMD at 65,536 particles). Consecutive threads hold
// slots of the same cell and walk the same neighbour ranges, so the candidate loads are shared
// through the cache.
#include <cstdlib>

#include "group.hpp"
#include "stencil.hpp"
#include "pair_terms.hpp"

namespace ffcm {

namespace {

template <typename T>
struct CorrectArgs {
  const typename Vec<T>::v4* Ys;
  const typename Vec<T>::v4* Fs;
  const int* cell_start;
  const int* cell_end;
  const int* perm;
  T* V;
  Scalars<T> s;
  const typename Vec<T>::v4* tab;  // Float32: the pair-scalar table (pair_scalars_table), or null
  T tab_scale;                     // kPairTableIntervals / R_c^2
};

// The pair scalars of one pair: from the table when the context has one, else the formulas.
template <typename T>
__device__ __forceinline__ void pair_AB(const CorrectArgs<T>& a, const PairConsts<T>& pk, T rr, T* A, T* B) {
  if constexpr (sizeof(T) == 4) {
    if (a.tab) {
      pair_scalars_table(rr, a.tab, a.tab_scale, A, B);
      return;
    }
  }
  correction_scalars(tsqrt(rr), pk, A, B);
}

template <typename T, int LPP>
__device__ __forceinline__ void correct_body(const Blk& blk, const CorrectArgs<T>& a) {
  const typename Vec<T>::v4* __restrict__ Ys = a.Ys;
  const typename Vec<T>::v4* __restrict__ Fs = a.Fs;
  const int* __restrict__ cell_start = a.cell_start;
  const int* __restrict__ cell_end = a.cell_end;
  const int* __restrict__ perm = a.perm;
  T* __restrict__ V = a.V;
  const Scalars<T>& s = a.s;
  __shared__ int list[kPairBatch * 256];
  const PairConsts<T> pk = pair_consts(s);
  const int total = s.N * LPP;
  for (int t = blk.id * blockDim.x + threadIdx.x; t < total; t += blk.n * blockDim.x) {
    const int slot = t / LPP, q = t % LPP;
    const typename Vec<T>::v4 yn = Ys[slot];
    const typename Vec<T>::v4 fn = Fs[slot];
    const T d = q == 0 ? s.delta : T(0);
    T v0 = d * fn.x, v1 = d * fn.y, v2 = d * fn.z;
    for_each_pair<T>(
        slot, yn, s, cell_start, cell_end, Ys, list,
        [&](int m, T x0, T x1, T x2, T rr) {
          T A, B;
          pair_AB(a, pk, rr, &A, &B);
          const typename Vec<T>::v4 fm = Fs[m];
          const T bx = B * (x0 * fm.x + x1 * fm.y + x2 * fm.z);
          v0 += A * fm.x + bx * x0;
          v1 += A * fm.y + bx * x1;
          v2 += A * fm.z + bx * x2;
        },
        q, LPP);
    v0 = lane_group_sum<LPP>(v0);
    v1 = lane_group_sum<LPP>(v1);
    v2 = lane_group_sum<LPP>(v2);
    if (q == 0) {
      const int n = perm[slot];
      V[3 * n + 0] += v0;
      V[3 * n + 1] += v1;
      V[3 * n + 2] += v2;
    }
  }
}

// RC right-hand sides at once: the pair scalars A(r), B(r) (the erf/exp work) are evaluated
// once per pair and applied to every right-hand side, with the lane split and operation order
// of correct_kernel<T, LPP>, so each right-hand side's result is bit-identical to a single-RHS
// launch. Forces of right-hand side r are at Fs[r * N + slot]; velocities at V[r * 3N + 3n + c].
template <typename T, int RC, int LPP>
__device__ __forceinline__ void correct_multi_body(const Blk& blk, const CorrectArgs<T>& a) {
  const typename Vec<T>::v4* __restrict__ Ys = a.Ys;
  const typename Vec<T>::v4* __restrict__ Fs = a.Fs;
  const int* __restrict__ cell_start = a.cell_start;
  const int* __restrict__ cell_end = a.cell_end;
  const int* __restrict__ perm = a.perm;
  T* __restrict__ V = a.V;
  const Scalars<T>& s = a.s;
  const size_t N = (size_t)s.N;
  __shared__ int list[kPairBatch * 256];
  const PairConsts<T> pk = pair_consts(s);
  const int total = s.N * LPP;
  for (int t = blk.id * blockDim.x + threadIdx.x; t < total; t += blk.n * blockDim.x) {
    const int slot = t / LPP, q = t % LPP;
    const typename Vec<T>::v4 yn = Ys[slot];
    const T d = q == 0 ? s.delta : T(0);
    T v[RC][3];
#pragma unroll
    for (int r = 0; r < RC; ++r) {
      const typename Vec<T>::v4 fn = Fs[r * N + slot];
      v[r][0] = d * fn.x;
      v[r][1] = d * fn.y;
      v[r][2] = d * fn.z;
    }
    for_each_pair<T>(
        slot, yn, s, cell_start, cell_end, Ys, list,
        [&](int s2, T x0, T x1, T x2, T rr) {
          T A, B;
          pair_AB(a, pk, rr, &A, &B);
#pragma unroll
          for (int r = 0; r < RC; ++r) {
            const typename Vec<T>::v4 fm = Fs[r * N + s2];
            const T bx = B * (x0 * fm.x + x1 * fm.y + x2 * fm.z);
            v[r][0] += A * fm.x + bx * x0;
            v[r][1] += A * fm.y + bx * x1;
            v[r][2] += A * fm.z + bx * x2;
          }
        },
        q, LPP);
#pragma unroll
    for (int r = 0; r < RC; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) v[r][c] = lane_group_sum<LPP>(v[r][c]);
    if (q == 0) {
      const int n = perm[slot];
#pragma unroll
      for (int r = 0; r < RC; ++r) {
        T* Vr = V + r * 3 * N;
        Vr[3 * n + 0] += v[r][0];
        Vr[3 * n + 1] += v[r][1];
        Vr[3 * n + 2] += v[r][2];
      }
    }
  }
}

template <typename T, int LPP>
__global__ void __launch_bounds__(256) correct_kernel(CorrectArgs<T> a) {
  correct_body<T, LPP>(plain_blk(), a);
}
template <typename T, int LPP>
__global__ void __launch_bounds__(256) correct_kernel_group(const GroupEntry<CorrectArgs<T>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const CorrectArgs<T>& a) { correct_body<T, LPP>(b, a); });
}
template <typename T, int RC, int LPP>
__global__ void __launch_bounds__(256) correct_multi_kernel(CorrectArgs<T> a) {
  correct_multi_body<T, RC, LPP>(plain_blk(), a);
}
template <typename T, int RC, int LPP>
__global__ void __launch_bounds__(256) correct_multi_kernel_group(const GroupEntry<CorrectArgs<T>>* __restrict__ g,
                                                                  int n) {
  group_run(g, n, [](const Blk& b, const CorrectArgs<T>& a) { correct_multi_body<T, RC, LPP>(b, a); });
}

// Lanes per slot: four (measured against one: 0.039 / 0.065 ms at 65,536 slots, where one
// thread per slot leaves one wave per SIMD; 0.49 / 0.56 ms at 2^20; against eight: 0.033 / 0.036
// ms and 0.41 / 0.47 ms). FFCM_CORRECT_LPP=1 or 8 forces one or eight.
int lanes_per_slot(int64_t) {
  if (const char* e = std::getenv("FFCM_CORRECT_LPP")) {
    const int v = std::atoi(e);
    if (v == 1 || v == 4 || v == 8) return v;
  }
  return 4;
}

unsigned correct_blocks(int64_t threads) {
  int64_t blocks = (threads + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > (1 << 20)) blocks = 1 << 20;
  return (unsigned)blocks;
}


// R right-hand sides (slot-ordered forces Fs[r * N + slot], velocities V[r * 3N + ...]) in
// chunks of 16, 8, 4, 2, 1.
template <typename T, int LPP>
hipError_t correct_multi_launch(ffcm_ctx* c, T* V, const void* Fs, int R) {
  using V4 = typename Vec<T>::v4;
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const unsigned nb = correct_blocks(c->N * LPP);
  for (int r0 = 0; r0 < R;) {
    const int left = R - r0;
    const int rc = left >= 16 ? 16 : left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
    const V4* F0 = (const V4*)Fs + (size_t)r0 * c->N;
    T* V0 = V + (size_t)r0 * 3 * c->N;
    const CorrectArgs<T> a{(const V4*)c->d.Y_sorted, F0, c->d.cell_start, c->d.cell_end, c->d.perm, V0, s,
                           (const V4*)c->d.pair_table, T(kPairTableIntervals) / s.Rc2};
    const int tag = kTagCorrect + r0;
    hipError_t e;
    switch (rc) {
      case 16: e = submit(c, tag, correct_multi_kernel<T, 16, LPP>, correct_multi_kernel_group<T, 16, LPP>, nb, 256, 0, a); break;
      case 8: e = submit(c, tag, correct_multi_kernel<T, 8, LPP>, correct_multi_kernel_group<T, 8, LPP>, nb, 256, 0, a); break;
      case 4: e = submit(c, tag, correct_multi_kernel<T, 4, LPP>, correct_multi_kernel_group<T, 4, LPP>, nb, 256, 0, a); break;
      case 2: e = submit(c, tag, correct_multi_kernel<T, 2, LPP>, correct_multi_kernel_group<T, 2, LPP>, nb, 256, 0, a); break;
      default: e = submit(c, tag, correct_kernel<T, LPP>, correct_kernel_group<T, LPP>, nb, 256, 0, a); break;
    }
    if (e != hipSuccess) return e;
    r0 += rc;
  }
  return hipSuccess;
}

}  // namespace

template <typename T>
hipError_t launch_correct_multi(ffcm_ctx* c, T* V, const void* Fs, int R) {
  const int lpp = lanes_per_slot(c->N);
  return lpp == 8 ? correct_multi_launch<T, 8>(c, V, Fs, R)
                  : (lpp == 4 ? correct_multi_launch<T, 4>(c, V, Fs, R) : correct_multi_launch<T, 1>(c, V, Fs, R));
}

template <typename T>
hipError_t launch_correct(ffcm_ctx* c, T* V, const void* Fs) {
  if (!Fs) Fs = c->d.F_sorted;
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  using V4 = typename Vec<T>::v4;
  const int lpp = lanes_per_slot(c->N);
  const CorrectArgs<T> a{(const V4*)c->d.Y_sorted, (const V4*)Fs, c->d.cell_start, c->d.cell_end, c->d.perm, V, s,
                         (const V4*)c->d.pair_table, T(kPairTableIntervals) / s.Rc2};
  if (lpp == 8)
    return submit(c, kTagCorrect, correct_kernel<T, 8>, correct_kernel_group<T, 8>, correct_blocks(c->N * 8), 256, 0, a);
  if (lpp == 4)
    return submit(c, kTagCorrect, correct_kernel<T, 4>, correct_kernel_group<T, 4>, correct_blocks(c->N * 4), 256, 0, a);
  return submit(c, kTagCorrect, correct_kernel<T, 1>, correct_kernel_group<T, 1>, correct_blocks(c->N), 256, 0, a);
}

template hipError_t launch_correct_multi<float>(ffcm_ctx*, float*, const void*, int);
template hipError_t launch_correct_multi<double>(ffcm_ctx*, double*, const void*, int);
template hipError_t launch_correct<float>(ffcm_ctx*, float*, const void*);
template hipError_t launch_correct<double>(ffcm_ctx*, double*, const void*);

}  // namespace ffcm
