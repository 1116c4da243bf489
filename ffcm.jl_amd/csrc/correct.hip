// Step 6 of the mobility apply: the self term and the real-space pair correction.
//
// Reference: correct_velocities! / _correct_velocities_kernel! CPU method
// (src/correct_velocities.jl:154-286), _self_correction (:21-27), _correction_scalars
// (:71-104), _min_image (:121-127). For every slot: v = delta * F_s, then the 27 periodic
// neighbour cells in (dz, dy, dx) order, slots in range order, skipping s itself, adding
// A(r) F_m + B(r) (x . F_m) x for every pair with r^2 < R_c^2; finally V[:, n] += v.
//
// Design: one thread per sorted slot gathers over its 27 cells (no atomics, deterministic, the
// reference's CPU visit order). Consecutive threads hold slots of the same cell and walk the
// same neighbour ranges, so the candidate loads are shared through the cache.
#include "pair_terms.hpp"

namespace ffcm {

namespace {

template <typename T>
__global__ void __launch_bounds__(256) correct_kernel(const typename Vec<T>::v4* __restrict__ Ys,
                                                      const typename Vec<T>::v4* __restrict__ Fs,
                                                      const int* __restrict__ cell_start,
                                                      const int* __restrict__ cell_end,
                                                      const int* __restrict__ perm, T* __restrict__ V,
                                                      Scalars<T> s) {
  __shared__ int list[kPairBatch * 256];
  for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < s.N; slot += gridDim.x * blockDim.x) {
    const typename Vec<T>::v4 yn = Ys[slot];
    const typename Vec<T>::v4 fn = Fs[slot];
    T v0 = s.delta * fn.x, v1 = s.delta * fn.y, v2 = s.delta * fn.z;
    for_each_pair<T>(slot, yn, s, cell_start, cell_end, Ys, list, [&](int m, T x0, T x1, T x2, T rr) {
      T A, B;
      correction_scalars(tsqrt(rr), s, &A, &B);
      const typename Vec<T>::v4 fm = Fs[m];
      const T bx = B * (x0 * fm.x + x1 * fm.y + x2 * fm.z);
      v0 += A * fm.x + bx * x0;
      v1 += A * fm.y + bx * x1;
      v2 += A * fm.z + bx * x2;
    });
    const int n = perm[slot];
    V[3 * n + 0] += v0;
    V[3 * n + 1] += v1;
    V[3 * n + 2] += v2;
  }
}

// RC right-hand sides at once: the pair scalars A(r), B(r) (the erf/exp work) are evaluated
// once per pair and applied to every right-hand side, in the same operation order as
// correct_kernel, so each right-hand side's result is bit-identical to a single-RHS launch.
// Forces of right-hand side q are at Fs[q * N + slot]; velocities at V[q * 3N + 3n + c].
template <typename T, int RC>
__global__ void __launch_bounds__(256) correct_multi_kernel(const typename Vec<T>::v4* __restrict__ Ys,
                                                            const typename Vec<T>::v4* __restrict__ Fs,
                                                            const int* __restrict__ cell_start,
                                                            const int* __restrict__ cell_end,
                                                            const int* __restrict__ perm, T* __restrict__ V,
                                                            Scalars<T> s) {
  const size_t N = (size_t)s.N;
  __shared__ int list[kPairBatch * 256];
  for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < s.N; slot += gridDim.x * blockDim.x) {
    const typename Vec<T>::v4 yn = Ys[slot];
    T v[RC][3];
#pragma unroll
    for (int q = 0; q < RC; ++q) {
      const typename Vec<T>::v4 fn = Fs[q * N + slot];
      v[q][0] = s.delta * fn.x;
      v[q][1] = s.delta * fn.y;
      v[q][2] = s.delta * fn.z;
    }
    for_each_pair<T>(slot, yn, s, cell_start, cell_end, Ys, list, [&](int s2, T x0, T x1, T x2, T rr) {
      T A, B;
      correction_scalars(tsqrt(rr), s, &A, &B);
#pragma unroll
      for (int q = 0; q < RC; ++q) {
        const typename Vec<T>::v4 fm = Fs[q * N + s2];
        const T bx = B * (x0 * fm.x + x1 * fm.y + x2 * fm.z);
        v[q][0] += A * fm.x + bx * x0;
        v[q][1] += A * fm.y + bx * x1;
        v[q][2] += A * fm.z + bx * x2;
      }
    });
    const int n = perm[slot];
#pragma unroll
    for (int q = 0; q < RC; ++q) {
      T* Vq = V + q * 3 * N;
      Vq[3 * n + 0] += v[q][0];
      Vq[3 * n + 1] += v[q][1];
      Vq[3 * n + 2] += v[q][2];
    }
  }
}

}  // namespace

// R right-hand sides (slot-ordered forces Fs[r * N + slot], velocities V[r * 3N + ...]) in
// chunks of 16, 8, 4, 2, 1.
template <typename T>
hipError_t launch_correct_multi(ffcm_ctx* c, T* V, const void* Fs, int R) {
  using V4 = typename Vec<T>::v4;
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  int64_t blocks = (c->N + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > (1 << 20)) blocks = 1 << 20;
  const unsigned nb = (unsigned)blocks;
  for (int r0 = 0; r0 < R;) {
    const int left = R - r0;
    const int rc = left >= 16 ? 16 : left >= 8 ? 8 : left >= 4 ? 4 : left >= 2 ? 2 : 1;
    const V4* F0 = (const V4*)Fs + (size_t)r0 * c->N;
    T* V0 = V + (size_t)r0 * 3 * c->N;
    const V4* Ys = (const V4*)c->d.Y_sorted;
    switch (rc) {
      case 16: correct_multi_kernel<T, 16><<<nb, 256, 0, c->stream>>>(Ys, F0, c->d.cell_start, c->d.cell_end, c->d.perm, V0, s); break;
      case 8: correct_multi_kernel<T, 8><<<nb, 256, 0, c->stream>>>(Ys, F0, c->d.cell_start, c->d.cell_end, c->d.perm, V0, s); break;
      case 4: correct_multi_kernel<T, 4><<<nb, 256, 0, c->stream>>>(Ys, F0, c->d.cell_start, c->d.cell_end, c->d.perm, V0, s); break;
      case 2: correct_multi_kernel<T, 2><<<nb, 256, 0, c->stream>>>(Ys, F0, c->d.cell_start, c->d.cell_end, c->d.perm, V0, s); break;
      default: correct_kernel<T><<<nb, 256, 0, c->stream>>>(Ys, F0, c->d.cell_start, c->d.cell_end, c->d.perm, V0, s); break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    r0 += rc;
  }
  return hipSuccess;
}

template <typename T>
hipError_t launch_correct(ffcm_ctx* c, T* V, const void* Fs) {
  if (!Fs) Fs = c->d.F_sorted;
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  int64_t blocks = (c->N + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > (1 << 20)) blocks = 1 << 20;
  correct_kernel<T><<<(unsigned)blocks, 256, 0, c->stream>>>((const typename Vec<T>::v4*)c->d.Y_sorted,
                                                               (const typename Vec<T>::v4*)Fs,
                                                               c->d.cell_start, c->d.cell_end, c->d.perm, V, s);
  return hipGetLastError();
}

template hipError_t launch_correct_multi<float>(ffcm_ctx*, float*, const void*, int);
template hipError_t launch_correct_multi<double>(ffcm_ctx*, double*, const void*, int);
template hipError_t launch_correct<float>(ffcm_ctx*, float*, const void*);
template hipError_t launch_correct<double>(ffcm_ctx*, double*, const void*);

}  // namespace ffcm
