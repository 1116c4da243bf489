// Step 1-2 of the mobility apply: periodic wrap, cell hashing, a stable LSD radix sort of
// (cell, particle) and the gather of positions and forces into cell order.
//
// Reference: src/cell_list.jl (wrap_positions! :22-31, _build_cell_list_kernel! :64-93,
// _gather_particles_kernel! :159-176, _assign_cells_kernel! :222-238). The reference's CPU
// counting sort is stable (ties in ascending original index); an LSD radix sort whose every
// pass is stable produces exactly that permutation, so slot order, cell_start and cell_end are
// bit-identical to the CPU path (the CUDA extension's atomic scatter is not stable).
#include "group.hpp"
#include "stencil.hpp"

namespace ffcm {

namespace {

// Julia `mod(x::T, y::T)` on floats: r = rem(x, y) (exact); r == 0 gives copysign(0, y); a
// remainder whose sign differs from y is shifted by y (one rounding, may round to y itself).
template <typename T>
__device__ __forceinline__ T jl_mod(T x, T y) {
  T r = fmod(x, y);
  if (r == T(0)) return copysign(T(0), y);
  if ((r > T(0)) != (y > T(0))) return r + y;
  return r;
}

template <typename T>
__device__ __forceinline__ int cell_of(T yw, T inv_cs, int m) {
  // floor(Int32, Y * inv_cell_size), clamped to m - 1 (src/cell_list.jl:229-234). The product
  // is a single rounded multiply; nothing can be contracted into it.
  int c = (int)floor(yw * inv_cs);
  // the lower clamp never triggers for wrapped input; it keeps a NaN position in bounds
  return c > m - 1 ? m - 1 : (c < 0 ? 0 : c);
}

template <typename T>
__global__ void wrap_kernel(const T* __restrict__ Y, T* __restrict__ Yw, int64_t n3, T L0, T L1, T L2) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n3; i += (int64_t)gridDim.x * blockDim.x) {
    int a = (int)(i % 3);
    T L = a == 0 ? L0 : (a == 1 ? L1 : L2);
    Yw[i] = jl_mod(Y[i], L);
  }
}

template <typename T>
__global__ void assign_kernel(const T* __restrict__ Yw, int* __restrict__ hash, int n, Scalars<T> s) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int cx = cell_of(Yw[3 * i + 0], s.inv_cs[0], s.nc[0]);
    int cy = cell_of(Yw[3 * i + 1], s.inv_cs[1], s.nc[1]);
    int cz = cell_of(Yw[3 * i + 2], s.inv_cs[2], s.nc[2]);
    hash[i] = cx + (cy + cz * s.nc[1]) * s.nc[0];
  }
}

// Fused wrap + hash for the pipeline: one pass over Y.
template <typename T>
struct WrapAssignArgs {
  const T* Y;
  T* Yw;
  int* hash;
  int n;
  Scalars<T> s;
  int* zero;
  int nzero;
};

template <typename T>
__device__ __forceinline__ void wrap_assign_body(const Blk& b, const WrapAssignArgs<T>& a) {
  const T* __restrict__ Y = a.Y;
  T* __restrict__ Yw = a.Yw;
  int* __restrict__ hash = a.hash;
  const Scalars<T>& s = a.s;
  const int n = a.n;
  // also zeroes zero[0, nzero) (the spread's tile counts, filled later in the apply)
  for (int i = b.id * blockDim.x + threadIdx.x; i < a.nzero; i += b.n * blockDim.x) a.zero[i] = 0;
  for (int i = b.id * blockDim.x + threadIdx.x; i < n; i += b.n * blockDim.x) {
    T y0 = jl_mod(Y[3 * i + 0], s.L[0]);
    T y1 = jl_mod(Y[3 * i + 1], s.L[1]);
    T y2 = jl_mod(Y[3 * i + 2], s.L[2]);
    Yw[3 * i + 0] = y0;
    Yw[3 * i + 1] = y1;
    Yw[3 * i + 2] = y2;
    int cx = cell_of(y0, s.inv_cs[0], s.nc[0]);
    int cy = cell_of(y1, s.inv_cs[1], s.nc[1]);
    int cz = cell_of(y2, s.inv_cs[2], s.nc[2]);
    hash[i] = cx + (cy + cz * s.nc[1]) * s.nc[0];
  }
}

template <typename T>
__global__ void wrap_assign_kernel(WrapAssignArgs<T> a) {
  wrap_assign_body<T>(plain_blk(), a);
}
template <typename T>
__global__ void wrap_assign_kernel_group(const GroupEntry<WrapAssignArgs<T>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const WrapAssignArgs<T>& a) { wrap_assign_body<T>(b, a); });
}

// ------------------------------------------------------------------ stable LSD radix sort
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kSortThreads = 256;
constexpr int kSortWaves = kSortThreads / 64;
// Keys per thread and tile: 16 (4096-key tiles) from 2^20 keys on, 8 (2048-key tiles) below,
// so that small inputs spread over more blocks (32 for 65,536 keys, whose 16-block scatter took
// 16.9 us per pass) while the digit histogram stays small for the single-block scan. Measured
// cell-list stage at 65,536 keys: 0.086 ms, against 0.090 with 4096-key and 0.104 with
// 1024-key tiles.

// Digit histogram of one tile; hist is digit-major (hist[d * nblocks + b]) so that one
// exclusive scan yields every (digit, block) output offset in stable order.
struct RadixHistArgs {
  const int* keys;
  int n, shift, mask;
  int* hist;
  int nblocks;
};

template <int kSortItems>
__device__ __forceinline__ void radix_hist_body(const Blk& b, const RadixHistArgs& a) {
  constexpr int kSortTile = kSortThreads * kSortItems;
  __shared__ int h[kRadix];
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) h[d] = 0;
  __syncthreads();
  const int base = b.id * kSortTile;
  for (int i = 0; i < kSortItems; ++i) {
    int e = base + i * kSortThreads + threadIdx.x;
    if (e < a.n) atomicAdd(&h[(a.keys[e] >> a.shift) & a.mask], 1);
  }
  __syncthreads();
  for (int d = threadIdx.x; d < kRadix; d += kSortThreads) a.hist[d * a.nblocks + b.id] = h[d];
}

template <int kSortItems>
__global__ void __launch_bounds__(kSortThreads) radix_hist_kernel(RadixHistArgs a) {
  radix_hist_body<kSortItems>(plain_blk(), a);
}
template <int kSortItems>
__global__ void __launch_bounds__(kSortThreads) radix_hist_kernel_group(const GroupEntry<RadixHistArgs>* __restrict__ g,
                                                                         int n) {
  group_run(g, n, [](const Blk& b, const RadixHistArgs& a) { radix_hist_body<kSortItems>(b, a); });
}

// Stable scatter of one tile. The tile is consumed in chunks of kSortThreads keys in input
// order; inside a chunk each wave ranks its keys by digit with ballots (lane order), waves are
// ranked in wave order and chunks in chunk order, so equal digits keep their input order.
// vals == nullptr means the identity permutation (first pass).
struct RadixScatterArgs {
  const int* keys_in;
  const int* vals_in;
  int n, shift, mask;
  const int* offsets;
  int nblocks;
  int* keys_out;
  int* vals_out;
};

template <int kSortItems>
__device__ __forceinline__ void radix_scatter_body(const Blk& blk, const RadixScatterArgs& a) {
  const int* __restrict__ keys_in = a.keys_in;
  const int* __restrict__ vals_in = a.vals_in;
  const int n = a.n, shift = a.shift, mask = a.mask, nblocks = a.nblocks;
  const int* __restrict__ offsets = a.offsets;
  int* __restrict__ keys_out = a.keys_out;
  int* __restrict__ vals_out = a.vals_out;
  constexpr int kSortTile = kSortThreads * kSortItems;
  __shared__ int digit_base[kRadix];
  __shared__ int wave_cnt[kSortWaves][kRadix];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const unsigned long long lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int d = tid; d < kRadix; d += kSortThreads) digit_base[d] = offsets[d * nblocks + blk.id];
  const int base = blk.id * kSortTile;
  for (int it = 0; it < kSortItems; ++it) {
    const int e = base + it * kSortThreads + tid;
    const bool valid = e < n;
    int key = 0, val = 0, d = 0;
    if (valid) {
      key = keys_in[e];
      val = vals_in ? vals_in[e] : e;
      d = (key >> shift) & mask;
    }
    for (int q = tid; q < kSortWaves * kRadix; q += kSortThreads) (&wave_cnt[0][0])[q] = 0;
    __syncthreads();
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < kRadixBits; ++b) {
      const bool bit = (d >> b) & 1;
      unsigned long long vote = __ballot(bit);
      m &= bit ? vote : ~vote;
    }
    const int rank = __popcll(m & lt_mask);
    if (valid && rank == 0) wave_cnt[wave][d] = __popcll(m);
    __syncthreads();
    int chunk_total = 0;
    if (tid < kRadix) {
      int run = 0;
#pragma unroll
      for (int w = 0; w < kSortWaves; ++w) {
        int cnt = wave_cnt[w][tid];
        wave_cnt[w][tid] = run;
        run += cnt;
      }
      chunk_total = run;
    }
    __syncthreads();
    if (valid) {
      const int pos = digit_base[d] + wave_cnt[wave][d] + rank;
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    if (tid < kRadix) digit_base[tid] += chunk_total;
    __syncthreads();
  }
}

template <int kSortItems>
__global__ void __launch_bounds__(kSortThreads) radix_scatter_kernel(RadixScatterArgs a) {
  radix_scatter_body<kSortItems>(plain_blk(), a);
}
template <int kSortItems>
__global__ void __launch_bounds__(kSortThreads)
    radix_scatter_kernel_group(const GroupEntry<RadixScatterArgs>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const RadixScatterArgs& a) { radix_scatter_body<kSortItems>(b, a); });
}

// cell_start / cell_end from the sorted keys: boundary s (0..N) owns the cells between the
// key before it and the key after it. Every cell is written exactly once in each array.
// cell_start is 0-based; cell_end is the exclusive end (= the reference's 1-based inclusive
// end), so an empty cell has end == start.
__global__ void cell_offsets_kernel(const int* __restrict__ keys, int n, int C, int* __restrict__ cell_start,
                                    int* __restrict__ cell_end) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s <= n; s += gridDim.x * blockDim.x) {
    const int prev = s > 0 ? keys[s - 1] : -1;
    const int cur = s < n ? keys[s] : C;
    const int hi = cur < C - 1 ? cur : C - 1;
    for (int c = prev + 1; c <= hi; ++c) cell_start[c] = s;
    for (int c = prev < 0 ? 0 : prev; c < cur; ++c) cell_end[c] = s;
  }
}

template <typename T>
__global__ void gather_kernel(const int* __restrict__ perm, const T* __restrict__ Yw, const T* __restrict__ F,
                              typename Vec<T>::v4* __restrict__ Ys, typename Vec<T>::v4* __restrict__ Fs, int n) {
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    const int i = perm[s];
    typename Vec<T>::v4 y, f;
    y.x = Yw[3 * i + 0];
    y.y = Yw[3 * i + 1];
    y.z = Yw[3 * i + 2];
    y.w = T(0);
    f.x = F[3 * i + 0];
    f.y = F[3 * i + 1];
    f.z = F[3 * i + 2];
    f.w = T(0);
    Ys[s] = y;
    Fs[s] = f;
  }
}

// The apply's step after the sort, one pass over the slots: cell ranges (cell_offsets_kernel's
// boundary s, s = 0..n), the gather of slot s (gather_kernel) and its spread binning
// (bin_count_kernel), with tile_count already zeroed.
template <typename T>
struct GatherBinArgs {
  const int* keys;
  int C;
  int* cell_start;
  int* cell_end;
  const int* perm;
  const T* Yw;
  const T* F;
  typename Vec<T>::v4* Ys;
  typename Vec<T>::v4* Fs;
  int n;
  Scalars<T> s;
  Tiling tl;
  int* tile_count;
  int2* slot_tile;
  int4* origin;
};

template <typename T>
__device__ __forceinline__ void gather_bin_body(const Blk& blk, const GatherBinArgs<T>& a) {
  const int* __restrict__ keys = a.keys;
  const int C = a.C, n = a.n;
  int* __restrict__ cell_start = a.cell_start;
  int* __restrict__ cell_end = a.cell_end;
  const int* __restrict__ perm = a.perm;
  const T* __restrict__ Yw = a.Yw;
  const T* __restrict__ F = a.F;
  typename Vec<T>::v4* __restrict__ Ys = a.Ys;
  typename Vec<T>::v4* __restrict__ Fs = a.Fs;
  const Scalars<T>& s = a.s;
  const Tiling& tl = a.tl;
  for (int b = blk.id * blockDim.x + threadIdx.x; b <= n; b += blk.n * blockDim.x) {
    const int prev = b > 0 ? keys[b - 1] : -1;
    const int cur = b < n ? keys[b] : C;
    const int hi = cur < C - 1 ? cur : C - 1;
    for (int c = prev + 1; c <= hi; ++c) cell_start[c] = b;
    for (int c = prev < 0 ? 0 : prev; c < cur; ++c) cell_end[c] = b;
    if (b < n) {
      const int i = perm[b];
      typename Vec<T>::v4 y;
      y.x = Yw[3 * i + 0];
      y.y = Yw[3 * i + 1];
      y.z = Yw[3 * i + 2];
      y.w = T(0);
      Ys[b] = y;
      if (F) {  // no forces: positions only (the forces are gathered per apply)
        typename Vec<T>::v4 f;
        f.x = F[3 * i + 0];
        f.y = F[3 * i + 1];
        f.z = F[3 * i + 2];
        f.w = T(0);
        Fs[b] = f;
      }
      bin_slot<T>(y, b, s, tl, a.tile_count, a.slot_tile, a.origin);
    }
  }
}

template <typename T>
__global__ void gather_bin_kernel(GatherBinArgs<T> a) {
  gather_bin_body<T>(plain_blk(), a);
}
template <typename T>
__global__ void gather_bin_kernel_group(const GroupEntry<GatherBinArgs<T>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const GatherBinArgs<T>& a) { gather_bin_body<T>(b, a); });
}

// nrhs force sets: set r is read at F + r * 3n and written at Fs + r * n.
template <typename T>
struct GatherForcesArgs {
  const int* perm;
  const T* F;
  typename Vec<T>::v4* Fs;
  int n, nrhs;
};

template <typename T>
__device__ __forceinline__ void gather_forces_body(const Blk& b, const GatherForcesArgs<T>& a) {
  const int n = a.n;
  for (int s = b.id * blockDim.x + threadIdx.x; s < n; s += b.n * blockDim.x) {
    const int i = a.perm[s];
    for (int r = 0; r < a.nrhs; ++r) {
      const T* Fr = a.F + (size_t)r * 3 * n;
      typename Vec<T>::v4 f;
      f.x = Fr[3 * i + 0];
      f.y = Fr[3 * i + 1];
      f.z = Fr[3 * i + 2];
      f.w = T(0);
      a.Fs[(size_t)r * n + s] = f;
    }
  }
}

template <typename T>
__global__ void gather_forces_kernel(GatherForcesArgs<T> a) {
  gather_forces_body<T>(plain_blk(), a);
}
template <typename T>
__global__ void gather_forces_kernel_group(const GroupEntry<GatherForcesArgs<T>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const GatherForcesArgs<T>& a) { gather_forces_body<T>(b, a); });
}

inline unsigned grid_for(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > 1 << 20) b = 1 << 20;
  return (unsigned)b;
}

}  // namespace

template <typename T>
hipError_t launch_wrap(ffcm_ctx* c, const T* Y) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  wrap_kernel<T><<<grid_for(3 * c->N, 256), 256, 0, c->stream>>>(Y, (T*)c->d.Y_wrapped, 3 * c->N, s.L[0], s.L[1],
                                                                   s.L[2]);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_assign_cells(ffcm_ctx* c) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  assign_kernel<T><<<grid_for(c->N, 256), 256, 0, c->stream>>>((const T*)c->d.Y_wrapped, c->d.hash, (int)c->N, s);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_wrap_assign(ffcm_ctx* c, const T* Y) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const WrapAssignArgs<T> a{Y, (T*)c->d.Y_wrapped, c->d.hash, (int)c->N, s, c->d.tile_count, c->tiling.ntiles};
  return submit(c, kTagWrapAssign, wrap_assign_kernel<T>, wrap_assign_kernel_group<T>, grid_for(c->N, 256), 256, 0, a);
}

int sort_items(int64_t n) { return n >= (int64_t(1) << 20) ? 16 : 8; }
int sort_blocks(int64_t n) {
  const int64_t tile = (int64_t)kSortThreads * sort_items(n);
  return (int)((n + tile - 1) / tile);
}

hipError_t launch_sort_cells(ffcm_ctx* c, bool offsets) {
  const int n = (int)c->N;
  const int items = sort_items(c->N);
  const int nblocks = sort_blocks(c->N);
  int* hist = c->d.radix_hist;
  int* offs = c->d.radix_hist + (size_t)kRadix * nblocks;
  const int passes = (c->key_bits + kRadixBits - 1) / kRadixBits;
  const int* kin = c->d.hash;
  const int* vin = nullptr;
  int* kout = c->d.keys_a;
  int* vout = c->d.perm_a;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * kRadixBits;
    int bits = c->key_bits - shift;
    if (bits > kRadixBits) bits = kRadixBits;
    const int mask = (1 << bits) - 1;
    const int tag = kTagSortPass + 10 * p;
    const RadixHistArgs ha{kin, n, shift, mask, hist, nblocks};
    hipError_t e = items == 16 ? submit(c, tag, radix_hist_kernel<16>, radix_hist_kernel_group<16>, nblocks,
                                        kSortThreads, 0, ha)
                               : submit(c, tag, radix_hist_kernel<8>, radix_hist_kernel_group<8>, nblocks,
                                        kSortThreads, 0, ha);
    if (e != hipSuccess) return e;
    e = launch_exclusive_scan(c, tag, hist, offs, (int64_t)kRadix * nblocks, c->d.scan_partials);
    if (e != hipSuccess) return e;
    const RadixScatterArgs sa{kin, vin, n, shift, mask, offs, nblocks, kout, vout};
    e = items == 16 ? submit(c, tag + 5, radix_scatter_kernel<16>, radix_scatter_kernel_group<16>, nblocks,
                             kSortThreads, 0, sa)
                    : submit(c, tag + 5, radix_scatter_kernel<8>, radix_scatter_kernel_group<8>, nblocks,
                             kSortThreads, 0, sa);
    if (e != hipSuccess) return e;
    kin = kout;
    vin = vout;
    if (kout == c->d.keys_a) {
      kout = c->d.keys_b;
      vout = c->d.perm_b;
    } else {
      kout = c->d.keys_a;
      vout = c->d.perm_a;
    }
  }
  c->d.keys_sorted = (int*)kin;
  c->d.perm = (int*)vin;
  if (offsets)
    cell_offsets_kernel<<<grid_for(c->N + 1, 256), 256, 0, c->stream>>>(kin, n, (int)c->C, c->d.cell_start,
                                                                         c->d.cell_end);
  c->tiles_valid = false;
  return hipGetLastError();
}

template <typename T>
hipError_t launch_gather(ffcm_ctx* c, const T* F) {
  gather_kernel<T><<<grid_for(c->N, 256), 256, 0, c->stream>>>(c->d.perm, (const T*)c->d.Y_wrapped, F,
                                                                 (typename Vec<T>::v4*)c->d.Y_sorted,
                                                                 (typename Vec<T>::v4*)c->d.F_sorted, (int)c->N);
  c->tiles_valid = false;
  return hipGetLastError();
}

template <typename T>
hipError_t launch_gather_bin(ffcm_ctx* c, const T* F) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const GatherBinArgs<T> a{c->d.keys_sorted, (int)c->C, c->d.cell_start, c->d.cell_end, c->d.perm,
                           (const T*)c->d.Y_wrapped, F, (typename Vec<T>::v4*)c->d.Y_sorted,
                           (typename Vec<T>::v4*)c->d.F_sorted, (int)c->N, s, c->tiling, c->d.tile_count,
                           (int2*)c->d.tile_cursor, c->d.anchor};
  c->tiles_valid = false;
  return submit(c, kTagGatherBin, gather_bin_kernel<T>, gather_bin_kernel_group<T>, grid_for(c->N + 1, 256), 256, 0,
                a);
}

template <typename T>
hipError_t launch_gather_forces(ffcm_ctx* c, const T* F, void* dst, int nrhs) {
  const GatherForcesArgs<T> a{c->d.perm, F, (typename Vec<T>::v4*)dst, (int)c->N, nrhs};
  return submit(c, kTagGatherForces, gather_forces_kernel<T>, gather_forces_kernel_group<T>, grid_for(c->N, 256), 256,
                0, a);
}

template hipError_t launch_gather_forces<float>(ffcm_ctx*, const float*, void*, int);
template hipError_t launch_gather_forces<double>(ffcm_ctx*, const double*, void*, int);
template hipError_t launch_wrap<float>(ffcm_ctx*, const float*);
template hipError_t launch_wrap<double>(ffcm_ctx*, const double*);
template hipError_t launch_assign_cells<float>(ffcm_ctx*);
template hipError_t launch_assign_cells<double>(ffcm_ctx*);
template hipError_t launch_wrap_assign<float>(ffcm_ctx*, const float*);
template hipError_t launch_wrap_assign<double>(ffcm_ctx*, const double*);
template hipError_t launch_gather<float>(ffcm_ctx*, const float*);
template hipError_t launch_gather_bin<float>(ffcm_ctx*, const float*);
template hipError_t launch_gather_bin<double>(ffcm_ctx*, const double*);
template hipError_t launch_gather<double>(ffcm_ctx*, const double*);

}  // namespace ffcm
