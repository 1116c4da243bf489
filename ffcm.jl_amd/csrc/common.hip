// Device utilities shared by the pipeline: fills and an exclusive scan of int32 counts (cell
// offsets, tile offsets, radix digit offsets): one block for small inputs, three kernels
// (block sums, scan of the sums, apply) otherwise. Neither uses hipMemsetAsync, so
// both may be captured in a HIP graph.
#include "group.hpp"

namespace ffcm {

namespace {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;  // elements per block

__global__ void fill_i32_kernel(int* __restrict__ p, int64_t n, int v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename T>
__global__ void fill_kernel(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Block-wide inclusive scan of one value per thread (kScanThreads threads); `lds` holds
// kScanThreads/64 + 1 ints. Returns the inclusive prefix; *total receives the block sum.
__device__ int block_inclusive_scan(int v, int* lds, int* total) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  if (lane == 63) lds[wave] = v;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int w = 0; w < kScanThreads / 64; ++w) {
      int t = lds[w];
      lds[w] = acc;
      acc += t;
    }
    lds[kScanThreads / 64] = acc;
  }
  __syncthreads();
  v += lds[wave];
  *total = lds[kScanThreads / 64];
  __syncthreads();
  return v;
}

// Arguments of every scan kernel (each uses the fields it needs).
struct ScanArgs {
  const int* in;
  int64_t n;
  int* partials;
  int* out;
};

__device__ __forceinline__ void scan_reduce_body(const Blk& b, const ScanArgs& a) {
  const int* __restrict__ in = a.in;
  const int64_t n = a.n;
  __shared__ int lds[kScanThreads / 64 + 1];
  const int64_t base = (int64_t)b.id * kScanTile;
  int sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t e = base + i * kScanThreads + threadIdx.x;
    if (e < n) sum += in[e];
  }
  int total;
  block_inclusive_scan(sum, lds, &total);
  if (threadIdx.x == 0) a.partials[b.id] = total;
}

// Exclusive scan of the block partials in place (one block, any count): a.n is the number of
// partials.
__device__ __forceinline__ void scan_partials_body(const Blk&, const ScanArgs& a) {
  int* __restrict__ partials = a.partials;
  const int64_t nb = a.n;
  __shared__ int lds[kScanThreads / 64 + 1];
  int carry = 0;
  for (int64_t base = 0; base < nb; base += kScanThreads) {
    int64_t e = base + threadIdx.x;
    int v = e < nb ? partials[e] : 0;
    int total;
    int incl = block_inclusive_scan(v, lds, &total);
    if (e < nb) partials[e] = carry + incl - v;
    carry += total;
  }
}

__device__ __forceinline__ void scan_apply_body(const Blk& b, const ScanArgs& a) {
  const int* __restrict__ in = a.in;
  const int64_t n = a.n;
  const int* __restrict__ partials = a.partials;
  int* __restrict__ out = a.out;
  __shared__ int lds[kScanThreads / 64 + 1];
  const int64_t base = (int64_t)b.id * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int v[kScanItems];
  int sum = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t e = base + i;
    v[i] = e < n ? in[e] : 0;
    sum += v[i];
  }
  int total;
  int incl = block_inclusive_scan(sum, lds, &total);
  int run = partials[b.id] + incl - sum;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t e = base + i;
    if (e < n) out[e] = run;
    run += v[i];
  }
  // The thread holding element n-1 writes the grand total.
  if (base <= n - 1 && n - 1 < base + kScanItems) out[n] = run;
}

// Whole scan in one block when the input is small (radix digit counts, tile counts): the input
// is loaded into LDS with coalesced loads all in flight at once, each of kSingleThreads threads
// scans a contiguous chunk there, one block-wide scan joins the chunks, and the result leaves
// with coalesced stores. Saves the two extra launches of the three-kernel form, which dominate
// at these sizes. LDS: kSingleMax values, padded by one word per 32 (chunk reads of consecutive
// threads fall on different banks): 132 KiB. (Up to 2^15 values: the digit counts of 186,510
// keys, 23,552 values, scanned by three kernels took 13.6 us per radix pass.)
constexpr int kSingleThreads = 1024;
constexpr int64_t kSingleMax = 1 << 15;
constexpr int kSingleItems = (int)(kSingleMax / kSingleThreads);
constexpr size_t kSingleLds = (kSingleMax + kSingleMax / 32) * sizeof(int);

__device__ __forceinline__ int single_pad(int i) { return i + (i >> 5); }

__device__ __forceinline__ void scan_single_body(const Blk&, const ScanArgs& a) {
  const int* __restrict__ in = a.in;
  const int64_t n64 = a.n;
  int* __restrict__ out = a.out;
  extern __shared__ int buf[];
  __shared__ int lds[kSingleThreads / 64 + 1];
  const int tid = threadIdx.x;
  const int n = (int)n64;
#pragma unroll
  for (int k = 0; k < kSingleItems; ++k) {
    const int i = k * kSingleThreads + tid;
    buf[single_pad(i)] = in[i < n ? i : n - 1];  // (entries past n are never read)
  }
  __syncthreads();
  const int chunk = (n + kSingleThreads - 1) / kSingleThreads;
  const int b = tid * chunk;
  const int e = b + chunk < n ? b + chunk : n;
  int sum = 0;
  for (int i = b; i < e; ++i) sum += buf[single_pad(i)];
  // block-wide inclusive scan of the chunk sums
  const int lane = tid & 63, wave = tid >> 6;
  int v = sum;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  if (lane == 63) lds[wave] = v;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int w = 0; w < kSingleThreads / 64; ++w) {
      const int t = lds[w];
      lds[w] = acc;
      acc += t;
    }
    lds[kSingleThreads / 64] = acc;
  }
  __syncthreads();
  int run = v - sum + lds[wave];
  for (int i = b; i < e; ++i) {
    const int x = buf[single_pad(i)];
    buf[single_pad(i)] = run;
    run += x;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSingleItems; ++k) {
    const int i = k * kSingleThreads + tid;
    if (i < n) out[i] = buf[single_pad(i)];
  }
  if (tid == 0) out[n] = lds[kSingleThreads / 64];
}

#define FFCM_SCAN_KERNELS(NAME, BOUNDS)                                                              \
  __global__ void __launch_bounds__(BOUNDS) NAME##_kernel(ScanArgs a) { NAME##_body(plain_blk(), a); } \
  __global__ void __launch_bounds__(BOUNDS) NAME##_kernel_group(const GroupEntry<ScanArgs>* __restrict__ g, int n) { \
    group_run(g, n, [](const Blk& b, const ScanArgs& a) { NAME##_body(b, a); });                 \
  }
FFCM_SCAN_KERNELS(scan_reduce, kScanThreads)
FFCM_SCAN_KERNELS(scan_partials, kScanThreads)
FFCM_SCAN_KERNELS(scan_apply, kScanThreads)
FFCM_SCAN_KERNELS(scan_single, kSingleThreads)
#undef FFCM_SCAN_KERNELS

}  // namespace

size_t scan_partials_needed(int64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

hipError_t launch_fill_i32(hipStream_t s, int* p, int64_t n, int v) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  fill_i32_kernel<<<(unsigned)blocks, 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_fill(hipStream_t s, T* p, int64_t n, T v) {
  if (n <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  fill_kernel<T><<<(unsigned)blocks, 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}
template hipError_t launch_fill<float>(hipStream_t, float*, int64_t, float);
template hipError_t launch_fill<double>(hipStream_t, double*, int64_t, double);

hipError_t launch_exclusive_scan(hipStream_t s, const int* in, int* out, int64_t n, int* partials) {
  if (n <= 0) return launch_fill_i32(s, out, 1, 0);
  if (n <= kSingleMax) {
    scan_single_kernel<<<1, kSingleThreads, kSingleLds, s>>>(ScanArgs{in, n, nullptr, out});
    return hipGetLastError();
  }
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  scan_reduce_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(ScanArgs{in, n, partials, nullptr});
  scan_partials_kernel<<<1, kScanThreads, 0, s>>>(ScanArgs{nullptr, nb, partials, nullptr});
  scan_apply_kernel<<<(unsigned)nb, kScanThreads, 0, s>>>(ScanArgs{in, n, partials, out});
  return hipGetLastError();
}

// The same scan on the context's stream, recorded under tag + kScan* when the context records
// (n >= 1 there: the counts scanned by the pipeline are never empty).
hipError_t launch_exclusive_scan(ffcm_ctx* c, int tag, const int* in, int* out, int64_t n, int* partials) {
  if (!c->rec) return launch_exclusive_scan(c->stream, in, out, n, partials);
  if (n <= 0) return hipErrorInvalidValue;
  if (n <= kSingleMax)
    return submit(c, tag + kScanSingle, scan_single_kernel, scan_single_kernel_group, 1, kSingleThreads, kSingleLds,
                  ScanArgs{in, n, nullptr, out});
  const int64_t nb = (n + kScanTile - 1) / kScanTile;
  hipError_t e = submit(c, tag + kScanReduce, scan_reduce_kernel, scan_reduce_kernel_group, (unsigned)nb,
                        kScanThreads, 0, ScanArgs{in, n, partials, nullptr});
  if (e == hipSuccess)
    e = submit(c, tag + kScanPartials, scan_partials_kernel, scan_partials_kernel_group, 1, kScanThreads, 0,
               ScanArgs{nullptr, nb, partials, nullptr});
  if (e == hipSuccess)
    e = submit(c, tag + kScanApply, scan_apply_kernel, scan_apply_kernel_group, (unsigned)nb, kScanThreads, 0,
               ScanArgs{in, n, partials, out});
  return e;
}

}  // namespace ffcm
