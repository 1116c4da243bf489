// The allocation-free mobility apply: ffcm_mobility_device's pipeline with every intermediate
// buffer (cell list, sort ping-pong, slot-ordered particles, tile lists, grids, spectrum,
// twiddles, wavevectors) placed in one file-scope device array instead of context memory, for
// callers that cannot allocate device memory (code under graph capture, serving loops,
// harnesses that forbid hipMalloc).
//
// The array is kScratchBytes (128 MiB) per device, which holds, for example, a 128^3 Float32
// grid with 2^17 particles. Every call on a device shares that device's array: calls that
// overlap in time on one device race (the caller orders them, e.g. on one stream). Nothing in
// the array needs to start at zero: the pipeline's kernels write every value they read (the
// tile counters are cleared by the first kernel, the grids are written whole).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ffcm_internal.hpp"

using namespace ffcm;

namespace {

constexpr size_t kScratchBytes = size_t(128) << 20;
constexpr size_t kAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices];  // device address of `scratch` on each device
int g_arch_ok[kMaxDevices];    // 0 unknown, 1 gfx950, -1 other
thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

// Device address of this device's copy of `scratch` (looked up once per device, with the
// device current) and whether the device is a gfx950.
int device_scratch(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(FFCM_ERR_DEVICE, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_arch_ok[dev] == 0) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return fail(FFCM_ERR_DEVICE, "hipGetDeviceProperties failed");
    g_arch_ok[dev] = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : -1;
  }
  if (g_arch_ok[dev] < 0)
    return fail(FFCM_ERR_DEVICE, "the device is not a gfx950; this library carries gfx950 code only");
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(FFCM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    }
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return FFCM_OK;
}

// The FFT twiddles exp(-2 pi i m / n) of the three axes (double, rounded to T) and the
// wavevectors of _derive_scalars (src/config.jl:352-356) and _wrap_wavevectors (:233-237):
// k = (T(2) T(pi)) T(m) / L in T, m = j - 1 on the leading half and j - 1 - M after it (x
// keeps the leading M/2 + 1). One thread per entry.
template <typename T>
__global__ void tables_kernel(T* __restrict__ tw, int n0, int n1, int n2, T* __restrict__ kx, T* __restrict__ ky,
                              T* __restrict__ kz, int Mh, int My, int Mz, T Lx, T Ly, T Lz) {
  const int total = n0 + n1 + n2 + Mh + My + Mz;
  const T twopi = T(2) * T(3.141592653589793238462643383279502884);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    int j = i;
    if (j < n0 + n1 + n2) {
      const int n = j < n0 ? n0 : (j < n0 + n1 ? n1 : n2);
      const int base = j < n0 ? 0 : (j < n0 + n1 ? n0 : n0 + n1);
      const int m = j - base;
      const double ang = 2.0 * 3.141592653589793238462643383279502884 * (double)m / (double)n;
      tw[2 * j] = (T)cos(ang);
      tw[2 * j + 1] = (T)-sin(ang);
      continue;
    }
    j -= n0 + n1 + n2;
    if (j < Mh) {
      kx[j] = twopi * (T)j / Lx;
      continue;
    }
    j -= Mh;
    if (j < My) {
      const int m = j <= My / 2 ? j : j - My;
      ky[j] = twopi * (T)m / Ly;
      continue;
    }
    j -= My;
    const int m = j <= Mz / 2 ? j : j - Mz;
    kz[j] = twopi * (T)m / Lz;
  }
}

template <typename T>
int run_noalloc(const ffcm_params* params, void* V, const void* Y, const void* F, hipStream_t stream, int dev,
                unsigned char* base) {
  ffcm_ctx c;
  c.dtype = params->dtype;
  c.host = *params;
  c.N = params->N;
  c.device = dev;
  std::memset(&c.d, 0, sizeof(c.d));
  std::memset(c.events, 0, sizeof(c.events));
  std::memset(c.pinned, 0, sizeof(c.pinned));
  c.timing = false;
  c.tiles_valid = false;
  c.positions_valid = false;
  c.nrhs = 1;
  c.borrowed_grids = false;
  c.borrowed_stream = true;
  c.stream = stream;
  Scalars<T> sc;
  fill_scalars<T>(params, &sc);
  c.scalars = &sc;
  if (int rc = plan_context<T>(&c)) return fail(rc, c.err);
  std::vector<BufReq> reqs;
  buffer_list<T>(&c, false, &reqs);
  size_t off = 0;
  for (const BufReq& r : reqs) {
    *r.slot = base + off;
    off += (r.bytes + kAlign - 1) / kAlign * kAlign;
  }
  if (off > kScratchBytes)
    return fail(FFCM_ERR_ARGUMENT, "this configuration needs " + std::to_string(off) +
                                       " bytes of device scratch; the allocation-free entry point holds " +
                                       std::to_string(kScratchBytes) + " (use ffcm_create and ffcm_mobility_device)");
  bind_aliases(&c);
  const int n0 = c.fft[0].n, n1 = c.fft[1].n, n2 = c.fft[2].n;
  const int My = params->M[1], Mz = params->M[2];
  const int total = n0 + n1 + n2 + c.Mh + My + Mz;
  tables_kernel<T><<<(total + 255) / 256, 256, 0, stream>>>((T*)c.d.twiddle_all, n0, n1, n2, (T*)c.d.k_x,
                                                             (T*)c.d.k_y, (T*)c.d.k_z, c.Mh, My, Mz, (T)params->L[0],
                                                             (T)params->L[1], (T)params->L[2]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(FFCM_ERR_DEVICE, std::string("tables_kernel: ") + hipGetErrorString(e));
  if (int rc = enqueue_mobility(&c, V, Y, F, 1)) return fail(rc, c.err);
  return FFCM_OK;
}

}  // namespace

extern "C" {

int ffcm_mobility_device_noalloc(const ffcm_params* params, void* V, const void* Y, const void* F, int64_t N,
                                 void* stream) {
  g_error.clear();
  std::string msg;
  if (int rc = validate_params(params, &msg)) return fail(rc, msg);
  if (N != params->N) return fail(FFCM_ERR_DIMENSION, "V, Y and F must each be 3xN for params->N");
  if (!V || !Y || !F) return fail(FFCM_ERR_ARGUMENT, "V, Y and F must not be NULL");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return fail(FFCM_ERR_DEVICE, "no HIP device is available (gpu_acceleration needs an MI355X)");
  }
  void* base = nullptr;
  if (int rc = device_scratch(dev, &base)) return rc;
  if (params->dtype == FFCM_F32)
    return run_noalloc<float>(params, V, Y, F, (hipStream_t)stream, dev, (unsigned char*)base);
  return run_noalloc<double>(params, V, Y, F, (hipStream_t)stream, dev, (unsigned char*)base);
}

const char* ffcm_noalloc_last_error(void) { return g_error.c_str(); }

size_t ffcm_noalloc_scratch_bytes(void) { return kScratchBytes; }

}  // extern "C"
