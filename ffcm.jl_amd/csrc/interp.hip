// Step 5 of the mobility apply: interpolation of the grid velocity back to the particles.
//
// Reference: interpolate_velocities! / _interpolate_velocities_kernel!
// (src/interpolate.jl:23-136): the same stencil and weights as spreading,
// v = sum w * u[ix, iy, iz], then V[:, original_index[s]] = h^3 * v, i.e. the output is written
// in the caller's original particle order.
//
// Design: one wave per particle; lanes cover the (kx, ky) plane of the stencil and loop
// over kz, reading the three interleaved velocity components of each point with adjacent
// lanes on adjacent x. Particles are visited in the spread's tile order (grouped by the grid
// tile holding their stencil's low corner), so consecutive waves read overlapping grid
// boxes. Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2; the
// block index is remapped so that every XCD walks one contiguous run of tiles and its L2
// keeps the rows its neighbours reuse. The wave sum is a butterfly reduction.
#include "stencil.hpp"

namespace ffcm {

namespace {

constexpr int kInterpThreads = 256;
constexpr int kXcds = 8;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T, int MGC>
__global__ void __launch_bounds__(kInterpThreads) interp_kernel(const typename Vec<T>::v4* __restrict__ Ys,
                                                                 const int* __restrict__ perm,
                                                                 const int* __restrict__ order,
                                                                 const T* __restrict__ grid, T* __restrict__ V,
                                                                 Scalars<T> s) {
  const int MG = MGC > 0 ? MGC : s.MG;
  const int lane = threadIdx.x & 63;
  const int waves_per_block = blockDim.x >> 6;
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  // gridDim.x is a multiple of kXcds: a bijection of [0, gridDim.x)
  const int per_xcd = gridDim.x / kXcds;
  const int lblock = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  for (int i = lblock * waves_per_block + (threadIdx.x >> 6); i < s.N; i += gridDim.x * waves_per_block) {
    const int slot = order[i];
    const typename Vec<T>::v4 y = Ys[slot];
    const AxisOrigin ax = axis_origin(y.x, s.inv_h, s.half, Mx);
    const AxisOrigin ay = axis_origin(y.y, s.inv_h, s.half, My);
    const AxisOrigin az = axis_origin(y.z, s.inv_h, s.half, Mz);
    AxisEntry<T> ex{T(0), T(0)}, ey{T(0), T(0)}, ez{T(0), T(0)};
    if (lane < MG) {
      ex = axis_entry(ax.lo, lane, y.x, s.h, s.inv_norm, s.inv_2S2);
      ey = axis_entry(ay.lo, lane, y.y, s.h, s.inv_norm, s.inv_2S2);
      ez = axis_entry(az.lo, lane, y.z, s.h, s.inv_norm, s.inv_2S2);
    }
    T v0 = T(0), v1 = T(0), v2 = T(0);
    if constexpr (MGC > 0 && MGC * MGC <= 64) {
      // one pass of the (kx, ky) plane: all MG z-loads of a lane are issued before any is
      // used, and the z entries (wave-uniform index) are scalar register reads
      const bool valid = lane < MGC * MGC;
      const int ky = valid ? lane / MGC : 0;
      const int kx = valid ? lane - ky * MGC : 0;
      const int ix = wrap_add(ax.low, kx, Mx);
      const int iy = wrap_add(ay.low, ky, My);
      const T gx = wave_bcast(ex.g, kx), r2x = wave_bcast(ex.r2, kx);
      const T gy = wave_bcast(ey.g, ky), r2y = wave_bcast(ey.r2, ky);
      const T* col = grid + ((size_t)iy * Mx + ix) * 3;
      const size_t plane = (size_t)My * Mx * 3;
      T u[MGC][3];
#pragma unroll
      for (int kz = 0; kz < MGC; ++kz) {
        const T* q = col + (size_t)wrap_add(az.low, kz, Mz) * plane;
        u[kz][0] = q[0];
        u[kz][1] = q[1];
        u[kz][2] = q[2];
      }
#pragma unroll
      for (int kz = 0; kz < MGC; ++kz) {
        const T gz = lane_bcast(ez.g, kz), r2z = lane_bcast(ez.r2, kz);
        const T w = ((s.a0 + s.a2 * (r2x + (r2y + r2z))) * gx) * (gy * gz);
        if (valid) {
          v0 += w * u[kz][0];
          v1 += w * u[kz][1];
          v2 += w * u[kz][2];
        }
      }
    } else
    for (int p0 = 0; p0 < MG * MG; p0 += 64) {
      const int p = p0 + lane;
      const bool valid = p < MG * MG;
      const int ky = valid ? p / MG : 0;
      const int kx = valid ? p - ky * MG : 0;
      const int ix = wrap_add(ax.low, kx, Mx);
      const int iy = wrap_add(ay.low, ky, My);
      const T gx = wave_bcast(ex.g, kx), r2x = wave_bcast(ex.r2, kx);
      const T gy = wave_bcast(ey.g, ky), r2y = wave_bcast(ey.r2, ky);
      const size_t rowxy = ((size_t)iy * Mx + ix) * 3;
      for (int kz = 0; kz < MG; ++kz) {
        const int iz = wrap_add(az.low, kz, Mz);
        const T gz = wave_bcast(ez.g, kz), r2z = wave_bcast(ez.r2, kz);
        if (valid) {
          const T w = ((s.a0 + s.a2 * (r2x + (r2y + r2z))) * gx) * (gy * gz);
          const T* u = grid + (size_t)iz * My * Mx * 3 + rowxy;
          v0 += w * u[0];
          v1 += w * u[1];
          v2 += w * u[2];
        }
      }
    }
    v0 = wave_sum(v0);
    v1 = wave_sum(v1);
    v2 = wave_sum(v2);
    if (lane == 0) {
      const int n = perm[slot];
      V[3 * n + 0] = s.h3 * v0;
      V[3 * n + 1] = s.h3 * v1;
      V[3 * n + 2] = s.h3 * v2;
    }
  }
}

}  // namespace

template <typename T>
hipError_t launch_interpolate(ffcm_ctx* c, T* V, const T* grid) {
  if (!grid) grid = (const T*)c->d.velocity;
  if (!c->tiles_valid) {  // the visiting order is the spread's tile lists
    hipError_t e = launch_bin_tiles<T>(c);
    if (e != hipSuccess) return e;
  }
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const int waves = kInterpThreads / 64;
  int64_t blocks = (c->N + waves - 1) / waves;
  if (blocks > (1 << 20)) blocks = 1 << 20;
  if (blocks < 1) blocks = 1;
  blocks = (blocks + kXcds - 1) / kXcds * kXcds;
  if (s.MG == 8)
    interp_kernel<T, 8><<<(unsigned)blocks, kInterpThreads, 0, c->stream>>>(
        (const typename Vec<T>::v4*)c->d.Y_sorted, c->d.perm, c->d.tile_list, grid, V, s);
  else
    interp_kernel<T, 0><<<(unsigned)blocks, kInterpThreads, 0, c->stream>>>(
        (const typename Vec<T>::v4*)c->d.Y_sorted, c->d.perm, c->d.tile_list, grid, V, s);
  return hipGetLastError();
}

template hipError_t launch_interpolate<float>(ffcm_ctx*, float*, const float*);
template hipError_t launch_interpolate<double>(ffcm_ctx*, double*, const double*);

}  // namespace ffcm
