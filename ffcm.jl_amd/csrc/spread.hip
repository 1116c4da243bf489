// Step 3 of the mobility apply: Gaussian spreading of the sorted forces onto the grid.
//
// Reference: spread_forces! / _spread_forces_kernel! (src/spread_forces.jl:17-130): zero the
// three force-density fields, then for every particle add F * w over its M_G^3 stencil.
//
// Design: owner-computes tiles. The grid is cut into tiles whose extent divides M and is at
// least M_G on every axis. Each particle is binned to the tile that holds its stencil's low
// corner (wrapped); its stencil then reaches only that tile and the next one on each axis, so
// a tile's contributors are the particles binned to the 2x2x2 tiles at or below it. One block
// owns one tile: it accumulates every contributor's clipped stencil in LDS (LDS atomics only)
// and writes the tile to HBM once, fully coalesced. The grid is never zero-filled separately
// and no global atomic is issued.
#include "stencil.hpp"

namespace ffcm {

namespace {

constexpr int kSpreadThreads = 256;

// Per slot: home tile of the stencil low corner, and the slot's rank inside that tile.
template <typename T>
__global__ void bin_count_kernel(const typename Vec<T>::v4* __restrict__ Ys, int n, Scalars<T> s, Tiling tl,
                                 int* __restrict__ tile_count, int2* __restrict__ slot_tile) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const typename Vec<T>::v4 y = Ys[i];
    const AxisOrigin ox = axis_origin(y.x, s.inv_h, s.half, s.M[0]);
    const AxisOrigin oy = axis_origin(y.y, s.inv_h, s.half, s.M[1]);
    const AxisOrigin oz = axis_origin(y.z, s.inv_h, s.half, s.M[2]);
    const int t = ox.low / tl.T[0] + tl.nt[0] * (oy.low / tl.T[1] + tl.nt[1] * (oz.low / tl.T[2]));
    const int r = atomicAdd(&tile_count[t], 1);
    slot_tile[i] = make_int2(t, r);
  }
}

__global__ void bin_scatter_kernel(const int2* __restrict__ slot_tile, const int* __restrict__ tile_start, int n,
                                   int* __restrict__ tile_list) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int2 tr = slot_tile[i];
    tile_list[tile_start[tr.x] + tr.y] = i;
  }
}

template <typename T, int MGC>
__global__ void __launch_bounds__(kSpreadThreads) spread_tile_kernel(const typename Vec<T>::v4* __restrict__ Ys,
                                                                      const typename Vec<T>::v4* __restrict__ Fs,
                                                                      const int* __restrict__ tile_start,
                                                                      const int* __restrict__ tile_list,
                                                                      T* __restrict__ grid, Scalars<T> s,
                                                                      Tiling tl) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* acc = (T*)smem_raw;
  const int MG = MGC > 0 ? MGC : s.MG;
  const int Tx = tl.T[0], Ty = tl.T[1], Tz = tl.T[2];
  const int ntx = tl.nt[0], nty = tl.nt[1], ntz = tl.nt[2];
  const int t = blockIdx.x;
  const int tx = t % ntx;
  const int ty = (t / ntx) % nty;
  const int tz = t / (ntx * nty);
  const int ox = tx * Tx, oy = ty * Ty, oz = tz * Tz;
  const int npts = Tx * Ty * Tz;
  for (int i = threadIdx.x; i < 3 * npts; i += blockDim.x) acc[i] = T(0);
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  for (int cand = 0; cand < 8; ++cand) {
    const int dx = cand & 1, dy = (cand >> 1) & 1, dz = cand >> 2;
    if ((dx && ntx == 1) || (dy && nty == 1) || (dz && ntz == 1)) continue;
    const int hx = tx - dx < 0 ? tx - dx + ntx : tx - dx;
    const int hy = ty - dy < 0 ? ty - dy + nty : ty - dy;
    const int hz = tz - dz < 0 ? tz - dz + ntz : tz - dz;
    const int h = hx + ntx * (hy + nty * hz);
    const int beg = tile_start[h], end = tile_start[h + 1];
    for (int i = beg + wave; i < end; i += nwaves) {
      const int slot = tile_list[i];
      const typename Vec<T>::v4 y = Ys[slot];
      const typename Vec<T>::v4 f = Fs[slot];
      const AxisOrigin ax = axis_origin(y.x, s.inv_h, s.half, Mx);
      const AxisOrigin ay = axis_origin(y.y, s.inv_h, s.half, My);
      const AxisOrigin az = axis_origin(y.z, s.inv_h, s.half, Mz);
      AxisEntry<T> ex{T(0), T(0)}, ey{T(0), T(0)}, ez{T(0), T(0)};
      if (lane < MG) {
        ex = axis_entry(ax.lo, lane, y.x, s.h, s.inv_norm, s.inv_2S2);
        ey = axis_entry(ay.lo, lane, y.y, s.h, s.inv_norm, s.inv_2S2);
        ez = axis_entry(az.lo, lane, y.z, s.h, s.inv_norm, s.inv_2S2);
      }
      for (int p0 = 0; p0 < MG * MG; p0 += 64) {
        const int p = p0 + lane;
        const int ky = p / MG;
        const int kx = p - ky * MG;
        const bool valid = p < MG * MG;
        const int lx = wrap_add(ax.low, kx, Mx) - ox;
        const int ly = wrap_add(ay.low, valid ? ky : 0, My) - oy;
        const bool in_xy = valid && lx >= 0 && lx < Tx && ly >= 0 && ly < Ty;
        const T gx = wave_bcast(ex.g, kx), r2x = wave_bcast(ex.r2, kx);
        const T gy = wave_bcast(ey.g, valid ? ky : 0), r2y = wave_bcast(ey.r2, valid ? ky : 0);
        for (int kz = 0; kz < MG; ++kz) {
          const int lz = wrap_add(az.low, kz, Mz) - oz;
          const T gz = wave_bcast(ez.g, kz), r2z = wave_bcast(ez.r2, kz);
          if (lz < 0 || lz >= Tz) continue;
          if (in_xy) {
            const T w = ((s.a0 + s.a2 * (r2x + (r2y + r2z))) * gx) * (gy * gz);
            T* a = acc + ((lz * Ty + ly) * Tx + lx) * 3;
            atomicAdd(a + 0, w * f.x);
            atomicAdd(a + 1, w * f.y);
            atomicAdd(a + 2, w * f.z);
          }
        }
      }
    }
  }
  __syncthreads();
  const int rowlen = 3 * Tx;
  for (int e = threadIdx.x; e < 3 * npts; e += blockDim.x) {
    const int r = e / rowlen;
    const int q = e - r * rowlen;
    const int lz = r / Ty;
    const int ly = r - lz * Ty;
    grid[(((size_t)(oz + lz) * My + (oy + ly)) * Mx + ox) * 3 + q] = acc[e];
  }
}

inline unsigned grid_for(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > 1 << 20) b = 1 << 20;
  return (unsigned)b;
}

// Tile extent for one axis: a divisor of M no smaller than M_G, as close to `target` as
// possible (M itself when nothing smaller divides it).
int pick_tile(int M, int MG, int target) {
  int best = M;
  for (int d = MG; d <= M; ++d) {
    if (M % d) continue;
    if (best == M || (d <= target ? target - d : d - target) < (best <= target ? target - best : best - target))
      best = d;
  }
  return best;
}

}  // namespace

template <typename T>
void plan_tiling(ffcm_ctx* c) {
  Tiling& tl = c->tiling;
  const int MG = c->host.M_G;
  const int targets[3] = {16, 16, 16};
  for (int a = 0; a < 3; ++a) {
    tl.T[a] = pick_tile(c->host.M[a], MG, targets[a]);
    tl.nt[a] = c->host.M[a] / tl.T[a];
  }
  tl.ntiles = tl.nt[0] * tl.nt[1] * tl.nt[2];
  tl.lds_bytes = 3 * tl.T[0] * tl.T[1] * tl.T[2] * (int)sizeof(T);
  (void)hipFuncSetAttribute((const void*)spread_tile_kernel<T, 8>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            tl.lds_bytes);
  (void)hipFuncSetAttribute((const void*)spread_tile_kernel<T, 0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            tl.lds_bytes);
}

template <typename T>
hipError_t launch_bin_tiles(ffcm_ctx* c) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const int n = (int)c->N;
  hipError_t e = launch_fill_i32(c->stream, c->d.tile_count, c->tiling.ntiles, 0);
  if (e != hipSuccess) return e;
  bin_count_kernel<T><<<grid_for(n, 256), 256, 0, c->stream>>>((const typename Vec<T>::v4*)c->d.Y_sorted, n, s,
                                                                 c->tiling, c->d.tile_count,
                                                                 (int2*)c->d.tile_cursor);
  e = launch_exclusive_scan(c->stream, c->d.tile_count, c->d.tile_start, c->tiling.ntiles, c->d.scan_partials);
  if (e != hipSuccess) return e;
  bin_scatter_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>((const int2*)c->d.tile_cursor, c->d.tile_start, n,
                                                               c->d.tile_list);
  c->tiles_valid = true;
  return hipGetLastError();
}

template <typename T>
hipError_t launch_spread(ffcm_ctx* c) {
  if (!c->tiles_valid) {
    hipError_t e = launch_bin_tiles<T>(c);
    if (e != hipSuccess) return e;
  }
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const Tiling& tl = c->tiling;
  if (s.MG == 8)
    spread_tile_kernel<T, 8><<<tl.ntiles, kSpreadThreads, tl.lds_bytes, c->stream>>>(
        (const typename Vec<T>::v4*)c->d.Y_sorted, (const typename Vec<T>::v4*)c->d.F_sorted, c->d.tile_start,
        c->d.tile_list, (T*)c->d.force, s, tl);
  else
    spread_tile_kernel<T, 0><<<tl.ntiles, kSpreadThreads, tl.lds_bytes, c->stream>>>(
        (const typename Vec<T>::v4*)c->d.Y_sorted, (const typename Vec<T>::v4*)c->d.F_sorted, c->d.tile_start,
        c->d.tile_list, (T*)c->d.force, s, tl);
  return hipGetLastError();
}

template void plan_tiling<float>(ffcm_ctx*);
template void plan_tiling<double>(ffcm_ctx*);
template hipError_t launch_bin_tiles<float>(ffcm_ctx*);
template hipError_t launch_bin_tiles<double>(ffcm_ctx*);
template hipError_t launch_spread<float>(ffcm_ctx*);
template hipError_t launch_spread<double>(ffcm_ctx*);

}  // namespace ffcm
