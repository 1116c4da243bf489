// Step 3 of the mobility apply: Gaussian spreading of the sorted forces onto the grid.
//
// Reference: spread_forces! / _spread_forces_kernel! (src/spread_forces.jl:17-130): zero the
// three force-density fields, then for every particle add F * w over its M_G^3 stencil.
//
// Design: owner-computes tiles. The grid is cut into tiles whose extent divides M on every
// axis. Each particle is binned to the tile that holds its stencil's low corner (wrapped); its
// stencil then reaches that tile and the next span - 1 tiles on each axis, so a tile's
// contributors are the particles binned to the span^3 tiles at or below it (2x2x2 for 16^3
// tiles and M_G = 8). One block owns one tile: it stages its contributors in LDS, accumulates
// their clipped stencils in LDS with each wave owning whole z planes (no atomics of any kind)
// and writes the tile to HBM once, fully coalesced. The grid is never zero-filled separately.
#include <cstdio>
#include <cstdlib>

#include "stencil.hpp"

namespace ffcm {

namespace {

constexpr int kSpreadThreads = 512;
constexpr int kMaxHome = 64;  // home tiles that can reach one tile

// Per slot: the wrapped stencil low corner on each axis and its home tile (kept for the
// spread kernel), and the slot's rank inside the home tile.
template <typename T>
__global__ void bin_count_kernel(const typename Vec<T>::v4* __restrict__ Ys, int n, Scalars<T> s, Tiling tl,
                                 int* __restrict__ tile_count, int2* __restrict__ slot_tile,
                                 int4* __restrict__ origin) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    bin_slot<T>(Ys[i], i, s, tl, tile_count, slot_tile, origin);
}

__global__ void bin_scatter_kernel(const int2* __restrict__ slot_tile, const int* __restrict__ tile_start, int n,
                                   int* __restrict__ tile_list) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int2 tr = slot_tile[i];
    tile_list[tile_start[tr.x] + tr.y] = i;
  }
}

// Does the cyclic stencil interval [low, low + MG) intersect the tile interval [o, o + T)?
__device__ __forceinline__ bool cyclic_overlap(int low, int MG, int o, int T, int M) {
  int d = o - low;
  d = d < 0 ? d + M : d;  // tile start relative to the stencil start
  int e = low - o;
  e = e < 0 ? e + M : e;  // stencil start relative to the tile start
  return d < MG || e < T;
}

// One block owns one tile. Per batch of contributors: (1) stage the particles binned to the
// home tiles that can reach this tile, filtered by an exact overlap test, in LDS; (2) compute
// every staged contributor's 1-D stencil tables (Gaussian and squared offset per axis entry, and
// the wrapped stencil origin) with all threads in parallel; (3) each wave adds all contributors
// to the z planes it owns. The tile is written to HBM once at the end.
// RC right-hand sides share the staging, the tables and each stencil weight: right-hand side
// q reads forces Fs[q * N + slot] and writes grid + q * 3 M^3.
template <typename T, int MGC, int RC>
__global__ void __launch_bounds__(kSpreadThreads) spread_tile_kernel(const typename Vec<T>::v4* __restrict__ Ys,
                                                                      const typename Vec<T>::v4* __restrict__ Fs,
                                                                      const int4* __restrict__ origin,
                                                                      const int* __restrict__ tile_start,
                                                                      const int* __restrict__ tile_list,
                                                                      T* __restrict__ grid, Scalars<T> s,
                                                                      Tiling tl) {
  using V4 = typename Vec<T>::v4;
  using C2 = typename Vec<T>::c2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int MG = MGC > 0 ? MGC : s.MG;
  const int Tx = tl.T[0], Ty = tl.T[1], Tz = tl.T[2];
  const int npts = Tx * Ty * Tz;
  const int cap = tl.cand_cap;
  V4* cand_y = (V4*)smem_raw;                  // cap
  V4* cand_f = cand_y + cap;                   // RC * cap (right-hand side major)
  C2* tab = (C2*)(cand_f + RC * cap);          // cap * 3 * MG: (gauss, x^2) per axis entry
  int* org = (int*)(tab + (size_t)cap * 3 * MG);  // cap * 4: wrapped stencil origin per axis
  T* acc = (T*)(org + 4 * cap);                // RC * 3 * npts (right-hand side major)
  const size_t Ns = (size_t)s.N;
  __shared__ int home_start[kMaxHome];
  __shared__ int home_prefix[kMaxHome + 1];
  __shared__ int ncand;
  const int ntx = tl.nt[0], nty = tl.nt[1];
  const int t = blockIdx.x;
  const int tx = t % ntx;
  const int ty = (t / ntx) % nty;
  const int tz = t / (ntx * nty);
  const int ox = tx * Tx, oy = ty * Ty, oz = tz * Tz;
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  const bool fast_xy = Mx >= Tx + 2 * MG && My >= Ty + 2 * MG;
  const int nhome = tl.span[0] * tl.span[1] * tl.span[2];
  for (int i = threadIdx.x; i < RC * 3 * npts; i += blockDim.x) acc[i] = T(0);
  for (int k = threadIdx.x; k < nhome; k += blockDim.x) {
    const int dx = k % tl.span[0];
    const int dy = (k / tl.span[0]) % tl.span[1];
    const int dz = k / (tl.span[0] * tl.span[1]);
    const int hx = (tx - dx + tl.nt[0]) % tl.nt[0];
    const int hy = (ty - dy + tl.nt[1]) % tl.nt[1];
    const int hz = (tz - dz + tl.nt[2]) % tl.nt[2];
    const int h = hx + tl.nt[0] * (hy + tl.nt[1] * hz);
    home_start[k] = tile_start[h];
    home_prefix[k + 1] = tile_start[h + 1] - tile_start[h];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    home_prefix[0] = 0;
    for (int k = 0; k < nhome; ++k) home_prefix[k + 1] += home_prefix[k];
  }
  __syncthreads();
  const int total = home_prefix[nhome];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int TyTx3 = Ty * Tx * 3;
  for (int base = 0; base < total; base += cap) {
    if (threadIdx.x == 0) ncand = 0;
    __syncthreads();
    // (1) stage the overlapping contributors
    const int lim = min(total, base + cap);
    for (int q = base + threadIdx.x; q < lim; q += blockDim.x) {
      int k = 0;
      while (home_prefix[k + 1] <= q) ++k;
      const int slot = tile_list[home_start[k] + (q - home_prefix[k])];
      const int4 o = origin[slot];
      if (cyclic_overlap(o.x, MG, ox, Tx, Mx) && cyclic_overlap(o.y, MG, oy, Ty, My) &&
          cyclic_overlap(o.z, MG, oz, Tz, Mz)) {
        const int j = atomicAdd(&ncand, 1);
        cand_y[j] = Ys[slot];
#pragma unroll
        for (int q = 0; q < RC; ++q) cand_f[q * cap + j] = Fs[q * Ns + slot];
        org[4 * j + 0] = o.x;
        org[4 * j + 1] = o.y;
        org[4 * j + 2] = o.z;
      }
    }
    __syncthreads();
    const int n = ncand;
    // (2) 1-D tables of every contributor, all threads in parallel
    for (int e = threadIdx.x; e < n * 3 * MG; e += blockDim.x) {
      const int c = e / (3 * MG);
      const int r = e - c * 3 * MG;
      const int a = r / MG;
      const int k = r - a * MG;
      const V4 y = cand_y[c];
      const T ya = a == 0 ? y.x : (a == 1 ? y.y : y.z);
      const int lo = (int)rint(ya * s.inv_h) - s.half;
      const AxisEntry<T> en = axis_entry(lo, k, ya, s.h, s.inv_norm, s.inv_2S2);
      tab[e] = C2{en.g, en.r2};
    }
    __syncthreads();
    // (3) accumulate. Each wave owns the tile's z planes lz = wave, wave + nwaves, ... and adds
    // every contributor's slice at that plane: lanes cover the (x, y) plane of the stencil, so
    // within one instruction every lane updates a different point and no other wave touches
    // the plane. The update is a plain LDS read-modify-write (LDS float atomics are serialised
    // per lane on this hardware).
    if constexpr (MGC == 8) {
      if (fast_xy) {
        // Contributors screened 64 at a time, one per lane: the plane's stencil entry kz and the
        // tile-local x, y offsets of the stencil origin (one offset per axis: the grid holds
        // at least T + 2 M_G points, so a stencil's points in the tile are consecutive). The
        // wave then visits only the contributors that reach its plane, reading their offsets as
        // scalars; each costs one round of LDS reads (tables, forces, accumulator).
        const int kxl = lane & 7, kyl = lane >> 3;
        for (int lz = wave; lz < Tz; lz += nwaves) {
          T* aplane = acc + lz * TyTx3;
          const int gz = oz + lz;
          for (int c0 = 0; c0 < n; c0 += 64) {
            const int c = c0 + lane;
            int dx = 0, dy = 0, kz = 8;
            if (c < n) {
              const int4 o = ((const int4*)org)[c];
              kz = gz - o.z;
              kz += kz < 0 ? Mz : 0;
              dx = o.x - ox;
              dx -= dx >= Tx ? Mx : 0;
              dy = o.y - oy;
              dy -= dy >= Ty ? My : 0;
            }
            unsigned long long mask = __ballot(kz < 8 && dx > -8 && dy > -8);
            while (mask) {
              const int l = __ffsll((long long)mask) - 1;
              mask &= mask - 1;
              const int cc = c0 + l;
              const int skz = __builtin_amdgcn_readlane(kz, l);
              const int lx = __builtin_amdgcn_readlane(dx, l) + kxl;
              const int ly = __builtin_amdgcn_readlane(dy, l) + kyl;
              if ((unsigned)lx < (unsigned)Tx && (unsigned)ly < (unsigned)Ty) {
                const C2* tc = tab + cc * 24;
                const C2 ex = tc[kxl];
                const C2 ey = tc[8 + kyl];
                const C2 ez = tc[16 + skz];
                const T w = (s.a0 + s.a2 * (ex.y + ey.y + ez.y)) * (ex.x * ey.x * ez.x);
                const int pt = (ly * Tx + lx) * 3;
#pragma unroll
                for (int q = 0; q < RC; ++q) {
                  const V4 f = cand_f[q * cap + cc];
                  T* a = aplane + pt + q * 3 * npts;
                  a[0] += w * f.x;
                  a[1] += w * f.y;
                  a[2] += w * f.z;
                }
              }
            }
          }
        }
        __syncthreads();
        continue;
      }
    }
    for (int lz = wave; lz < Tz; lz += nwaves) {
      T* aplane = acc + lz * TyTx3;
      for (int c = 0; c < n; ++c) {
        const int lowz = org[4 * c + 2];
        int kz = oz + lz - lowz;
        kz = kz < 0 ? kz + Mz : kz;
        if (kz >= MG) continue;  // wave-uniform
        const int lowx = org[4 * c + 0], lowy = org[4 * c + 1];
        const C2* tc = tab + (size_t)c * 3 * MG;
        const C2 ez = tc[2 * MG + kz];
        for (int p0 = 0; p0 < MG * MG; p0 += 64) {
          const int p = p0 + lane;
          const bool valid = p < MG * MG;
          const int ky = valid ? p / MG : 0;
          const int kx = valid ? p - ky * MG : 0;
          const int lx = wrap_add(lowx, kx, Mx) - ox;
          const int ly = wrap_add(lowy, ky, My) - oy;
          if (valid && lx >= 0 && lx < Tx && ly >= 0 && ly < Ty) {
            const C2 ex = tc[kx];
            const C2 ey = tc[MG + ky];
            // reassociated: a0 + a2 (r2x + r2y + r2z), gx gy gz
            const T w = (s.a0 + s.a2 * (ex.y + ey.y + ez.y)) * (ex.x * ey.x * ez.x);
#pragma unroll
            for (int q = 0; q < RC; ++q) {
              const V4 f = cand_f[q * cap + c];
              T* a = aplane + q * 3 * npts + (ly * Tx + lx) * 3;
              a[0] += w * f.x;
              a[1] += w * f.y;
              a[2] += w * f.z;
            }
          }
        }
      }
    }
    __syncthreads();
  }
  const int rowlen = 3 * Tx;
  const size_t gstride = (size_t)3 * Mx * My * Mz;
  if constexpr (sizeof(T) == 4) {
    if (Tx % 4 == 0 && Mx % 4 == 0) {
      // tile rows (3 Tx values) start on 16-byte boundaries in the grid and in LDS: four
      // values per store
#pragma unroll
      for (int rq = 0; rq < RC; ++rq)
        for (int e = 4 * threadIdx.x; e < 3 * npts; e += 4 * blockDim.x) {
          const int r = e / rowlen;
          const int q = e - r * rowlen;
          const int lz = r / Ty;
          const int ly = r - lz * Ty;
          *(float4*)&grid[rq * gstride + (((size_t)(oz + lz) * My + (oy + ly)) * Mx + ox) * 3 + q] =
              *(const float4*)&acc[rq * 3 * npts + e];
        }
      return;
    }
  }
#pragma unroll
  for (int rq = 0; rq < RC; ++rq)
    for (int e = threadIdx.x; e < 3 * npts; e += blockDim.x) {
      const int r = e / rowlen;
      const int q = e - r * rowlen;
      const int lz = r / Ty;
      const int ly = r - lz * Ty;
      grid[rq * gstride + (((size_t)(oz + lz) * My + (oy + ly)) * Mx + ox) * 3 + q] = acc[rq * 3 * npts + e];
    }
}

inline unsigned grid_for(int64_t n, int threads) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > 1 << 20) b = 1 << 20;
  return (unsigned)b;
}

// Right-hand sides one spread launch accumulates together (the tiling budgets LDS for them).
int spread_rhs_per_pass(const ffcm_ctx* c) { return c->nrhs >= 4 ? 4 : (c->nrhs >= 2 ? 2 : 1); }

// Largest divisor of M not above `cap`.
int divisor_at_most(int M, int cap) {
  for (int d = cap < M ? cap : M; d >= 1; --d)
    if (M % d == 0) return d;
  return 1;
}

}  // namespace

// Tiles: per axis the largest divisor of M not above 16, shrunk on the widest axis until the
// accumulator and the staged contributors fit the LDS budget. A stencil starting in home tile h
// reaches tiles h .. h + span - 1 on each axis.
template <typename T>
bool plan_tiling(ffcm_ctx* c, int lds_budget) {
  Tiling& tl = c->tiling;
  const int RC = spread_rhs_per_pass(c);
  const int MG = c->host.M_G;
  // flat tiles: the lanes of a wave cover the (x, y) plane of a stencil and loop over z, so
  // clipping in z costs only skipped iterations while clipping in x, y idles lanes
  int target[3] = {16, 16, 8};
  if (const char* e = std::getenv("FFCM_SPREAD_TILE")) {  // tuning override "Tx,Ty,Tz"
    int t[3];
    if (std::sscanf(e, "%d,%d,%d", &t[0], &t[1], &t[2]) == 3 && t[0] > 0 && t[1] > 0 && t[2] > 0)
      for (int a = 0; a < 3; ++a) target[a] = t[a];
  }
  for (int a = 0; a < 3; ++a) tl.T[a] = divisor_at_most(c->host.M[a], target[a]);
  // per staged contributor: position and force (two T4), 3 * M_G table entries (two T each)
  // and the wrapped origin (four ints)
  const int per_cand = (4 + 4 * RC) * (int)sizeof(T) + 6 * MG * (int)sizeof(T) + 16;
  tl.cand_cap = 64;
  auto bytes = [&]() { return RC * 3 * tl.T[0] * tl.T[1] * tl.T[2] * (int)sizeof(T) + tl.cand_cap * per_cand; };
  while (bytes() > lds_budget) {
    if (tl.cand_cap > 32) {
      tl.cand_cap /= 2;
      continue;
    }
    int a = 2;
    for (int b = 1; b >= 0; --b)
      if (tl.T[b] > tl.T[a]) a = b;
    if (tl.T[a] == 1) return false;
    tl.T[a] = divisor_at_most(c->host.M[a], tl.T[a] - 1);
  }
  for (int a = 0; a < 3; ++a) {
    tl.nt[a] = c->host.M[a] / tl.T[a];
    const int kmax = (tl.T[a] - 1 + MG - 1) / tl.T[a];
    tl.span[a] = kmax + 1 < tl.nt[a] ? kmax + 1 : tl.nt[a];
  }
  tl.ntiles = tl.nt[0] * tl.nt[1] * tl.nt[2];
  tl.lds_bytes = bytes();
  return tl.span[0] * tl.span[1] * tl.span[2] <= kMaxHome;
}

template <typename T>
hipError_t launch_bin_tiles(ffcm_ctx* c, bool counted) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const int n = (int)c->N;
  hipError_t e;
  if (!counted) {
    e = launch_fill_i32(c->stream, c->d.tile_count, c->tiling.ntiles, 0);
    if (e != hipSuccess) return e;
    bin_count_kernel<T><<<grid_for(n, 256), 256, 0, c->stream>>>((const typename Vec<T>::v4*)c->d.Y_sorted, n, s,
                                                                   c->tiling, c->d.tile_count,
                                                                   (int2*)c->d.tile_cursor, c->d.anchor);
  }
  e = launch_exclusive_scan(c->stream, c->d.tile_count, c->d.tile_start, c->tiling.ntiles, c->d.scan_partials);
  if (e != hipSuccess) return e;
  bin_scatter_kernel<<<grid_for(n, 256), 256, 0, c->stream>>>((const int2*)c->d.tile_cursor, c->d.tile_start, n,
                                                               c->d.tile_list);
  c->tiles_valid = true;
  return hipGetLastError();
}

namespace {
template <typename T, int RC>
void spread_launch(ffcm_ctx* c, const typename Vec<T>::v4* Fs, T* grid) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const Tiling& tl = c->tiling;
  const typename Vec<T>::v4* Ys = (const typename Vec<T>::v4*)c->d.Y_sorted;
  if (s.MG == 8)
    spread_tile_kernel<T, 8, RC><<<tl.ntiles, kSpreadThreads, tl.lds_bytes, c->stream>>>(
        Ys, Fs, c->d.anchor, c->d.tile_start, c->d.tile_list, grid, s, tl);
  else
    spread_tile_kernel<T, 0, RC><<<tl.ntiles, kSpreadThreads, tl.lds_bytes, c->stream>>>(
        Ys, Fs, c->d.anchor, c->d.tile_start, c->d.tile_list, grid, s, tl);
}
}  // namespace

// R right-hand sides (forces Fs[r * N + slot], grids grid + r * 3 M^3), up to
// spread_rhs_per_pass per launch.
template <typename T>
hipError_t launch_spread(ffcm_ctx* c, const void* Fs, T* grid, int R) {
  if (!c->tiles_valid) {
    hipError_t e = launch_bin_tiles<T>(c);
    if (e != hipSuccess) return e;
  }
  if (!Fs) Fs = c->d.F_sorted;
  if (!grid) grid = (T*)c->d.force;
  using V4 = typename Vec<T>::v4;
  const int rcmax = spread_rhs_per_pass(c);
  for (int r0 = 0; r0 < R;) {
    const int left = R - r0;
    const int rc = (rcmax >= 4 && left >= 4) ? 4 : ((rcmax >= 2 && left >= 2) ? 2 : 1);
    const V4* F0 = (const V4*)Fs + (size_t)r0 * c->N;
    T* g0 = grid + (size_t)r0 * 3 * c->M3;
    if (rc == 4)
      spread_launch<T, 4>(c, F0, g0);
    else if (rc == 2)
      spread_launch<T, 2>(c, F0, g0);
    else
      spread_launch<T, 1>(c, F0, g0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    r0 += rc;
  }
  return hipSuccess;
}

template bool plan_tiling<float>(ffcm_ctx*, int);
template bool plan_tiling<double>(ffcm_ctx*, int);
template hipError_t launch_bin_tiles<float>(ffcm_ctx*, bool);
template hipError_t launch_bin_tiles<double>(ffcm_ctx*, bool);
template hipError_t launch_spread<float>(ffcm_ctx*, const void*, float*, int);
template hipError_t launch_spread<double>(ffcm_ctx*, const void*, double*, int);

}  // namespace ffcm
