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
#include <cstring>
#include <algorithm>

#include "group.hpp"
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

struct BinScatterArgs {
  const int2* slot_tile;
  const int* tile_start;
  int n;
  int* tile_list;
};

__device__ __forceinline__ void bin_scatter_body(const Blk& b, const BinScatterArgs& a) {
  for (int i = b.id * blockDim.x + threadIdx.x; i < a.n; i += b.n * blockDim.x) {
    const int2 tr = a.slot_tile[i];
    a.tile_list[a.tile_start[tr.x] + tr.y] = i;
  }
}
__global__ void bin_scatter_kernel(BinScatterArgs a) { bin_scatter_body(plain_blk(), a); }
__global__ void bin_scatter_kernel_group(const GroupEntry<BinScatterArgs>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const BinScatterArgs& a) { bin_scatter_body(b, a); });
}

// Arguments of both spread kernels.
template <typename T>
struct SpreadArgs {
  const typename Vec<T>::v4* Ys;
  const typename Vec<T>::v4* Fs;
  const int4* origin;
  const int* tile_start;
  const int* tile_list;
  T* grid;
  Scalars<T> s;
  Tiling tl;
  // matrix-core kernel: one tile counter per XCD (zeroed before the launch), or null for the
  // static tile order
  int* queue = nullptr;
};

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
__device__ __forceinline__ void spread_tile_body(const Blk& blk, const SpreadArgs<T>& a) {
  const typename Vec<T>::v4* __restrict__ Ys = a.Ys;
  const typename Vec<T>::v4* __restrict__ Fs = a.Fs;
  const int4* __restrict__ origin = a.origin;
  const int* __restrict__ tile_start = a.tile_start;
  const int* __restrict__ tile_list = a.tile_list;
  T* __restrict__ grid = a.grid;
  const Scalars<T>& s = a.s;
  const Tiling& tl = a.tl;
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
  const int t = blk.id;
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
    if constexpr (MGC == 0) {
      if (fast_xy && MG <= 16) {
        // The same screening for any M_G up to 16: lane p of step k (p = lane + 64 k, up to four
        // steps) covers stencil point (kx, ky) = (p mod M_G, p / M_G), computed once; points past
        // M_G^2 get a y offset no tile holds. (The per-contributor loop divided p by M_G and
        // wrapped both offsets at every point: 186,510 particles at M_G = 10 on 256^3 spread in
        // 1.00 ms.)
        const int steps = (MG * MG + 63) >> 6;
        int kxs[4], kys[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const int p = lane + 64 * k;
          kys[k] = p / MG;
          kxs[k] = p - kys[k] * MG;
          if (p >= MG * MG) kys[k] = 1 << 20;
        }
        for (int lz = wave; lz < Tz; lz += nwaves) {
          T* aplane = acc + lz * TyTx3;
          const int gz = oz + lz;
          for (int c0 = 0; c0 < n; c0 += 64) {
            const int c = c0 + lane;
            int dx = 0, dy = 0, kz = MG;
            if (c < n) {
              const int4 o = ((const int4*)org)[c];
              kz = gz - o.z;
              kz += kz < 0 ? Mz : 0;
              dx = o.x - ox;
              dx -= dx >= Tx ? Mx : 0;
              dy = o.y - oy;
              dy -= dy >= Ty ? My : 0;
            }
            unsigned long long mask = __ballot(kz < MG && dx > -MG && dy > -MG);
            while (mask) {
              const int l = __ffsll((long long)mask) - 1;
              mask &= mask - 1;
              const int cc = c0 + l;
              const int skz = __builtin_amdgcn_readlane(kz, l);
              const int sdx = __builtin_amdgcn_readlane(dx, l);
              const int sdy = __builtin_amdgcn_readlane(dy, l);
              const C2* tc = tab + (size_t)cc * 3 * MG;
              const C2 ez = tc[2 * MG + skz];
#pragma unroll
              for (int k = 0; k < 4; ++k) {
                if (k >= steps) break;
                const int lx = sdx + kxs[k], ly = sdy + kys[k];
                if ((unsigned)lx < (unsigned)Tx && (unsigned)ly < (unsigned)Ty) {
                  const C2 ex = tc[kxs[k]];
                  const C2 ey = tc[MG + kys[k]];
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

template <typename T, int MGC, int RC>
__global__ void __launch_bounds__(kSpreadThreads) spread_tile_kernel(SpreadArgs<T> a) {
  spread_tile_body<T, MGC, RC>(plain_blk(), a);
}
template <typename T, int MGC, int RC>
__global__ void __launch_bounds__(kSpreadThreads)
    spread_tile_kernel_group(const GroupEntry<SpreadArgs<T>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const SpreadArgs<T>& a) { spread_tile_body<T, MGC, RC>(b, a); });
}

// ------------------------------------------------------------------ matrix-core spreading
// Float32, one right-hand side: M_G = 8 on Tx x Ty x 8 tiles (8 <= Tx, Ty <= 16), or M_G from 9 to
// 16 on 16 x 16 x 8 tiles, on grids with at least T + 2 M_G points along x and y. The stencil weight factorises per axis into
//   w = (a0 + a2 (rx^2 + ry^2 + rz^2)) gx gy gz
//     = [gy gz (a0 + a2 rz^2) + gy ry^2 a2 gz] * gx  +  [gy gz] * a2 rx^2 gx,
// two products of a (y, z) factor and an x factor. One z plane of a tile, 16 rows (y) by
// 48 columns (16 x points times 3 components, the grid's own order; rows and columns past the
// tile's Ty and 3 Tx have zero factors and are not stored), therefore receives from
// its contributors the matrix product A B with A[y][2c + t] the (y, z) factor of term t of
// contributor c and B[2c + t][(x, comp)] its x factor times force component comp. Each wave
// owns one plane and accumulates the product in registers with v_mfma_f32_16x16x4f32 (two
// contributors per product, K = 4; three 16-column blocks): no LDS read-modify-write. Lane l
// supplies A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[4 (l / 16) + r][l % 16],
// r = 0..3, of each block.
typedef float mfma_f4 __attribute__((ext_vector_type(4)));
constexpr int kMfmaTile = 16;   // rows and x points of one plane's product (tile extents up to 16)
constexpr int kMfmaPlanes = 8;  // tile planes along z, one per wave
constexpr int kMfmaCols = 3 * kMfmaTile;
constexpr int kMfmaWaves = kSpreadThreads / 64;
constexpr int kXcds = 8;  // L2 domains; workgroups are dispatched to them round-robin
// LDS per contributor: the factor tables (x: 2 x 48 values; y: 16 pairs (gy, gy ry^2); z: per
// stencil entry and term the pair that multiplies the y pair), plus position, force and origin
// offsets and one entry of each wave's plane list. One more all-zero set of factor tables stands
// in for missing contributors, and each plane list has 4 entries of padding.
// kWide: M_G from 9 to 16 on 16 x 16 x 8 tiles: 16 z entries per contributor instead of 8, and
// twelve home tiles (three along z)
template <bool kWide>
constexpr int kMfmaZEntries = kWide ? 16 : 8;
template <bool kWide>
constexpr int kMfmaTableBytes = (2 * kMfmaCols + 2 * kMfmaTile + 2 * 2 * kMfmaZEntries<kWide>) * 4;
template <bool kWide>
constexpr int kMfmaCandBytes = kMfmaTableBytes<kWide> + 3 * 16 + kMfmaWaves * 4;
template <bool kWide>
constexpr int kMfmaFixedBytes = kMfmaTableBytes<kWide> + kMfmaWaves * 4 * 4;

// Candidates of a tile: the slots binned to its home tiles (the tile and its predecessors within
// the stencil span, k = dx + 2 dy + 4 dz, up to eight, or twelve with three along z), in home order. Lane k < nhome of a wave
// loads home k's list bounds; home_map turns them into the wave-uniform prefix of the home counts
// and, per thread, the list position of candidate q.
template <int K>
struct HomeMap {
  int pre[K];   // candidates before home k
  int step[K];  // list position of candidate q: q + sum of step[k] over the homes k with pre[k] <= q
  int total;
};

__device__ __forceinline__ int2 load_home_bounds(const int* __restrict__ tile_start, const Tiling& tl, int t,
                                                 int nhome, int lane) {
  if (t >= tl.ntiles || lane >= nhome) return make_int2(0, 0);
  const int ntx = tl.nt[0], nty = tl.nt[1], ntz = tl.nt[2];
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  int hx = tx - (lane & 1), hy = ty - ((lane >> 1) & 1), hz = tz - (lane >> 2);
  hx += hx < 0 ? ntx : 0;
  hy += hy < 0 ? nty : 0;
  hz += hz < 0 ? ntz : 0;
  const int h = hx + ntx * (hy + nty * hz);
  return make_int2(tile_start[h], tile_start[h + 1]);
}

template <int K>
__device__ __forceinline__ HomeMap<K> home_map(int2 b) {
  HomeMap<K> m;
  int acc = 0, prev = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int a = __builtin_amdgcn_readlane(b.x, k);
    const int e = __builtin_amdgcn_readlane(b.y, k);
    m.pre[k] = acc;
    m.step[k] = (a - acc) - prev;
    prev = a - acc;
    acc += e - a;
  }
  m.total = acc;
  return m;
}

// (a sum of masked steps, not a select chain: the latter is turned into an indexed table)
template <int K>
__device__ __forceinline__ int candidate_pos(const HomeMap<K>& m, int q) {
  int pos = q + m.step[0];
#pragma unroll
  for (int k = 1; k < K; ++k) pos += q >= m.pre[k] ? m.step[k] : 0;
  return pos;
}

// Tile i of XCD x's queue: the tiles of the rows with (ty + tz) mod 8 = x, layer by layer, rows
// in order, x fastest (ntiles past the end). Each XCD's share is then spread over the whole
// volume (the XCDs' loads differ by under 1% on the clustered 2^20-particle suspension, against
// 15% for contiguous rows), and neighbouring tiles of a row, which share partial cache lines of
// the grid, are drawn by blocks of the same L2 at about the same time. Needs nty % 8 == 0.
__device__ __forceinline__ int queue_tile(int i, int x, const Tiling& tl) {
  const int ntx = tl.nt[0], nty = tl.nt[1];
  const int rows = nty >> 3;  // rows of one layer per XCD
  const int per_layer = rows * ntx;
  const int tz = i / per_layer;
  if (tz >= tl.nt[2]) return tl.ntiles;
  const int r = i - tz * per_layer;
  const int j = r / ntx;
  const int tx = r - j * ntx;
  const int ty = ((x - tz) & 7) + 8 * j;
  return tx + ntx * (ty + nty * tz);
}

// Persistent blocks: block b spreads tiles t = b', b' + G, ... (b' the XCD-aware remap of b), or,
// with a queue, the tiles it draws one at a time from its XCD's queue (the blocks of an XCD then
// finish together however unevenly the tiles' candidates are spread; a static order left the
// busiest block with 1.3 times the mean candidate count at 2^20 clustered particles).
// The chain home bounds -> slot -> (position, force, stencil origin) is three dependent loads;
// it runs three tiles ahead: while tile t is accumulated, the records of the next tile, the
// slots of the one after and the home bounds of the third are in flight, and the previous
// tile's output is being stored.
// Thread q holds the record of candidate q; the candidates whose stencils reach the tile are
// compacted (block-wide ballot) and staged to LDS in batches of cap. Candidates past the block's
// thread count are loaded in further rounds.
// kPartial: tile extents below 16 on x or y (positions past them masked); otherwise 16 x 16.
// kWide: M_G from 9 to 16 (16 x 16 tiles only: then two home tiles along x and y, up to three
// along z).
template <bool kPartial, bool kWide>
__device__ __forceinline__ void spread_mfma_body(const Blk& blk, const SpreadArgs<float>& a) {
  static_assert(!(kPartial && kWide), "wide stencils take full 16 x 16 tiles");
  constexpr int kHomes = kWide ? 12 : 8;
  constexpr int kZE = kMfmaZEntries<kWide>;  // z entries per contributor (plane-list code c kZE + kz)
  const int MG = kWide ? a.s.MG : 8;
  const float4* __restrict__ Ys = a.Ys;
  const float4* __restrict__ Fs = a.Fs;
  const int4* __restrict__ origin = a.origin;
  const int* __restrict__ tile_start = a.tile_start;
  const int* __restrict__ tile_list = a.tile_list;
  float* __restrict__ grid = a.grid;
  const Scalars<float>& s = a.s;
  const Tiling& tl = a.tl;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int wave_kept[kMfmaWaves];
  const int cap = tl.cand_cap;
  // factor tables of cap + 1 contributors (the last one all zero)
  float* bx = (float*)smem_raw;                                    // x: term 0 (gx F), term 1 (a2 rx^2 gx F)
  float2* ay = (float2*)(bx + (size_t)(cap + 1) * 2 * kMfmaCols);  // y: (gy, gy ry^2) per tile row
  float2* zf = ay + (size_t)(cap + 1) * kMfmaTile;  // z: [entry][term] (gz (a0 + a2 rz^2), a2 gz), (gz, 0)
  float4* cy = (float4*)(zf + (size_t)(cap + 1) * 2 * kZE);  // cap positions
  float4* cf = cy + cap;                                // cap forces
  int4* org = (int4*)(cf + cap);  // cap: x, y stencil origin in the tile, wrapped z origin
  int* plist = (int*)(org + cap) + (threadIdx.x >> 6) * (cap + 4);  // this wave's plane list
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int row = lane & 15;  // A row (y) and B column within a block
  const int sel = lane >> 5;  // which contributor of a product
  const int term = (lane >> 4) & 1;
  const int ntx = tl.nt[0], nty = tl.nt[1], ntiles = tl.ntiles;
  const int Tx = kPartial ? tl.T[0] : kMfmaTile, Ty = kPartial ? tl.T[1] : kMfmaTile;
  const int nhome = tl.span[0] * tl.span[1] * tl.span[2];
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  const int G = blk.n;
  if (tid < 2 * kMfmaCols) bx[cap * 2 * kMfmaCols + tid] = 0.f;
  if (tid < kMfmaTile) ay[cap * kMfmaTile + tid] = make_float2(0.f, 0.f);
  if (tid < 2 * kZE) zf[cap * 2 * kZE + tid] = make_float2(0.f, 0.f);
  // slot of candidate q (entry 0, a valid slot, stands in past the end: every load is
  // unconditional and in bounds)
  auto slot_of = [&](const HomeMap<kHomes>& m, int q) { return tile_list[q < m.total ? candidate_pos(m, q) : 0]; };
  // the tile being spread and the next three (records, slots and home bounds in flight)
  int t, t1, t2, t3;
  __shared__ int drawn[2];  // per-tile draws, by parity
  int* const queue = a.queue ? a.queue + blk.id % kXcds : nullptr;
  const int xcd = blk.id % kXcds;
  const int bpx = blk.n / kXcds;  // blocks per XCD (the launch is a multiple of kXcds)
  if (queue) {
    // the first four tiles without atomics: entries b, b + bpx, ... of the XCD's order; the
    // queue counter then numbers the entries from 4 bpx on
    const int b = blk.id / kXcds;
    t = queue_tile(b, xcd, tl), t1 = queue_tile(b + bpx, xcd, tl);
    t2 = queue_tile(b + 2 * bpx, xcd, tl), t3 = queue_tile(b + 3 * bpx, xcd, tl);
  } else {
    t = xcd_first_tile(blk.id, blk.n);
    t1 = t + G, t2 = t + 2 * G, t3 = t + 3 * G;
  }
  int parity = 0;
  const HomeMap<kHomes> m0 = home_map<kHomes>(load_home_bounds(tile_start, tl, t, nhome, lane));
  const HomeMap<kHomes> m1 = home_map<kHomes>(load_home_bounds(tile_start, tl, t1, nhome, lane));
  int2 hb2 = load_home_bounds(tile_start, tl, t2, nhome, lane);
  const int slot0 = slot_of(m0, tid);
  int slot1 = slot_of(m1, tid);
  int total0 = m0.total, total1 = m1.total;
  float4 ry = Ys[slot0], rf = Fs[slot0];
  int4 ro = origin[slot0];
  // previous tile's plane, stored while the next one is accumulated
  mfma_f4 out0 = {0.f, 0.f, 0.f, 0.f}, out1 = out0, out2 = out0;
  float* out_dst = nullptr;
  int out_rows = 0;  // rows of this lane's four that lie in the tile
  const size_t rs = (size_t)Mx * 3;
  const int cols = 3 * Tx;
  auto store_plane = [&]() {
#pragma unroll
    for (int r = 0; r < 4; ++r)
      if (!kPartial || r < out_rows) {
        if (!kPartial || row < cols) out_dst[r * rs + 0] = out0[r];
        if (!kPartial || row + 16 < cols) out_dst[r * rs + 16] = out1[r];
        if (!kPartial || row + 32 < cols) out_dst[r * rs + 32] = out2[r];
      }
  };
  while (t < ntiles) {
    vm_wait_all();
    if (out_dst) store_plane();
    // in flight during this tile: records of t1, slots of t2, home bounds of t3 and the draw of
    // the tile after t3 (read at the end of the tile)
    const float4 ny = Ys[slot1], nf = Fs[slot1];
    const int4 no = origin[slot1];
    const HomeMap<kHomes> m2 = home_map<kHomes>(hb2);
    const int slot2 = slot_of(m2, tid);
    const int2 hb3 = load_home_bounds(tile_start, tl, t3, nhome, lane);
    int next_draw = 0;
    if (queue && tid == 0) next_draw = atomicAdd(queue, 1);
    const int tx = t % ntx;
    const int ty = (t / ntx) % nty;
    const int tz = t / (ntx * nty);
    const int ox = tx * Tx, oy = ty * Ty, oz = tz * kMfmaPlanes;
    const int gz = oz + wave;  // this wave's plane
    mfma_f4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = acc0, acc2 = acc0;
    // one round of up to kSpreadThreads candidates, candidate q0 + tid's record in (y, f, o)
    auto round = [&](const float4& y, const float4& f, const int4& o, int q0) {
      // (1) keep the candidates whose stencils reach the tile; survivor index by block-wide ballot
      const bool keep = q0 + tid < total0 && cyclic_overlap(o.x, MG, ox, Tx, Mx) &&
                        cyclic_overlap(o.y, MG, oy, Ty, My) && cyclic_overlap(o.z, MG, oz, kMfmaPlanes, Mz);
      const unsigned long long km = __ballot(keep);
      if (lane == 0) wave_kept[wave] = __popcll(km);
      __syncthreads();
      int ns = 0, before = 0;
#pragma unroll
      for (int w = 0; w < kMfmaWaves; ++w) {
        const int c = wave_kept[w];
        before += w < wave ? c : 0;
        ns += c;
      }
      const int si = before + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(km >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((unsigned)km, 0u));
      // the grid holds at least T + 2 M_G points on x and y, so a reaching stencil's points in
      // the tile are consecutive: one offset per axis in (-M_G, T)
      int dx = o.x - ox;
      dx -= dx >= Tx ? Mx : 0;
      int dy = o.y - oy;
      dy -= dy >= Ty ? My : 0;
      for (int b0 = 0; b0 < ns; b0 += cap) {
        const int nb = min(cap, ns - b0);
        if (keep && si >= b0 && si < b0 + nb) {
          const int j = si - b0;
          cy[j] = y;
          cf[j] = f;
          org[j] = make_int4(dx, dy, o.z, 0);
        }
        __syncthreads();
        // (2) factor tables, one uniform loop over (contributor, j). Tile positions j and j + 8
        // are 8 apart, so at most one of them lies in a stencil's 8 points: one entry per pair of
        // positions, the other written as zero (as are positions past the tile's extent)
        // (one loop for the three axes, each contributor's position and origin read once; the
        // Gaussians as hardware base-2 exponentials, as in the interpolation: 512^3 spread 1.02 ->
        // 0.99 ms)
        if constexpr (kWide) {
          // wide stencils: every tile position j of x and y its own entry (j and j + 8 can both
          // lie in a stencil of more than 8 points), z entries kz = j < M_G, the rest zero
          const float c2 = s.inv_2S2 * 1.4426950408889634f;
          for (int e = tid; e < nb * 16; e += kSpreadThreads) {
            const int c = e >> 4, j = e & 15;
            const float4 y = cy[c];
            const int4 o = org[c];
            {
              const int k = j - o.x;
              const bool in = (unsigned)k < (unsigned)MG;
              const AxisEntry<float> en = axis_entry_exp2((int)rint(y.x * s.inv_h) - s.half, in ? k : 0, y.x, s.h,
                                                          s.inv_norm, c2);
              const float g = in ? en.g : 0.f;
              const float g1 = s.a2 * en.r2 * g;
              const float4 f = cf[c];
              float* b1 = bx + c * 2 * kMfmaCols + 3 * j;
              b1[0] = g * f.x;
              b1[1] = g * f.y;
              b1[2] = g * f.z;
              b1[kMfmaCols + 0] = g1 * f.x;
              b1[kMfmaCols + 1] = g1 * f.y;
              b1[kMfmaCols + 2] = g1 * f.z;
            }
            {
              const int k = j - o.y;
              const bool in = (unsigned)k < (unsigned)MG;
              const AxisEntry<float> en = axis_entry_exp2((int)rint(y.y * s.inv_h) - s.half, in ? k : 0, y.y, s.h,
                                                          s.inv_norm, c2);
              ay[c * kMfmaTile + j] = in ? make_float2(en.g, en.g * en.r2) : make_float2(0.f, 0.f);
            }
            {
              const bool in = j < MG;
              const AxisEntry<float> en =
                  axis_entry_exp2((int)rint(y.z * s.inv_h) - s.half, in ? j : 0, y.z, s.h, s.inv_norm, c2);
              *(float4*)&zf[2 * e] = in ? make_float4(en.g * (s.a0 + s.a2 * en.r2), s.a2 * en.g, en.g, 0.f)
                                        : make_float4(0.f, 0.f, 0.f, 0.f);
            }
          }
        } else {
          const float c2 = s.inv_2S2 * 1.4426950408889634f;  // log2(e) / 2 Sigma^2
          for (int e = tid; e < nb * 8; e += kSpreadThreads) {
            const int c = e >> 3, j = e & 7;
            const float4 y = cy[c];
            const int4 o = org[c];
            // x factors
            {
              const int k1 = j - o.x, k2 = k1 + 8;
              const bool in1 = (unsigned)k1 < 8u && j < Tx, in2 = (unsigned)k2 < 8u && j + 8 < Tx;
              const AxisEntry<float> en = axis_entry_exp2((int)rint(y.x * s.inv_h) - s.half,
                                                          in1 ? k1 : (in2 ? k2 : 0), y.x, s.h, s.inv_norm, c2);
              const float g = (in1 || in2) ? en.g : 0.f;
              const float g1 = s.a2 * en.r2 * g;
              const float4 f = cf[c];
              const float v[6] = {g * f.x, g * f.y, g * f.z, g1 * f.x, g1 * f.y, g1 * f.z};
              float* b1 = bx + c * 2 * kMfmaCols + 3 * j;
              float* b2 = b1 + 24;
#pragma unroll
              for (int i = 0; i < 3; ++i) {
                b1[i] = in1 ? v[i] : 0.f;
                b1[kMfmaCols + i] = in1 ? v[3 + i] : 0.f;
                b2[i] = in2 ? v[i] : 0.f;
                b2[kMfmaCols + i] = in2 ? v[3 + i] : 0.f;
              }
            }
            // y factors
            {
              const int k1 = j - o.y, k2 = k1 + 8;
              const bool in1 = (unsigned)k1 < 8u && j < Ty, in2 = (unsigned)k2 < 8u && j + 8 < Ty;
              const AxisEntry<float> en = axis_entry_exp2((int)rint(y.y * s.inv_h) - s.half,
                                                          in1 ? k1 : (in2 ? k2 : 0), y.y, s.h, s.inv_norm, c2);
              const float2 v = make_float2(en.g, en.g * en.r2);
              const float2 z = make_float2(0.f, 0.f);
              ay[c * kMfmaTile + j] = in1 ? v : z;
              ay[c * kMfmaTile + j + 8] = in2 ? v : z;
            }
            // z factors (entry kz = j)
            {
              const AxisEntry<float> en =
                  axis_entry_exp2((int)rint(y.z * s.inv_h) - s.half, j, y.z, s.h, s.inv_norm, c2);
              *(float4*)&zf[2 * e] = make_float4(en.g * (s.a0 + s.a2 * en.r2), s.a2 * en.g, en.g, 0.f);
            }
          }
        }
        // (3) this wave's plane list: the contributors that reach plane gz, packed as
        // kZE c + stencil entry, in steps of four stored as (0, 2, 1, 3) so that half `sel` of the wave
        // reads the contributors of both its products with one 8-byte load; padded with the zero
        // contributor
        int cnt = 0;
        for (int c0 = 0; c0 < nb; c0 += 64) {
          const int c = c0 + lane;
          int kz = MG;
          if (c < nb) {
            kz = gz - org[c].z;
            kz += kz < 0 ? Mz : 0;
          }
          const bool hit = kz < MG;
          const unsigned long long m = __ballot(hit);
          if (hit) {
            const int i = cnt + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                                __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            plist[(i & ~3) | ((i & 1) << 1) | ((i >> 1) & 1)] = c * kZE + kz;
          }
          cnt += __popcll(m);
        }
        const int padded = (cnt + 3) & ~3;
        if (lane < padded - cnt) {
          const int i = cnt + lane;
          plist[(i & ~3) | ((i & 1) << 1) | ((i >> 1) & 1)] = cap * kZE;
        }
        __syncthreads();
        // (4) accumulate: two products (four contributors) per step
        for (int st = 0; st < padded; st += 4) {
          const int2 e = *(const int2*)&plist[st + 2 * sel];
          const unsigned ca = (unsigned)e.x / kZE, cb = (unsigned)e.y / kZE;
          const float* pa = bx + ca * (2 * kMfmaCols) + term * kMfmaCols + row;
          const float* pb = bx + cb * (2 * kMfmaCols) + term * kMfmaCols + row;
          const float a0 = pa[0], a1 = pa[16], a2 = pa[32];
          const float b0v = pb[0], b1v = pb[16], b2v = pb[32];
          const float2 ya = ay[ca * kMfmaTile + row], yb = ay[cb * kMfmaTile + row];
          const float2 za = zf[2 * e.x + term], zb = zf[2 * e.y + term];
          const float fa = ya.x * za.x + ya.y * za.y;
          const float fb = yb.x * zb.x + yb.y * zb.y;
          acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(fa, a0, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(fa, a1, acc1, 0, 0, 0);
          acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(fa, a2, acc2, 0, 0, 0);
          acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(fb, b0v, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(fb, b1v, acc1, 0, 0, 0);
          acc2 = __builtin_amdgcn_mfma_f32_16x16x4f32(fb, b2v, acc2, 0, 0, 0);
        }
        __syncthreads();
      }
      __syncthreads();  // wave_kept is rewritten by the next round
    };
    // the first round from the prefetched records (a separate call, so that no path merges
    // them with loads still in flight), further rounds for dense tiles load their own
    if (total0 > 0) round(ry, rf, ro, 0);
    for (int q0 = kSpreadThreads; q0 < total0; q0 += kSpreadThreads) {
      const HomeMap<kHomes> mt = home_map<kHomes>(load_home_bounds(tile_start, tl, t, nhome, lane));
      const int sl = slot_of(mt, q0 + tid);
      round(Ys[sl], Fs[sl], origin[sl], q0);
    }
    // rows 4 (lane / 16) + r of the plane, column lane % 16 of each 16-column block
    out0 = acc0;
    out1 = acc1;
    out2 = acc2;
    out_dst = grid + (((size_t)gz * My + oy + 4 * (lane >> 4)) * Mx + ox) * 3 + row;
    out_rows = Ty - 4 * (lane >> 4);
    int t4;
    if (queue) {
      // a slot is rewritten two tiles later, after this barrier and the next tile's
      if (tid == 0) drawn[parity] = queue_tile(4 * bpx + next_draw, xcd, tl);
      __syncthreads();
      t4 = drawn[parity];
      parity ^= 1;
    } else {
      t4 = t3 + G;
    }
    t = t1, t1 = t2, t2 = t3, t3 = t4;
    ry = ny;
    rf = nf;
    ro = no;
    total0 = total1;
    slot1 = slot2;
    total1 = m2.total;
    hb2 = hb3;
  }
  if (out_dst) store_plane();
}

template <bool kPartial, bool kWide>
__global__ void __launch_bounds__(kSpreadThreads) spread_mfma_kernel(SpreadArgs<float> a) {
  spread_mfma_body<kPartial, kWide>(plain_blk(), a);
}
template <bool kPartial, bool kWide>
__global__ void __launch_bounds__(kSpreadThreads)
    spread_mfma_kernel_group(const GroupEntry<SpreadArgs<float>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const SpreadArgs<float>& a) { spread_mfma_body<kPartial, kWide>(b, a); });
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
  // 16 slots per tile (2^20 particles on 512^3): cell-list stage 0.243 -> 0.205 ms; at 8
  // (65,536 on 256^3) the wave's tiles are many and the per-tile loop lost 0.012 ms
  tl.wave_ranks = c->N >= 12 * (int64_t)tl.ntiles;
  // matrix-core spreading: 96 contributors per LDS batch (about 70 KiB, two blocks per CU);
  // FFCM_SPREAD=lds keeps the LDS accumulation, FFCM_SPREAD_CAP sets the batch
  tl.mfma = 0;
  tl.blocks = 0;
  const char* mode = std::getenv("FFCM_SPREAD");
  // M_G = 8 on Tx x Ty x 8 tiles with 8 <= Tx, Ty <= 16, or (wide) 9 <= M_G <= 16 on 16 x 16 x 8
  // tiles; grids of at least T + 2 M_G points along x and y
  const bool wide = MG > 8 && MG <= 16 && tl.T[0] == kMfmaTile && tl.T[1] == kMfmaTile;
  if (sizeof(T) == 4 && (MG == 8 || wide) && RC == 1 && tl.T[0] >= 8 && tl.T[0] <= kMfmaTile && tl.T[1] >= 8 &&
      tl.T[1] <= kMfmaTile && tl.T[2] == kMfmaPlanes && c->host.M[0] >= tl.T[0] + 2 * MG &&
      c->host.M[1] >= tl.T[1] + 2 * MG && c->host.M[2] >= MG &&
      !(mode && std::strcmp(mode, "lds") == 0)) {
    // (wide: 88 contributors of 848 bytes keep two blocks on a CU)
    int mcap = wide ? 88 : 96;
    // (a multiple of 4: the plane lists are read in 8-byte pairs)
    if (const char* e = std::getenv("FFCM_SPREAD_CAP")) mcap = std::max(8, std::atoi(e)) & ~3;
    const int cand_bytes = wide ? kMfmaCandBytes<true> : kMfmaCandBytes<false>;
    const int fixed_bytes = wide ? kMfmaFixedBytes<true> : kMfmaFixedBytes<false>;
    const int lds_max = c->lds_limit - 1024;
    while (mcap > 8 && mcap * cand_bytes + fixed_bytes > lds_max) mcap /= 2;
    int per_cu = 0, cus = 256;
    const int lds = mcap * cand_bytes + fixed_bytes;
    const void* kern = wide ? (const void*)spread_mfma_kernel<false, true>
                       : tl.T[0] == kMfmaTile && tl.T[1] == kMfmaTile ? (const void*)spread_mfma_kernel<false, false>
                                                                      : (const void*)spread_mfma_kernel<true, false>;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kSpreadThreads, lds) != hipSuccess)
      per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) cus = 256;
    (void)hipGetLastError();
    if (lds <= lds_max && per_cu > 0) {
      tl.mfma = wide ? 2 : 1;
      tl.cand_cap = mcap;
      tl.lds_bytes = lds;
      tl.blocks = std::min(tl.ntiles, per_cu * cus);
    }
  }
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
  e = launch_exclusive_scan(c, kTagBinScan, c->d.tile_count, c->d.tile_start, c->tiling.ntiles, c->d.scan_partials);
  if (e != hipSuccess) return e;
  c->tiles_valid = true;
  return submit(c, kTagBinScatter, bin_scatter_kernel, bin_scatter_kernel_group, grid_for(n, 256), 256, 0,
                BinScatterArgs{(const int2*)c->d.tile_cursor, c->d.tile_start, n, c->d.tile_list});
}

namespace {
template <typename T, int RC>
hipError_t spread_launch(ffcm_ctx* c, const typename Vec<T>::v4* Fs, T* grid, int tag) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const Tiling& tl = c->tiling;
  const typename Vec<T>::v4* Ys = (const typename Vec<T>::v4*)c->d.Y_sorted;
  SpreadArgs<T> a{Ys, Fs, c->d.anchor, c->d.tile_start, c->d.tile_list, grid, s, tl};
  if constexpr (sizeof(T) == 4 && RC == 1) {
    if (tl.mfma) {
      // per-XCD tile queues for a plain launch (recorded batch launches keep the static order)
      // from 32 tiles per block on: with fewer, a block's draws are not hidden behind its tiles
      // (65,536 particles on 256^3, 16 tiles per block: 0.134 ms against 0.096 ms static).
      // FFCM_SPREAD_ORDER=queue or static overrides.
      const char* order = std::getenv("FFCM_SPREAD_ORDER");
      bool use_queue = tl.ntiles >= 32 * tl.blocks;
      if (order) use_queue = std::strcmp(order, "queue") == 0;
      if (use_queue && !c->rec && c->d.spread_queue && tl.nt[1] % 8 == 0 && tl.blocks % kXcds == 0) {
        hipError_t e = launch_fill_i32(c->stream, c->d.spread_queue, kXcds, 0);
        if (e != hipSuccess) return e;
        a.queue = c->d.spread_queue;
      }
      // persistent: tl.blocks resident blocks over tl.ntiles tiles
      if (tl.mfma == 2)
        return submit(c, tag, spread_mfma_kernel<false, true>, spread_mfma_kernel_group<false, true>, tl.blocks,
                      kSpreadThreads, tl.lds_bytes, a, tl.ntiles);
      if (tl.T[0] == kMfmaTile && tl.T[1] == kMfmaTile)
        return submit(c, tag, spread_mfma_kernel<false, false>, spread_mfma_kernel_group<false, false>, tl.blocks,
                      kSpreadThreads, tl.lds_bytes, a, tl.ntiles);
      return submit(c, tag, spread_mfma_kernel<true, false>, spread_mfma_kernel_group<true, false>, tl.blocks,
                    kSpreadThreads, tl.lds_bytes, a, tl.ntiles);
    }
  }
  if (s.MG == 8)
    return submit(c, tag, spread_tile_kernel<T, 8, RC>, spread_tile_kernel_group<T, 8, RC>, tl.ntiles,
                  kSpreadThreads, tl.lds_bytes, a);
  return submit(c, tag, spread_tile_kernel<T, 0, RC>, spread_tile_kernel_group<T, 0, RC>, tl.ntiles, kSpreadThreads,
                tl.lds_bytes, a);
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
    const int tag = kTagSpread + r0;
    const hipError_t e = rc == 4 ? spread_launch<T, 4>(c, F0, g0, tag)
                                 : (rc == 2 ? spread_launch<T, 2>(c, F0, g0, tag) : spread_launch<T, 1>(c, F0, g0, tag));
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
