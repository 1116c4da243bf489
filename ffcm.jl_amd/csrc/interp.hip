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
// keeps the rows its neighbours reuse. The wave sum is a row-level DPP reduction.
#include <cstdlib>
#include <cstring>

#include "group.hpp"
#include "stencil.hpp"

namespace ffcm {

namespace {

constexpr int kInterpThreads = 256;
constexpr int kXcds = 8;

// Arguments of both interpolation kernels (the wave kernel reads the slot order from order, the
// column kernel the tile lists from tile_start / order, tl and rp).
template <typename T>
struct InterpArgs {
  const typename Vec<T>::v4* Ys;
  const int* perm;
  const int* tile_start;
  const int* order;
  const T* grid;
  T* V;
  Scalars<T> s;
  Tiling tl;
  RingPlan rp;
  // gradient mode of the DMA ring kernel (the pullback's grid part): weight vectors by slot, and
  // whether V (then dL/dY) is added to rather than written
  const typename Vec<T>::v4* W = nullptr;
  int accumulate = 0;
};

template <typename T, int MGC>
__device__ __forceinline__ void interp_body(const Blk& blk, const InterpArgs<T>& a) {
  const typename Vec<T>::v4* __restrict__ Ys = a.Ys;
  const int* __restrict__ perm = a.perm;
  const int* __restrict__ order = a.order;
  const T* __restrict__ grid = a.grid;
  T* __restrict__ V = a.V;
  const Scalars<T>& s = a.s;
  const int MG = MGC > 0 ? MGC : s.MG;
  const int lane = threadIdx.x & 63;
  const int waves_per_block = blockDim.x >> 6;
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  // the block count is a multiple of kXcds: a bijection of [0, blk.n)
  const int per_xcd = blk.n / kXcds;
  const int lblock = (blk.id % kXcds) * per_xcd + blk.id / kXcds;
  for (int i = lblock * waves_per_block + (threadIdx.x >> 6); i < s.N; i += blk.n * waves_per_block) {
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
    if constexpr (MGC > 0) {
      // the (kx, ky) plane in ceil(MG^2 / 64) passes of point p = lane + 64 pass: all MG z-loads
      // of a lane are issued before any is used, and the z entries (wave-uniform index) are
      // scalar register reads
      constexpr int kPasses = (MGC * MGC + 63) / 64;
#pragma unroll
      for (int ps = 0; ps < kPasses; ++ps) {
        const int p = lane + 64 * ps;
        const bool valid = p < MGC * MGC;
        const int ky = valid ? p / MGC : 0;
        const int kx = valid ? p - ky * MGC : 0;
        const int ix = wrap_add(ax.low, kx, Mx);
        const int iy = wrap_add(ay.low, ky, My);
        // this lane's own x and y entries (no cross-lane permute)
        const AxisEntry<T> exk = axis_entry(ax.lo, kx, y.x, s.h, s.inv_norm, s.inv_2S2);
        const AxisEntry<T> eyk = axis_entry(ay.lo, ky, y.y, s.h, s.inv_norm, s.inv_2S2);
        const T gx = exk.g, r2x = exk.r2;
        const T gy = eyk.g, r2y = eyk.r2;
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
    v0 = wave_total(v0);
    v1 = wave_total(v1);
    v2 = wave_total(v2);
    if (lane == 0) {
      const int n = perm[slot];
      V[3 * n + 0] = s.h3 * v0;
      V[3 * n + 1] = s.h3 * v1;
      V[3 * n + 2] = s.h3 * v2;
    }
  }
}


// ---- column-streamed interpolation ----
//
// One block owns one column of 16 x 16 x 8 spread tiles (x, y) over a run of tile layers in z
// and visits the particles binned to those tiles layer by layer. The grid points they read,
// the tile plus a 7-point halo on each axis, sit in an LDS ring of 15 z planes of 23 x 23
// points. While layer j is interpolated from the ring, every thread holds in registers its
// share of the 8 planes that layers j + 1 and j + 2 add (two layers of loads in flight); after
// the layer it stores those of layer j + 1 over the planes layer j was the last to need. Each
// plane of the column is read from memory once, in rows of 23 consecutive points.
//
// Half a wave per particle: lane (kx, kq) sums the stencil points (kx, 2 kq .. 2 kq + 1, 0 .. 7)
// from two ring rows at once, so that a layer of a few tens of particles keeps every wave of
// the block busy; per-particle work is three exponentials per lane plus row broadcasts of the
// z entries, and the half's sum is five row-level moves.

constexpr int kRingStaged = 512;  // particles of a step whose positions are staged in LDS
constexpr int kRingMaxLayers = 256;
constexpr int kColumnLayers = 2;  // tile layers per step (a 23-plane ring: 146 KB of LDS)

// Velocity of one particle from a ring of grid planes in LDS ([RZ][PY][PX][3], the plane of
// run-relative z P in slot P mod RZ), half a wave per particle: lane (kx, kq) of the half sums
// the stencil points (kx, 2 kq .. 2 kq + 1, 0 .. 7). Sums complete in lane 31 of the half:
//   sum_{ky} gx gy(ky) sum_{kz} (R(ky) gz(kz) + a2 r2z(kz) gz(kz)) u(kx, ky, kz)
// with R(ky) = a0 + a2 (r2x + r2y(ky)), i.e. the spreading weight
// (a0 + a2 (r2x + r2y + r2z)) gx gy gz. (ox, oy, oz0): grid point of the ring's origin.
// Positions are wrapped into [0, L) (Y_sorted), and the Gaussians are base-2 exponentials.
template <int PX, int PLANE, int RZ>
__device__ __forceinline__ void ring_half_velocity(const float* ring, const Vec<float>::v4& y, const Scalars<float>& s,
                                                   int ox, int oy, int oz0, int kx, int kq, float& v0, float& v1,
                                                   float& v2) {
  using T = float;
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  const AxisOrigin ax = axis_origin_wrapped(y.x, s.inv_h, s.half, Mx);
  const AxisOrigin ay = axis_origin_wrapped(y.y, s.inv_h, s.half, My);
  const AxisOrigin az = axis_origin_wrapped(y.z, s.inv_h, s.half, Mz);
  const T c2 = s.inv_2S2 * 1.4426950408889634f;  // log2(e) / 2 Sigma^2
  const AxisEntry<T> ex = axis_entry_exp2(ax.lo, kx, y.x, s.h, s.inv_norm, c2);
  const AxisEntry<T> ey0 = axis_entry_exp2(ay.lo, 2 * kq, y.y, s.h, s.inv_norm, c2);
  const AxisEntry<T> ey1 = axis_entry_exp2(ay.lo, 2 * kq + 1, y.y, s.h, s.inv_norm, c2);
  // z entry kz evaluated in lane kx = kz of each row, then broadcast along the row
  const AxisEntry<T> ez = axis_entry_exp2(az.lo, kx, y.z, s.h, s.inv_norm, c2);
  const T Pl = ez.g, Ql = s.a2 * ez.r2 * ez.g;
  T P[8], Q[8];
  P[0] = dpp_move<0x150, 0xF>(Pl), Q[0] = dpp_move<0x150, 0xF>(Ql);  // row_newbcast:0
  P[1] = dpp_move<0x151, 0xF>(Pl), Q[1] = dpp_move<0x151, 0xF>(Ql);
  P[2] = dpp_move<0x152, 0xF>(Pl), Q[2] = dpp_move<0x152, 0xF>(Ql);
  P[3] = dpp_move<0x153, 0xF>(Pl), Q[3] = dpp_move<0x153, 0xF>(Ql);
  P[4] = dpp_move<0x154, 0xF>(Pl), Q[4] = dpp_move<0x154, 0xF>(Ql);
  P[5] = dpp_move<0x155, 0xF>(Pl), Q[5] = dpp_move<0x155, 0xF>(Ql);
  P[6] = dpp_move<0x156, 0xF>(Pl), Q[6] = dpp_move<0x156, 0xF>(Ql);
  P[7] = dpp_move<0x157, 0xF>(Pl), Q[7] = dpp_move<0x157, 0xF>(Ql);
  const int z0 = (az.low - oz0) % RZ;
  // rows ky = 2 kq and 2 kq + 1 of the stencil, PX points apart in the ring
  const T* base = ring + ((ay.low - oy + 2 * kq) * PX + (ax.low - ox) + kx) * 3;
  const T R0 = s.a0 + s.a2 * (ex.r2 + ey0.r2);
  const T R1 = s.a0 + s.a2 * (ex.r2 + ey1.r2);
  T a0 = T(0), a1 = T(0), a2 = T(0), b0 = T(0), b1 = T(0), b2 = T(0);
#pragma unroll
  for (int kz = 0; kz < 8; ++kz) {
    int zs = z0 + kz;
    zs -= zs >= RZ ? RZ : 0;
    const T* q = base + zs * (PLANE * 3);
    const T t0 = R0 * P[kz] + Q[kz];
    const T t1 = R1 * P[kz] + Q[kz];
    a0 += t0 * q[0];
    a1 += t0 * q[1];
    a2 += t0 * q[2];
    b0 += t1 * q[3 * PX + 0];
    b1 += t1 * q[3 * PX + 1];
    b2 += t1 * q[3 * PX + 2];
  }
  const T G0 = ex.g * ey0.g, G1 = ex.g * ey1.g;
  v0 = G0 * a0 + G1 * b0;
  v1 = G0 * a1 + G1 * b1;
  v2 = G0 * a2 + G1 * b2;
  // sum over the half: pairs, quads, half-rows, rows (16 lanes), then row 0 -> 1 (2 -> 3)
  v0 += dpp_move<0xB1, 0xF>(v0);
  v1 += dpp_move<0xB1, 0xF>(v1);
  v2 += dpp_move<0xB1, 0xF>(v2);
  v0 += dpp_move<0x4E, 0xF>(v0);
  v1 += dpp_move<0x4E, 0xF>(v1);
  v2 += dpp_move<0x4E, 0xF>(v2);
  v0 += dpp_move<0x141, 0xF>(v0);
  v1 += dpp_move<0x141, 0xF>(v1);
  v2 += dpp_move<0x141, 0xF>(v2);
  v0 += dpp_move<0x140, 0xF>(v0);
  v1 += dpp_move<0x140, 0xF>(v1);
  v2 += dpp_move<0x140, 0xF>(v2);
  v0 += dpp_move<0x142, 0xA>(v0);
  v1 += dpp_move<0x142, 0xA>(v1);
  v2 += dpp_move<0x142, 0xA>(v2);
}

// Pullback grid part of one particle from the same ring, half a wave per particle: for the
// weight vector w (G or F) the sum over the stencil of dW/dY(g) (w . u(g)), with
//   dW/dY_i = x_i gx gy gz c,  c = (a0 + a2 (r2x + r2y + r2z)) / Sigma^2 - 2 a2
// (x_i = g_i h - Y_i). Lane (kx, kq) sums rows ky = 2 kq, 2 kq + 1 over kz with t = w . u:
//   A = sum gz t,  A' = sum r2z gz t,  B = sum xz gz t,  B' = sum xz r2z gz t,
// and with C(ky) = (a0 + a2 (r2x + r2y)) / Sigma^2 - 2 a2, k = a2 / Sigma^2:
//   d_x += xx gx gy (C A + k A'),  d_y += xy gx gy (C A + k A'),  d_z += gx gy (C B + k B').
// Sums complete in lane 31 of the half.
template <int PX, int PLANE, int RZ>
__device__ __forceinline__ void ring_half_grad(const float* ring, const Vec<float>::v4& y, const Vec<float>::v4& w,
                                               const Scalars<float>& s, int ox, int oy, int oz0, int kx, int kq,
                                               float& d0, float& d1, float& d2) {
  using T = float;
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  const AxisOrigin ax = axis_origin_wrapped(y.x, s.inv_h, s.half, Mx);
  const AxisOrigin ay = axis_origin_wrapped(y.y, s.inv_h, s.half, My);
  const AxisOrigin az = axis_origin_wrapped(y.z, s.inv_h, s.half, Mz);
  const T c2 = s.inv_2S2 * 1.4426950408889634f;  // log2(e) / 2 Sigma^2
  const AxisEntry<T> ex = axis_entry_exp2(ax.lo, kx, y.x, s.h, s.inv_norm, c2);
  const AxisEntry<T> ey0 = axis_entry_exp2(ay.lo, 2 * kq, y.y, s.h, s.inv_norm, c2);
  const AxisEntry<T> ey1 = axis_entry_exp2(ay.lo, 2 * kq + 1, y.y, s.h, s.inv_norm, c2);
  const AxisEntry<T> ez = axis_entry_exp2(az.lo, kx, y.z, s.h, s.inv_norm, c2);
  const T xx = (T)(ax.lo + kx) * s.h - y.x;
  const T xy0 = (T)(ay.lo + 2 * kq) * s.h - y.y;
  const T xy1 = (T)(ay.lo + 2 * kq + 1) * s.h - y.y;
  const T xzl = (T)(az.lo + kx) * s.h - y.z;
  // z entry kz from lane kx = kz of each row
  const T Pl = ez.g, Ql = ez.r2 * ez.g, Xl = xzl * ez.g, Zl = xzl * ez.r2 * ez.g;
  T P[8], Q[8], X[8], Z[8];
#define FFCM_ROW_BCAST(k)                    \
  P[k] = dpp_move<0x150 + k, 0xF>(Pl);      \
  Q[k] = dpp_move<0x150 + k, 0xF>(Ql);      \
  X[k] = dpp_move<0x150 + k, 0xF>(Xl);      \
  Z[k] = dpp_move<0x150 + k, 0xF>(Zl);
  FFCM_ROW_BCAST(0)
  FFCM_ROW_BCAST(1)
  FFCM_ROW_BCAST(2)
  FFCM_ROW_BCAST(3)
  FFCM_ROW_BCAST(4)
  FFCM_ROW_BCAST(5)
  FFCM_ROW_BCAST(6)
  FFCM_ROW_BCAST(7)
#undef FFCM_ROW_BCAST
  const int z0 = (az.low - oz0) % RZ;
  const T* base = ring + ((ay.low - oy + 2 * kq) * PX + (ax.low - ox) + kx) * 3;
  T A0 = T(0), A0p = T(0), B0 = T(0), B0p = T(0);  // row ky = 2 kq
  T A1 = T(0), A1p = T(0), B1 = T(0), B1p = T(0);  // row ky = 2 kq + 1
#pragma unroll
  for (int kz = 0; kz < 8; ++kz) {
    int zs = z0 + kz;
    zs -= zs >= RZ ? RZ : 0;
    const T* q = base + zs * (PLANE * 3);
    const T t0 = w.x * q[0] + w.y * q[1] + w.z * q[2];
    const T t1 = w.x * q[3 * PX + 0] + w.y * q[3 * PX + 1] + w.z * q[3 * PX + 2];
    A0 += P[kz] * t0;
    A0p += Q[kz] * t0;
    B0 += X[kz] * t0;
    B0p += Z[kz] * t0;
    A1 += P[kz] * t1;
    A1p += Q[kz] * t1;
    B1 += X[kz] * t1;
    B1p += Z[kz] * t1;
  }
  const T k = s.a2 * s.inv_S2, two_a2 = T(2) * s.a2;
  const T C0 = (s.a0 + s.a2 * (ex.r2 + ey0.r2)) * s.inv_S2 - two_a2;
  const T C1 = (s.a0 + s.a2 * (ex.r2 + ey1.r2)) * s.inv_S2 - two_a2;
  const T G0 = ex.g * ey0.g, G1 = ex.g * ey1.g;
  const T SA0 = G0 * (C0 * A0 + k * A0p), SA1 = G1 * (C1 * A1 + k * A1p);
  d0 = xx * (SA0 + SA1);
  d1 = xy0 * SA0 + xy1 * SA1;
  d2 = G0 * (C0 * B0 + k * B0p) + G1 * (C1 * B1 + k * B1p);
  d0 += dpp_move<0xB1, 0xF>(d0);
  d1 += dpp_move<0xB1, 0xF>(d1);
  d2 += dpp_move<0xB1, 0xF>(d2);
  d0 += dpp_move<0x4E, 0xF>(d0);
  d1 += dpp_move<0x4E, 0xF>(d1);
  d2 += dpp_move<0x4E, 0xF>(d2);
  d0 += dpp_move<0x141, 0xF>(d0);
  d1 += dpp_move<0x141, 0xF>(d1);
  d2 += dpp_move<0x141, 0xF>(d2);
  d0 += dpp_move<0x140, 0xF>(d0);
  d1 += dpp_move<0x140, 0xF>(d1);
  d2 += dpp_move<0x140, 0xF>(d2);
  d0 += dpp_move<0x142, 0xA>(d0);
  d1 += dpp_move<0x142, 0xA>(d1);
  d2 += dpp_move<0x142, 0xA>(d2);
}

template <int TX, int TY, int TZ, int NT>
struct ColumnShape {
  static constexpr int PX = TX + 7, PY = TY + 7, RZ = TZ + 7;
  static constexpr int PLANE = PX * PY;
  static constexpr int NPTS = TZ * PLANE;  // points a layer adds
  static constexpr int KS = (NPTS + NT - 1) / NT;  // of them per thread
  static constexpr int RING = RZ * PLANE;
};

// kRingThreads threads per block: the ring fills the LDS of a CU, so the block is all the CU
// runs; 1024 threads keep four waves per SIMD in flight, 512 two.
template <typename T, int TX, int TY, int TZ, int SL, int kRingThreads>
__device__ __forceinline__ void interp_column_body(const Blk& blk, const InterpArgs<T>& a) {
  const typename Vec<T>::v4* __restrict__ Ys = a.Ys;
  const int* __restrict__ perm = a.perm;
  const int* __restrict__ tile_start = a.tile_start;
  const int* __restrict__ tile_list = a.order;
  const T* __restrict__ grid = a.grid;
  T* __restrict__ V = a.V;
  const Scalars<T>& s = a.s;
  const Tiling& tl = a.tl;
  const RingPlan& rp = a.rp;
  using V4 = typename Vec<T>::v4;
  constexpr int SZ = SL * TZ;  // planes one step adds (SL tile layers)
  using Sh = ColumnShape<TX, TY, SZ, kRingThreads>;
  constexpr int PX = Sh::PX, RZ = Sh::RZ, PLANE = Sh::PLANE, NPTS = Sh::NPTS, KS = Sh::KS;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int lay_start[kRingMaxLayers];
  __shared__ int lay_count[kRingMaxLayers];
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  const int per_xcd = blk.n / kXcds;  // the block count is a multiple of kXcds
  const int lb = (blk.id % kXcds) * per_xcd + blk.id / kXcds;
  if (lb >= rp.ncols * rp.nseg) return;
  // segment-major, so that the blocks in flight share a z range; inside a segment, columns in
  // patches of 8 x 4 (when the tile counts allow) so that they also share halos
  const int seg = lb / rp.ncols;
  const int col = lb - seg * rp.ncols;
  int cx, cy;
  if (rp.patch) {
    const int p = col >> 5, w = col & 31, npx = tl.nt[0] >> 3;
    cx = (p % npx) * 8 + (w & 7);
    cy = (p / npx) * 4 + (w >> 3);
  } else {
    cx = col % tl.nt[0];
    cy = col / tl.nt[0];
  }
  const int ox = cx * TX, oy = cy * TY;
  const int tz0 = seg * rp.seg_layers;
  const int nl = min(rp.seg_layers, tl.nt[2] - tz0);  // tile layers of the run
  const int ns = (nl + SL - 1) / SL;                  // steps
  const int oz0 = tz0 * TZ;
  constexpr int kRingBytes = (3 * Sh::RING * (int)sizeof(T) + 15) / 16 * 16;
  T* ring = (T*)smem_raw;                 // [RZ][PY][PX][3]
  V4* pY = (V4*)(smem_raw + kRingBytes);  // [kRingStaged] positions of a step's particles, with
                                          // the original index in w (as int bits)
  const int tid = threadIdx.x;
  for (int j = tid; j < nl; j += kRingThreads) {
    const int t = cx + tl.nt[0] * (cy + tl.nt[1] * (tz0 + j));
    const int a = tile_start[t];
    lay_start[j] = a;
    lay_count[j] = tile_start[t + 1] - a;
  }
  const unsigned plane_stride = (unsigned)My * Mx * 3;
  // grid offset of (x, y) of in-plane point rem of the column
  auto xy_offset = [&](int rem) -> unsigned {
    const int yi = rem / PX;
    const int xi = rem - yi * PX;
    int gy = oy + yi;
    gy -= gy >= My ? My : 0;
    int gx = ox + xi;
    gx -= gx >= Mx ? Mx : 0;
    return ((unsigned)gy * Mx + gx) * 3;
  };
  auto z_offset = [&](int gz) -> unsigned {
    gz -= gz >= Mz ? Mz : 0;
    return (unsigned)gz * plane_stride;
  };
  // prologue: the RZ planes of step 0, through registers
  {
    T st[KS][3];
    for (int base = 0; base < Sh::RING; base += KS * kRingThreads) {
#pragma unroll
      for (int k = 0; k < KS; ++k) {
        const int p = min(base + tid + k * kRingThreads, Sh::RING - 1);
        const int zi = p / PLANE;
        const T* q = grid + z_offset(oz0 + zi) + xy_offset(p - zi * PLANE);
        st[k][0] = q[0];
        st[k][1] = q[1];
        st[k][2] = q[2];
      }
#pragma unroll
      for (int k = 0; k < KS; ++k) {
        const int p = base + tid + k * kRingThreads;
        if (p < Sh::RING) {
          T* r = ring + p * 3;
          r[0] = st[k][0];
          r[1] = st[k][1];
          r[2] = st[k][2];
        }
      }
    }
  }
  // this thread's points of every step's new planes: point tid + k kRingThreads (clamped);
  // their in-plane grid offsets are kept in registers by the 512-thread block and recomputed
  // per step by the 1024-thread one, whose 128 registers per lane have no room for them
  constexpr bool kKeepOffsets = kRingThreads <= 512;
  unsigned goff[kKeepOffsets ? KS : 1];
#pragma unroll
  for (int k = 0; k < (kKeepOffsets ? KS : 1); ++k) {
    const int p = min(tid + k * kRingThreads, NPTS - 1);
    goff[k] = xy_offset(p % PLANE);
  }
  auto point_offset = [&](int k) -> unsigned {
    if constexpr (kKeepOffsets) return goff[k];
    return xy_offset(min(tid + k * kRingThreads, NPTS - 1) % PLANE);
  };
  __syncthreads();  // lay_start / lay_count
  // particles of step j: those of tile layers SL j .. SL j + SL - 1, layer after layer
  auto step_count = [&](int j) -> int {
    int n = 0;
#pragma unroll
    for (int t = 0; t < SL; ++t) n += SL * j + t < nl ? lay_count[SL * j + t] : 0;
    return n;
  };
  // tile-list index of particle i of step j (i < step_count(j))
  auto step_entry = [&](int j, int i) -> int {
#pragma unroll
    for (int t = 0; t < SL - 1; ++t) {
      const int c = lay_count[SL * j + t];
      if (i < c) return lay_start[SL * j + t] + i;
      i -= c;
    }
    return lay_start[SL * j + SL - 1] + i;
  };
  if (tid < min(step_count(0), kRingStaged)) {
    const int sl = tile_list[step_entry(0, tid)];
    V4 y = Ys[sl];
    y.w = __int_as_float(perm[sl]);
    pY[tid] = y;
  }
  vm_wait_all();
  // This thread's particle slot of step j, or slot 0 where it has none (the load is issued
  // either way, so that every step issues the same loads).
  auto slot_of = [&](int j) -> int {
    const bool has = j < ns && tid < min(step_count(min(j, ns - 1)), kRingStaged);
    return tile_list[has ? step_entry(j, tid) : 0];
  };
  // Loads for a later step, held in registers: this thread's points of the planes it adds and
  // the position and original index of this thread's particle of it.
  struct Stage {
    T g[KS][3];
    T y[3];  // components, not a vector member: keeps the stage in registers
    int n;
  };
  auto issue_particle = [&](Stage& S, int slot) {
    const V4 y = Ys[slot];
    S.y[0] = y.x;
    S.y[1] = y.y;
    S.y[2] = y.z;
    S.n = perm[slot];
  };
  // the planes step j adds (rz from (j - 1) SZ + RZ on); a step past the run loads one grid
  // point in every lane
  auto issue_planes = [&](Stage& S, int j) {
    const bool live = j < ns;
    const int gz0 = oz0 + (j - 1) * SZ + RZ;
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int zi = min(tid + k * kRingThreads, NPTS - 1) / PLANE;
      const T* q = live ? grid + z_offset(gz0 + zi) + point_offset(k) : grid;
      S.g[k][0] = q[0];
      S.g[k][1] = q[1];
      S.g[k][2] = q[2];
    }
  };
  // after the barrier that ends step j - 1: the new planes over the ring slots of planes
  // (j - 1) SZ .. j SZ - 1, the particle to the LDS buffer
  auto commit = [&](const Stage& S, int j) {
    if (j >= ns) return;
    const int slot0 = ((j - 1) * SZ) % RZ;
#pragma unroll
    for (int k = 0; k < KS; ++k) {
      const int p = tid + k * kRingThreads;
      if (p < NPTS) {
        const int zi = p / PLANE;
        int zs = slot0 + zi;
        zs -= zs >= RZ ? RZ : 0;
        T* r = ring + (zs * PLANE + (p - zi * PLANE)) * 3;
        r[0] = S.g[k][0];
        r[1] = S.g[k][1];
        r[2] = S.g[k][2];
      }
    }
    if (tid < min(step_count(j), kRingStaged)) {
      V4 y;
      y.x = S.y[0];
      y.y = S.y[1];
      y.z = S.y[2];
      y.w = __int_as_float(S.n);
      pY[tid] = y;
    }
  };
  const int lane = tid & 63;
  const int wave = tid >> 6;
  constexpr int nwaves = kRingThreads / 64;
  // two particles per wave, one per half; in a half, lane kx + 8 kq covers stencil points
  // (kx, ky, kz) with ky = 2 kq, 2 kq + 1 and all kz. Each 16-lane row holds one particle.
  const int half = lane >> 5;
  const int kx = lane & 7;
  const int kq = (lane >> 3) & 3;
  auto half_velocity = [&](const V4& y, T& v0, T& v1, T& v2) {
    ring_half_velocity<PX, PLANE, RZ>(ring, y, s, ox, oy, oz0, kx, kq, v0, v1, v2);
  };
  const bool writer = (lane & 31) == 31;
  // step j from the ring, two particles per wave; velocities straight to memory
  auto compute = [&](int j) {
    if (j >= ns) return;
    const int n = step_count(j);
    const int nlds = min(n, kRingStaged);
    for (int i0 = wave * 2; i0 < nlds; i0 += nwaves * 2) {
      const int i = i0 + half;
      const bool ok = i < nlds;
      const V4 y = pY[ok ? i : i0];
      T v0, v1, v2;
      half_velocity(y, v0, v1, v2);
      if (ok && writer) {
        const int on = __float_as_int(y.w);
        V[3 * on + 0] = s.h3 * v0;
        V[3 * on + 1] = s.h3 * v1;
        V[3 * on + 2] = s.h3 * v2;
      }
    }
    // particles beyond the LDS buffer (a step holding more than kRingStaged)
    for (int i0 = kRingStaged + wave * 2; i0 < n; i0 += nwaves * 2) {
      const int i = i0 + half;
      const bool ok = i < n;
      const int sl = tile_list[step_entry(j, ok ? i : i0)];
      T v0, v1, v2;
      half_velocity(Ys[sl], v0, v1, v2);
      if (ok && writer) {
        const int on = perm[sl];
        V[3 * on + 0] = s.h3 * v0;
        V[3 * on + 1] = s.h3 * v1;
        V[3 * on + 2] = s.h3 * v2;
      }
    }
  };
  // Two steps of loads in flight: step j interpolates from the ring while the loads of step
  // j + 1 (issued the step before, in X) and of step j + 2 (issued now, into Y) are
  // outstanding, then waits only for X. Slots run one step further ahead, in `carry`.
  if constexpr (kKeepOffsets) {
    Stage A, B;
    issue_particle(A, slot_of(1));
    int carry = slot_of(2);
    issue_planes(A, 1);
    __syncthreads();
    auto step = [&](int j, Stage& X, Stage& Y) {
      issue_particle(Y, carry);
      carry = slot_of(j + 3);
      issue_planes(Y, j + 2);
      compute(j);
      __syncthreads();
      commit(X, j + 1);
      __syncthreads();
    };
    // both steps of every iteration run (a step past the run issues only placeholder loads),
    // so that the loads outstanding at the loop head are the same on every path
    for (int j = 0; j < ns; j += 2) {
      step(j, A, B);
      step(j + 1, B, A);
    }
  } else {
    // one step of loads in flight: four waves per SIMD cover the latency that the second
    // stage covers for the 512-thread block
    Stage A;
    __syncthreads();
    for (int j = 0; j < ns; ++j) {
      issue_particle(A, slot_of(j + 1));
      issue_planes(A, j + 1);
      compute(j);
      __syncthreads();
      commit(A, j + 1);
      __syncthreads();
    }
  }
}
// ---- column-streamed interpolation, ring filled by LDS DMA ----
//
// The same column walk as above, but the grid planes go from memory straight into the LDS
// ring (global_load_lds_dwordx4: each lane moves 16 bytes, a wave 1 KiB of consecutive LDS)
// with no register staging and no copy-in phase. A ring plane is 23 rows of 24 points (the
// 16 x 16 column, its 7-point halo and one unused point, so that a row is 72 floats, 18 whole
// 16-byte units; a unit never straddles the periodic wrap in x because the wrap falls on a
// multiple of 48 floats). The ring holds 23 planes, plane P of the run in slot P mod 23.
//
// A step interpolates the particles whose stencils start in SP consecutive planes (a tile
// layer holds 8 / SP steps); they read SP + 7 planes. The other 16 - SP slots hold planes in
// flight: at the start of step j the SP slots that step j - 1 was the last to read receive the
// planes SP j + 23 - SP .. SP j + 22, which step j + 1 + ... reads several steps later, so each
// transfer has several steps to land.
//
// Roles (wave index made uniform): waves 0..12 interpolate, two particles per wave at a time;
// waves 13 and 14 fill the planes of parity 0 and 1 (seven 1 KiB blocks per 6,624-byte plane,
// the last one partly) and at the end of a step wait only for the transfers the next step
// reads (vmcnt counts the newer ones, issued in order); wave 15 stages the positions and
// original indices of the next tile layer's first 64 particles in LDS, at the first step of
// each layer, and waits for them. Each role runs its own loop with one barrier per step, so
// the compiler never sees a transfer in flight in the interpolating waves and adds no wait
// before their LDS reads.
constexpr int kDmaPX = 24;                               // ring points per row
constexpr int kDmaPY = 23;                               // ring rows per plane
constexpr int kDmaRZ = 23;                               // ring planes
constexpr int kDmaRowBytes = kDmaPX * 12;                // 288
constexpr int kDmaPlaneBytes = kDmaPY * kDmaRowBytes;    // 6624
constexpr int kDmaUnitsPerPlane = (kDmaPlaneBytes + 1023) / 1024;  // 7
constexpr int kDmaLastLanes = (kDmaPlaneBytes - (kDmaUnitsPerPlane - 1) * 1024) / 16;  // 30
constexpr int kDmaStaged = 64;                           // particles per layer staged in LDS
constexpr int kDmaThreads = 1024;
constexpr int kDmaWaves = kDmaThreads / 64;
constexpr int kDmaComputeWaves = kDmaWaves - 3;
constexpr int kDmaRingBytes = kDmaRZ * kDmaPlaneBytes;   // 152,352
constexpr int kDmaLdsBytes = kDmaRingBytes + 2 * kDmaStaged * 16 + 2 * kDmaStaged * 4;
constexpr int kDmaGradLdsBytes = kDmaLdsBytes + 2 * kDmaStaged * 16;  // and the weight vectors
static_assert(kDmaRowBytes % 16 == 0 && kDmaPlaneBytes % 16 == 0, "16-byte DMA units");

// Steps (of SP planes) that a plane-filling wave may leave in flight at the end of a step:
// the planes issued at step k are SP (k - 1) + 23 .. SP k + 22, and step j + 1 reads up to
// plane SP (j + 2) + 6.
template <int SP>
constexpr int dma_steps_in_flight() {
  int k = -64;
  while (SP * k + 22 < SP * 2 + 6) ++k;  // first step (relative to j = 0) whose planes step 1 needs
  return -k;                              // steps k + 1 .. 0 may still be in flight
}

// kGrad: the pullback's grid part (ring_half_grad with the weight vectors a.W, staged beside
// the positions) in place of the velocity.
template <int SP, bool kGrad>
__device__ __forceinline__ void interp_dma_body(const Blk& blk, const InterpArgs<float>& a) {
  using T = float;
  using V4 = Vec<T>::v4;
  static_assert(SP == 2 || SP == 4 || SP == 8, "steps of 2, 4 or 8 planes");
  constexpr int kSub = 8 / SP;                                  // steps per tile layer
  constexpr int kWaitCnt = dma_steps_in_flight<SP>() * (SP / 2) * kDmaUnitsPerPlane;
  static_assert(kWaitCnt <= 63, "vmcnt range");
  const V4* __restrict__ Ys = a.Ys;
  const int* __restrict__ perm = a.perm;
  const int* __restrict__ tile_start = a.tile_start;
  const int* __restrict__ tile_list = a.order;
  const T* __restrict__ grid = a.grid;
  T* __restrict__ V = a.V;
  const Scalars<T>& s = a.s;
  const Tiling& tl = a.tl;
  const RingPlan& rp = a.rp;
  constexpr int TX = 16, TY = 16, TZ = 8;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int lay_start[kRingMaxLayers];
  __shared__ int lay_count[kRingMaxLayers];
  const int Mx = s.M[0], My = s.M[1], Mz = s.M[2];
  const int per_xcd = blk.n / kXcds;  // the block count is a multiple of kXcds
  const int lb = (blk.id % kXcds) * per_xcd + blk.id / kXcds;
  if (lb >= rp.ncols * rp.nseg) return;
  // segment-major, so that the blocks in flight share a z range; inside a segment, columns in
  // patches of 8 x 4 (when the tile counts allow) so that they also share halos in L2
  const int seg = lb / rp.ncols;
  const int col = lb - seg * rp.ncols;
  int cx, cy;
  if (rp.patch) {
    const int p = col >> 5, w = col & 31, npx = tl.nt[0] >> 3;
    cx = (p % npx) * 8 + (w & 7);
    cy = (p / npx) * 4 + (w >> 3);
  } else {
    cx = col % tl.nt[0];
    cy = col / tl.nt[0];
  }
  const int ox = cx * TX, oy = cy * TY;
  const int tz0 = seg * rp.seg_layers;
  const int nl = min(rp.seg_layers, tl.nt[2] - tz0);  // tile layers of the run
  const int ns = nl * kSub;                           // steps
  const int oz0 = tz0 * TZ;
  const int pend = TZ * nl + 7;                       // planes the run reads
  unsigned char* ring_b = smem_raw;
  const T* ring = (const T*)smem_raw;
  V4* pY = (V4*)(smem_raw + kDmaRingBytes);  // [2][kDmaStaged]
  int* pN = (int*)(pY + 2 * kDmaStaged);      // [2][kDmaStaged]
  V4* pW = (V4*)(pN + 2 * kDmaStaged);        // [2][kDmaStaged], gradient mode only
  const V4* __restrict__ Ws = a.W;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int j = tid; j < nl; j += kDmaThreads) {
    const int t = cx + tl.nt[0] * (cy + tl.nt[1] * (tz0 + j));
    const int st = tile_start[t];
    lay_start[j] = st;
    lay_count[j] = tile_start[t + 1] - st;
  }
  lds_sync();  // lay_start / lay_count
  if (wave >= kDmaComputeWaves && wave < kDmaComputeWaves + 2) {
    // ---- plane filler q: planes of parity q
    const int q = wave - kDmaComputeWaves;
    unsigned xy[kDmaUnitsPerPlane];  // grid offset (floats, in a plane) of this lane's unit of each block
#pragma unroll
    for (int u = 0; u < kDmaUnitsPerPlane; ++u) {
      const int b = min(u * 1024 + lane * 16, kDmaPlaneBytes - 16);
      const int row = b / kDmaRowBytes;
      const int fx = (b - row * kDmaRowBytes) >> 2;
      int gy = oy + row;
      gy -= gy >= My ? My : 0;
      int gxf = 3 * ox + fx;
      gxf -= gxf >= 3 * Mx ? 3 * Mx : 0;
      xy[u] = (unsigned)gy * (unsigned)(3 * Mx) + (unsigned)gxf;
    }
    const size_t plane_stride = (size_t)My * Mx * 3;
    auto fill_plane = [&](int P) {
      if ((P & 1) != q || P >= pend) return;
      int gz = oz0 + P;
      gz -= gz >= Mz ? Mz : 0;
      const T* src = grid + (size_t)gz * plane_stride;
      unsigned char* dst = ring_b + (P % kDmaRZ) * kDmaPlaneBytes;
#pragma unroll
      for (int u = 0; u < kDmaUnitsPerPlane - 1; ++u)
        __builtin_amdgcn_global_load_lds(src + xy[u], dst + u * 1024, 16, 0, 0);
      if (lane < kDmaLastLanes)
        __builtin_amdgcn_global_load_lds(src + xy[kDmaUnitsPerPlane - 1], dst + (kDmaUnitsPerPlane - 1) * 1024, 16,
                                         0, 0);
    };
    // prologue: the planes step 0 reads, waited for; then those that steps k = -D .. -1 of the
    // pipeline would have issued (SP (k - 1) + 23 .. SP k + 22, clipped below at SP + 7)
    for (int P = 0; P < SP + 7; ++P) fill_plane(P);
    vm_wait_all();
    constexpr int D = dma_steps_in_flight<SP>();
    for (int k = -D; k <= -1; ++k)
      for (int P = max(SP * (k - 1) + 23, SP + 7); P < SP * k + 23; ++P) fill_plane(P);
    // barriers without the LDS release fence, which would wait for every transfer in flight
    // (the fence counts a transfer's LDS write as this wave's own); the waves that read the
    // ring have their own fences
    __builtin_amdgcn_s_barrier();
    for (int j = 0; j < ns; ++j) {
      // over the slots of the planes step j - 1 was the last to read
#pragma unroll
      for (int i = 0; i < SP; ++i) fill_plane(SP * (j - 1) + 23 + i);
      __builtin_amdgcn_s_waitcnt(0x0F70 | (kWaitCnt & 15) | ((kWaitCnt >> 4) << 14));
      __builtin_amdgcn_s_barrier();
    }
  } else if (wave == kDmaComputeWaves + 2) {
    // ---- particle stager: layer L + 1 at the first step of layer L
    auto stage_layer = [&](int L) {
      if (L >= nl) return;
      const int slot = lane < min(lay_count[L], kDmaStaged) ? tile_list[lay_start[L] + lane] : -1;
      vm_wait_all();
      if (slot >= 0) {
        __builtin_amdgcn_global_load_lds(Ys + slot, pY + (L & 1) * kDmaStaged, 16, 0, 0);
        __builtin_amdgcn_global_load_lds(perm + slot, pN + (L & 1) * kDmaStaged, 4, 0, 0);
        if constexpr (kGrad) __builtin_amdgcn_global_load_lds(Ws + slot, pW + (L & 1) * kDmaStaged, 16, 0, 0);
      }
      vm_wait_all();
    };
    stage_layer(0);
    lds_sync();
    for (int j = 0; j < ns; ++j) {
      if (j % kSub == 0) stage_layer(j / kSub + 1);
      lds_sync();
    }
  } else {
    // ---- interpolation: the particles of layer L whose stencils start in the step's planes
    const int half = lane >> 5;
    const int kx = lane & 7;
    const int kq = (lane >> 3) & 3;
    const bool writer = (lane & 31) == 31;
    const bool acc = kGrad && a.accumulate;
    auto dma_write = [&](int on, T v0, T v1, T v2) {
      T* o = V + 3 * on;
      if (acc) {
        o[0] += s.h3 * v0;
        o[1] += s.h3 * v1;
        o[2] += s.h3 * v2;
      } else {
        o[0] = s.h3 * v0;
        o[1] = s.h3 * v1;
        o[2] = s.h3 * v2;
      }
    };
    lds_sync();
    for (int j = 0; j < ns; ++j) {
      const int L = j / kSub;
      const int z_lo = oz0 + TZ * L + SP * (j - L * kSub);  // first stencil plane of the step
      const int n = lay_count[L];
      const int nst = min(n, kDmaStaged);
      const V4* py = pY + (L & 1) * kDmaStaged;
      const int* pn = pN + (L & 1) * kDmaStaged;
      // staged particles from LDS (i0 even and nst <= 64: i is staged where it is < nst)
      for (int i0 = wave * 2; i0 < nst; i0 += kDmaComputeWaves * 2) {
        const int i = i0 + half;
        const bool ok = i < nst;
        const int ii = ok ? i : i0;
        const V4 y = py[ii];
        bool take = ok;
        if constexpr (SP < TZ) {  // steps finer than a tile layer: the particles starting in them
          const int zl = axis_origin_wrapped(y.z, s.inv_h, s.half, Mz).low;
          take = ok && (unsigned)(zl - z_lo) < (unsigned)SP;
          if (!__any(take)) continue;
        }
        T v0, v1, v2;
        if constexpr (kGrad)
          ring_half_grad<kDmaPX, kDmaPX * kDmaPY, kDmaRZ>(ring, y, pW[(L & 1) * kDmaStaged + ii], s, ox, oy, oz0, kx,
                                                          kq, v0, v1, v2);
        else
          ring_half_velocity<kDmaPX, kDmaPX * kDmaPY, kDmaRZ>(ring, y, s, ox, oy, oz0, kx, kq, v0, v1, v2);
        if (take && writer) dma_write(pn[ii], v0, v1, v2);
      }
      // beyond the staged count, from memory
      for (int i0 = kDmaStaged + wave * 2; i0 < n; i0 += kDmaComputeWaves * 2) {
        const int i = i0 + half;
        const bool ok = i < n;
        const int sl = tile_list[lay_start[L] + (ok ? i : i0)];
        const V4 y = Ys[sl];
        bool take = ok;
        if constexpr (SP < TZ) {
          const int zl = axis_origin_wrapped(y.z, s.inv_h, s.half, Mz).low;
          take = ok && (unsigned)(zl - z_lo) < (unsigned)SP;
          if (!__any(take)) continue;
        }
        T v0, v1, v2;
        if constexpr (kGrad)
          ring_half_grad<kDmaPX, kDmaPX * kDmaPY, kDmaRZ>(ring, y, Ws[sl], s, ox, oy, oz0, kx, kq, v0, v1, v2);
        else
          ring_half_velocity<kDmaPX, kDmaPX * kDmaPY, kDmaRZ>(ring, y, s, ox, oy, oz0, kx, kq, v0, v1, v2);
        if (take && writer) dma_write(perm[sl], v0, v1, v2);
      }
      lds_sync();
    }
  }
}

template <typename T, int MGC>
__global__ void __launch_bounds__(kInterpThreads) interp_kernel(InterpArgs<T> a) {
  interp_body<T, MGC>(plain_blk(), a);
}
template <typename T, int MGC>
__global__ void __launch_bounds__(kInterpThreads) interp_kernel_group(const GroupEntry<InterpArgs<T>>* __restrict__ g,
                                                                       int n) {
  group_run(g, n, [](const Blk& b, const InterpArgs<T>& a) { interp_body<T, MGC>(b, a); });
}
template <typename T, int TX, int TY, int TZ, int SL, int kRingThreads>
__global__ void __launch_bounds__(kRingThreads) interp_column_kernel(InterpArgs<T> a) {
  interp_column_body<T, TX, TY, TZ, SL, kRingThreads>(plain_blk(), a);
}
template <typename T, int TX, int TY, int TZ, int SL, int kRingThreads>
__global__ void __launch_bounds__(kRingThreads)
    interp_column_kernel_group(const GroupEntry<InterpArgs<T>>* __restrict__ g, int n) {
  group_run(g, n,
            [](const Blk& b, const InterpArgs<T>& a) { interp_column_body<T, TX, TY, TZ, SL, kRingThreads>(b, a); });
}
template <int SP>
__global__ void __launch_bounds__(kDmaThreads) interp_dma_kernel(InterpArgs<float> a) {
  interp_dma_body<SP, false>(plain_blk(), a);
}
template <int SP>
__global__ void __launch_bounds__(kDmaThreads)
    interp_dma_kernel_group(const GroupEntry<InterpArgs<float>>* __restrict__ g, int n) {
  group_run(g, n, [](const Blk& b, const InterpArgs<float>& a) { interp_dma_body<SP, false>(b, a); });
}
__global__ void __launch_bounds__(kDmaThreads) interp_grad_dma_kernel(InterpArgs<float> a) {
  interp_dma_body<8, true>(plain_blk(), a);
}
}  // namespace

// Column-streamed plan: 16 x 16 x 8 spread tiles and 8^3 stencils in Float32, the grid
// addressable with 32-bit offsets and the LDS within the device limit; false otherwise.
template <typename T>
bool ring_plan(const ffcm_ctx* c, RingPlan* rp) {
  using Sh = ColumnShape<16, 16, 8 * kColumnLayers, 512>;
  const Tiling& tl = c->tiling;
  if (sizeof(T) != 4 || c->host.M_G != 8) return false;
  if (tl.T[0] != 16 || tl.T[1] != 16 || tl.T[2] != 8) return false;
  if (3 * c->M3 >= (int64_t(1) << 31)) return false;
  // the ring's planes and halo reach at most Mz + 14 planes past the run start (and Mx + 6,
  // My + 6 columns), which z_offset and xy_offset fold back with one subtraction of the
  // extent: exact from 16 planes per axis on
  if (c->host.M[0] < 16 || c->host.M[1] < 16 || c->host.M[2] < 16) return false;
  rp->lds_bytes = (3 * Sh::RING * (int)sizeof(T) + 15) / 16 * 16 + kRingStaged * (int)sizeof(typename Vec<T>::v4);
  rp->threads = 1024;
  if (const char* e = std::getenv("FFCM_INTERP_THREADS")) rp->threads = std::atoi(e) == 512 ? 512 : 1024;
  rp->dma = 0;
  if (rp->lds_bytes + 2 * kRingMaxLayers * (int)sizeof(int) > c->lds_limit) return false;
  rp->ncols = tl.nt[0] * tl.nt[1];
  // about 1024 blocks: whole columns where there are that many, else runs of layers
  int want = (1024 + rp->ncols - 1) / rp->ncols;
  if (const char* e = std::getenv("FFCM_INTERP_SEGS")) want = std::atoi(e);
  if (want < 1) want = 1;
  if (want > tl.nt[2]) want = tl.nt[2];
  rp->seg_layers = (tl.nt[2] + want - 1) / want;
  if (rp->seg_layers > kRingMaxLayers) rp->seg_layers = kRingMaxLayers;
  rp->nseg = (tl.nt[2] + rp->seg_layers - 1) / rp->seg_layers;
  rp->patch = (tl.nt[0] % 8 == 0 && tl.nt[1] % 4 == 0) ? 1 : 0;
  rp->blocks = (rp->ncols * rp->nseg + kXcds - 1) / kXcds * kXcds;
  return true;
}

// Column streaming reads the grid about twice (the tile halos) from L2 and once from memory
// whatever the number of particles; the per-particle gather reads every stencil row of every
// particle. Measured on 2^20 particles on 512^3 (stencils covering the grid 4 times): 0.84 ms
// against 0.97 ms; 2^16 on 256^3 (twice): 0.11 ms against 0.084 ms. Column streaming is used
// from 3 coverings on; FFCM_INTERP=column or =wave forces either where the column fits.
template <typename T>
bool use_ring(const ffcm_ctx* c, RingPlan* rp) {
  const char* e = std::getenv("FFCM_INTERP");
  if (e && std::strcmp(e, "wave") == 0) return false;
  if (!ring_plan<T>(c, rp)) return false;
  const bool staged = e && std::strcmp(e, "column") == 0;
  rp->dma = 0;
  if (!staged && kDmaLdsBytes + 2 * kRingMaxLayers * (int)sizeof(int) <= c->lds_limit) {
    rp->dma = 8;
    if (const char* d = std::getenv("FFCM_INTERP_STEP")) {
      const int v = std::atoi(d);
      if (v == 2 || v == 4 || v == 8) rp->dma = v;
    }
  }
  if (staged || (e && std::strcmp(e, "dma") == 0)) return true;
  const double cover = (double)c->N * 512.0 / (double)c->M3;
  return cover >= 3.0;
}

int interp_kernel_kind(const ffcm_ctx* c) {
  RingPlan rp;
  if (c->dtype != FFCM_F32 || !use_ring<float>(c, &rp)) return 0;
  return rp.dma ? 2 : 1;
}

template <typename T>
hipError_t launch_interpolate(ffcm_ctx* c, T* V, const T* grid, int tag) {
  if (!grid) grid = (const T*)c->d.velocity;
  if (!c->tiles_valid) {  // the visiting order is the spread's tile lists
    hipError_t e = launch_bin_tiles<T>(c);
    if (e != hipSuccess) return e;
  }
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  RingPlan rp;
  if (c->N <= 0) return hipSuccess;
  InterpArgs<T> a{(const typename Vec<T>::v4*)c->d.Y_sorted, c->d.perm, c->d.tile_start, c->d.tile_list, grid, V, s,
                  c->tiling, RingPlan{}};
  if constexpr (sizeof(T) == 4) {
    if (use_ring<T>(c, &rp)) {
      a.rp = rp;
      if (rp.dma == 2)
        return submit(c, tag, interp_dma_kernel<2>, interp_dma_kernel_group<2>, (unsigned)rp.blocks, kDmaThreads,
                      kDmaLdsBytes, a);
      if (rp.dma == 4)
        return submit(c, tag, interp_dma_kernel<4>, interp_dma_kernel_group<4>, (unsigned)rp.blocks, kDmaThreads,
                      kDmaLdsBytes, a);
      if (rp.dma == 8)
        return submit(c, tag, interp_dma_kernel<8>, interp_dma_kernel_group<8>, (unsigned)rp.blocks, kDmaThreads,
                      kDmaLdsBytes, a);
      if (rp.threads == 1024)
        return submit(c, tag, interp_column_kernel<T, 16, 16, 8, kColumnLayers, 1024>,
                      interp_column_kernel_group<T, 16, 16, 8, kColumnLayers, 1024>, (unsigned)rp.blocks, 1024,
                      rp.lds_bytes, a);
      return submit(c, tag, interp_column_kernel<T, 16, 16, 8, kColumnLayers, 512>,
                    interp_column_kernel_group<T, 16, 16, 8, kColumnLayers, 512>, (unsigned)rp.blocks, 512,
                    rp.lds_bytes, a);
    }
  }
  const int waves = kInterpThreads / 64;
  int64_t blocks = (c->N + waves - 1) / waves;
  if (blocks > (1 << 20)) blocks = 1 << 20;
  if (blocks < 1) blocks = 1;
  blocks = (blocks + kXcds - 1) / kXcds * kXcds;
  if (s.MG == 8)
    return submit(c, tag, interp_kernel<T, 8>, interp_kernel_group<T, 8>, (unsigned)blocks, kInterpThreads, 0, a);
  // M_G = 10: the published GPU comparison configuration (two passes per plane)
  if (s.MG == 10)
    return submit(c, tag, interp_kernel<T, 10>, interp_kernel_group<T, 10>, (unsigned)blocks, kInterpThreads, 0, a);
  return submit(c, tag, interp_kernel<T, 0>, interp_kernel_group<T, 0>, (unsigned)blocks, kInterpThreads, 0, a);
}

// The pullback's grid part through the DMA ring, in two passes: dY = h^3 sum dW/dY (G . u_F),
// then dY += h^3 sum dW/dY (F . u_G). Returns false, launching nothing, where the forward
// interpolation does not take the DMA ring (the caller then runs the per-particle kernel).
bool launch_interp_grad_ring(ffcm_ctx* c, float* dY, const float* uF, const float* uG, hipError_t* err) {
  RingPlan rp;
  if (c->N <= 0 || !c->tiles_valid || !use_ring<float>(c, &rp) || rp.dma == 0) return false;
  if (kDmaGradLdsBytes + 2 * kRingMaxLayers * (int)sizeof(int) > c->lds_limit) return false;
  using V4 = Vec<float>::v4;
  const Scalars<float>& s = *(const Scalars<float>*)c->scalars;
  InterpArgs<float> a{(const V4*)c->d.Y_sorted, c->d.perm, c->d.tile_start, c->d.tile_list, uF, dY, s, c->tiling, rp};
  a.W = (const V4*)c->d.G_sorted;
  a.accumulate = 0;
  interp_grad_dma_kernel<<<(unsigned)rp.blocks, kDmaThreads, kDmaGradLdsBytes, c->stream>>>(a);
  a.grid = uG;
  a.W = (const V4*)c->d.F_sorted;
  a.accumulate = 1;
  interp_grad_dma_kernel<<<(unsigned)rp.blocks, kDmaThreads, kDmaGradLdsBytes, c->stream>>>(a);
  *err = hipGetLastError();
  return true;
}

template hipError_t launch_interpolate<float>(ffcm_ctx*, float*, const float*, int);
template hipError_t launch_interpolate<double>(ffcm_ctx*, double*, const double*, int);

}  // namespace ffcm
