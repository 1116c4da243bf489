// The per-particle separable stencil shared by spreading and interpolation, so that
// interpolation stays the discrete adjoint of spreading.
//
// Reference: _fill_particle_stencil! (src/stencil.jl:80-112):
//   anchor  j = round(Y * inv_h), ties to even
//   point   g = j - M_G/2 + k, k = 0..M_G-1 (unwrapped), grid index mod(g, M)
//   weight  x = T(g) * h - Y;  gauss = inv_norm * exp(-x^2 * inv_2S2);  r2 = x^2
// and the weight product of src/spread_forces.jl:115-123:
//   w = (a0 + a2 * (r2x + (r2y + r2z))) * gx * (gy * gz)
// One wave handles one particle: lane k < M_G evaluates entry k of each axis, other lanes read
// it with a cross-lane permute.
#pragma once

#include "ffcm_internal.hpp"

namespace ffcm {

template <typename T>
__device__ __forceinline__ T wave_bcast(T v, int src) {
  return __shfl(v, src, 64);
}

// Broadcast of one lane's value when the lane index is wave-uniform: a scalar read of that
// lane's register (v_readlane), no LDS permute.
template <typename T>
__device__ __forceinline__ T lane_bcast(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

// Row-level data-parallel move of a 32- or 64-bit value: v_mov_dpp with control CTRL; lanes in
// rows outside ROW_MASK receive 0.
template <int CTRL, int ROW_MASK, typename T>
__device__ __forceinline__ T dpp_move(T v) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROW_MASK, 0xF, false));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

// Sum over the 64 lanes, returned in every lane: adjacent-lane and pair swaps, half-row and
// row mirrors (every lane then holds its 16-lane row's sum), then row 0 -> 1 and 2 -> 3
// broadcasts of lane 15 and a rows 0-1 -> 2-3 broadcast of lane 31 leave the total in lane
// 63, which is read back as a scalar. All steps are VALU data-parallel moves (no LDS).
template <typename T>
__device__ __forceinline__ T wave_total(T v) {
  v += dpp_move<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
  v += dpp_move<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
  v += dpp_move<0x141, 0xF>(v);  // row_half_mirror
  v += dpp_move<0x140, 0xF>(v);  // row_mirror
  v += dpp_move<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
  return lane_bcast(v, 63);
}

// Sum over each group of LPP consecutive lanes (LPP = 1, 2 or 4), returned in every lane.
template <int LPP, typename T>
__device__ __forceinline__ T lane_group_sum(T v) {
  static_assert(LPP == 1 || LPP == 2 || LPP == 4 || LPP == 8, "groups of 1, 2, 4 or 8 lanes");
  if constexpr (LPP >= 2) v += dpp_move<0xB1, 0xF>(v);   // quad_perm [1, 0, 3, 2]
  if constexpr (LPP >= 4) v += dpp_move<0x4E, 0xF>(v);   // quad_perm [2, 3, 0, 1]
  if constexpr (LPP >= 8) v += dpp_move<0x141, 0xF>(v);  // row_half_mirror: the other quad
  return v;
}

// s_waitcnt vmcnt(0): every vector memory operation of the wave has completed. Placed where a
// loop's loads must be complete on every path, so that the compiler does not add waits for
// them at the top of the next iteration.
__device__ __forceinline__ void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Workgroup barrier that orders LDS only: global loads in flight (a prefetch) are not
// drained by it.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T>
struct AxisEntry {
  T g, r2;
};

// Stencil origin of one axis: unwrapped low corner and its wrapped grid index.
struct AxisOrigin {
  int lo;   // j - half (unwrapped)
  int low;  // mod(lo, M)
};

template <typename T>
__device__ __forceinline__ AxisOrigin axis_origin(T y, T inv_h, int half, int M) {
  AxisOrigin o;
  // single rounded product, then round-half-even: no contraction possible
  o.lo = (int)rint(y * inv_h) - half;
  int w = o.lo % M;
  o.low = w < 0 ? w + M : w;
  return o;
}

// axis_origin for a position already wrapped into [0, L): the unwrapped low corner then lies
// in [-half, M - half], so one conditional addition of M wraps it.
template <typename T>
__device__ __forceinline__ AxisOrigin axis_origin_wrapped(T y, T inv_h, int half, int M) {
  AxisOrigin o;
  o.lo = (int)rint(y * inv_h) - half;
  o.low = o.lo < 0 ? o.lo + M : o.lo;
  return o;
}

// axis_entry with the Gaussian as one hardware base-2 exponential, exp(-r2 / 2S^2) =
// 2^(-r2 log2(e) / 2S^2) (v_exp_f32; relative error a few 1e-7 over the stencil's range),
// for the Float32 interpolation kernels.
__device__ __forceinline__ AxisEntry<float> axis_entry_exp2(int lo, int k, float y, float h, float inv_norm,
                                                            float inv_2S2_log2e) {
  AxisEntry<float> e;
  const float x = (float)(lo + k) * h - y;
  e.r2 = x * x;
  e.g = inv_norm * __builtin_amdgcn_exp2f(-e.r2 * inv_2S2_log2e);
  return e;
}

template <typename T>
__device__ __forceinline__ AxisEntry<T> axis_entry(int lo, int k, T y, T h, T inv_norm, T inv_2S2) {
  AxisEntry<T> e;
  const T x = (T)(lo + k) * h - y;
  e.r2 = x * x;
  e.g = inv_norm * exp(-e.r2 * inv_2S2);
  return e;
}

// Spread binning of one slot: the wrapped stencil low corner on each axis and its home tile
// (the tile holding the low corner), counted into tile_count; slot_tile[i] gets (tile, rank in
// the tile) and origin[i] (low x, low y, low z, tile).
template <typename T>
__device__ __forceinline__ void bin_slot(const typename Vec<T>::v4& y, int i, const Scalars<T>& s, const Tiling& tl,
                                         int* __restrict__ tile_count, int2* __restrict__ slot_tile,
                                         int4* __restrict__ origin) {
  const AxisOrigin ox = axis_origin(y.x, s.inv_h, s.half, s.M[0]);
  const AxisOrigin oy = axis_origin(y.y, s.inv_h, s.half, s.M[1]);
  const AxisOrigin oz = axis_origin(y.z, s.inv_h, s.half, s.M[2]);
  const int t = ox.low / tl.T[0] + tl.nt[0] * (oy.low / tl.T[1] + tl.nt[1] * (oz.low / tl.T[2]));
  // rank in the tile. With many slots per tile (tl.wave_ranks), the lanes of a wave that share
  // a tile take one atomic between them and are ranked in lane order (consecutive sorted slots
  // mostly share tiles, and each tile's counter is contended); otherwise one atomic per slot
  const int lane = (int)(threadIdx.x & 63);
  int r = 0;
  if (!tl.wave_ranks) {
    r = atomicAdd(&tile_count[t], 1);
  } else for (;;) {
    const int lt = __builtin_amdgcn_readfirstlane(t);
    const unsigned long long same = __ballot(t == lt);
    if (t == lt) {
      const int first = __ffsll((long long)same) - 1;
      int base = 0;
      if (lane == first) base = atomicAdd(&tile_count[t], __popcll(same));
      base = __builtin_amdgcn_readlane(base, first);
      r = base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(same >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)same, 0u));
      break;
    }
  }
  slot_tile[i] = make_int2(t, r);
  origin[i] = make_int4(ox.low, oy.low, oz.low, t);
}

// First tile of a persistent block. Workgroups are dispatched round-robin over the 8 XCDs,
// each with its own L2; remapping the block index gives every XCD a contiguous run of tiles
// at each step (tiles b, b + gridDim, ...), so neighbouring tiles that share 128-byte lines
// (narrow column groups of the FFT's y and z passes, overlapping stencils of spread tiles) are fetched through the same L2.
// Block `id` of `n` (a block range of a grouped launch starts on a multiple of 8, so id % 8 is
// still the block's XCD).
__device__ __forceinline__ int xcd_first_tile(int id, int n) {
  constexpr int kXcds = 8;
  if (n % kXcds != 0) return id;
  return (id % kXcds) * (n / kXcds) + id / kXcds;
}
__device__ __forceinline__ int xcd_first_tile() { return xcd_first_tile((int)blockIdx.x, (int)gridDim.x); }

__device__ __forceinline__ int wrap_add(int base, int k, int M) {
  int i = base + k;
  return i >= M ? i - M : i;
}

}  // namespace ffcm
