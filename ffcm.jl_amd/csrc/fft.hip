// Step 4 of the mobility apply: the Stokes solve as a hand-written 3-D real FFT.
//
// Reference: stokes_solve! (src/stokes_solve.jl:21-42) with FFTW plan_rfft / plan_brfft
// (src/config.jl:418-423) and the projection _apply_inverse_stokes_kernel!
// (src/stokes_solve.jl:70-110). Conventions kept from the reference:
//   - forward transforms use exp(-2 pi i jk/n), unnormalised; the backward transform is the
//     unnormalised c2r and the 1/(Mx My Mz) factor lives in the projection scalar inv_M/mu;
//   - the projection is f <- (inv_M/mu * 1/k^2) (f - k (k.f)/k^2) with k = 0 set to exactly 0
//     and the wavevectors passed in verbatim (Nyquist entries are +pi/h);
//   - the c2r treats each x row as Hermitian, using the real parts of its DC and (even Mx)
//     Nyquist bins only, after complex inverse passes along z and y.
//
// Structure (five HBM passes over the spectrum):
//   x_fwd   real rows -> half-spectrum rows; two real rows share one complex transform
//   y       complex transforms along y (forward)
//   z       forward along z, projection of the three components, inverse along z, fused
//   y       complex transforms along y (inverse)
//   x_inv   half-spectrum rows -> real rows, two rows per complex transform
// Each block stages its pencils in LDS; the transforms are Stockham radix-2/3/4/5/7/8/11/13
// stages executed from registers with the LDS array as the exchange buffer.
#include <algorithm>
#include <cmath>
#include <type_traits>

#include <cstdlib>

#include "fft_kernels.hpp"

namespace ffcm {

using namespace fftk;

bool plan_axis(int n, FftAxis* ax) {
  ax->n = n;
  ax->nstages = 0;
  int r = n;
  const int order[] = {8, 4, 2, 3, 5, 7, 11, 13};
  ax->packed = 0;
  for (int f : order) {
    while (r % f == 0 && ax->nstages < kMaxStages) {
      ax->packed |= (unsigned long long)f << (4 * ax->nstages);
      ax->radix[ax->nstages++] = f;
      r /= f;
    }
  }
  return r == 1;
}

namespace {

// Kernel table for a transform of length n: the compile-time plan when one is instantiated
// (Float32), else the runtime-radix plan.
template <typename T>
FftKernels<T> kernels_for(int n) {
  if constexpr (sizeof(T) == 4) {
    FftKernels<float> k;
    if (static_kernels_part0(n, &k) || static_kernels_part1(n, &k) || static_kernels_part2(n, &k) ||
        static_kernels_part3(n, &k))
      return k;
  }
  return make_kernels<T, DynPlan>();
}

// Threads per pencil: the compile-time plan's own (a power of two, so that its stages
// synchronise per wave), else the registers-per-stage bound of the runtime plan.
template <typename T>
int pencil_threads(const FftAxis& ax, const FftKernels<T>& k) {
  return k.is_static ? k.tp : threads_per_pencil(ax);
}

int log2i(int w) {
  int l = 0;
  while ((1 << l) < w) ++l;
  return l;
}

// Blocks of a persistent kernel that are resident at once (occupancy query).
// Persistent grid of a pass: blocks resident per CU (occupancy), at most `cap` (0: no cap;
// FFCM_FFT_BLOCKS_PER_CU overrides every pass's cap for tuning), times the CUs.
int resident_blocks(const void* kernel, int threads, int lds, int cus, int cap) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess || per_cu < 1) {
    (void)hipGetLastError();
    per_cu = 1;
  }
  if (const char* e = std::getenv("FFCM_FFT_BLOCKS_PER_CU")) cap = std::atoi(e);
  if (cap > 0 && per_cu > cap) per_cu = cap;
  return per_cu * cus;
}
}  // namespace

template <typename T>
void plan_passes(ffcm_ctx* c, int lds_budget) {
  const int Mx = c->host.M[0], My = c->host.M[1], Mz = c->host.M[2];
  const int Mh = Mx / 2 + 1;
  const int cs = (int)sizeof(C2<T>);
  const FftKernels<T> kx = kernels_for<T>(Mx), ky = kernels_for<T>(My), kz = kernels_for<T>(Mz);
  const int tpx = pencil_threads(c->fft[0], kx);
  const int tpy = pencil_threads(c->fft[1], ky);
  const int tpz = pencil_threads(c->fft[2], kz);
  // Large tiles keep up to ~140 KiB of each block's next tile in flight and make the strided
  // segments of the y and z passes long (256 B and 192 B at 512^3).
  const int lds_cap = lds_budget < 144 * 1024 ? lds_budget : 144 * 1024;
  int cus = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess) cus = 256;
  (void)hipGetLastError();
  // threads: up to `most` (512; the wide z-pass kernel kZWideThreads), a whole number of
  // pencils per round, and (for the fetch addressing) a multiple of `mult`
  auto threads_for = [](int pencils, int tp, int mult = 1, int most = 512) {
    for (int ng = std::min(most / tp, pencils); ng >= 1; --ng)
      if ((ng * tp) % mult == 0) return ng * tp;
    return tp;  // tp * pencils is always a multiple of the pencil count
  };
  // x passes: P row pairs (3P pencils); each row pair is fetched in ceil(row / threads) chunks
  for (int inverse = 0; inverse < 2; ++inverse) {
    const int pairs = (My + 1) / 2;
    const int row = inverse ? 3 * Mh : 3 * Mx;
    const int slots = inverse ? kx.ept / 2 + 1 : kx.ept;
    int P = 16;
    if (P > pairs) P = pairs;
    auto fits = [&](int P) {
      const int thr = threads_for(3 * P, tpx);
      const int U1 = (row + thr - 1) / thr;
      return (3 * P * pencil_stride(Mx, kx.padded) + Mx) * cs <= lds_cap && P * U1 <= slots;
    };
    while (P > 1 && !fits(P)) --P;
    FftPass fp;
    fp.width = P;
    fp.log2w = 0;
    fp.pencils = 3 * P;
    fp.tp = tpx;
    fp.threads = threads_for(3 * P, tpx);
    fp.lds_bytes = (3 * P * pencil_stride(Mx, kx.padded) + Mx) * cs;
    fp.tiles = Mz * ((pairs + P - 1) / P);
    fp.covered = fits(P);
    const void* kern = inverse ? (const void*)kx.x_inv : (const void*)kx.x_fwd;
    // blocks per CU: the forward pass (real rows in, padded half-spectrum rows out) is fastest
    // at two (measured at 256^3 and 512^3: 0.110 / 0.745 ms against 0.131 / 0.880 at the
    // occupancy limit), the inverse at the occupancy limit (0.716 ms against 0.964 at two)
    fp.resident = resident_blocks(kern, fp.threads, fp.lds_bytes, cus, inverse ? 0 : 2);
    fp.blocks = std::min(fp.resident, fp.tiles);
    (inverse ? c->pass_xi : c->pass_xf) = fp;
  }
  // y passes: W columns (power of two)
  {
    const int Q = 3 * Mh;
    int W = 32;
    while (W > 1 && ((W * pencil_stride(My, ky.padded) + My) * cs > lds_cap || W * My > ky.ept_strided * threads_for(W, tpy, W)))
      W >>= 1;
    FftPass fp;
    fp.width = W;
    fp.log2w = log2i(W);
    fp.pencils = W;
    fp.tp = tpy;
    fp.threads = threads_for(W, tpy, W);
    fp.lds_bytes = (W * pencil_stride(My, ky.padded) + My) * cs;
    fp.tiles = Mz * ((Q + W - 1) / W);
    fp.covered = W * My <= ky.ept_strided * fp.threads &&
                 (int64_t)(My + fp.threads / W) * c->Mhp * 3 < (int64_t(1) << 32);  // 32-bit fetch offsets
    // one block per CU once a tile's prefetch reaches 64 KiB (512-point columns: 0.69 ms
    // against 0.77 at two), the occupancy limit for smaller tiles
    const int tile_bytes = W * My * cs;
    fp.resident = resident_blocks((const void*)ky.y_fwd, fp.threads, fp.lds_bytes, cus, tile_bytes >= 65536 ? 1 : 0);
    fp.blocks = std::min(fp.resident, fp.tiles);
    c->pass_y = fp;
  }
  // z pass: W kx columns x 3 components (W a power of two). The plans with a wide z-pass
  // kernel (512 points in Float32, 32 threads per pencil) may take up to kZWideThreads threads:
  // at 512^3 that is 8 columns (24 pencils, 768 threads, 106 KB of LDS, one block per CU), whose
  // 192-byte kz rows leave two neighbouring tiles sharing one 128-byte line, where 4 columns
  // (96-byte rows, four tiles over three lines) fetched 1.65x the bytes the pass reads.
  {
    const int kslots = z_pass_k_slots<T>(Mh, Mz);
    const int zmost = (kz.z_wide && !std::getenv("FFCM_FFT_Z_NARROW")) ? kZWideThreads : 512;
    int W = 16;
    while (W > 1 && ((3 * W * pencil_stride(Mz, kz.padded) + Mz + kslots) * cs > lds_cap ||
                     3 * W * Mz > kz.ept_strided * threads_for(3 * W, tpz, 3 * W, zmost)))
      W >>= 1;
    FftPass fp;
    fp.width = W;
    fp.log2w = log2i(W);
    fp.pencils = 3 * W;
    fp.tp = tpz;
    fp.threads = threads_for(3 * W, tpz, 3 * W, zmost);
    fp.wide = fp.threads > 512;
    fp.lds_bytes = (3 * W * pencil_stride(Mz, kz.padded) + Mz + kslots) * cs;
    fp.tiles = My * ((Mh + W - 1) / W);
    fp.covered = 3 * W * Mz <= kz.ept_strided * fp.threads &&
                 (int64_t)(Mz + fp.threads / (3 * W)) * My * c->Mhp * 3 < (int64_t(1) << 32);
    // two blocks per CU (256^3: 0.127 ms against 0.149 at the occupancy limit; 512^3: equal)
    fp.resident = resident_blocks(fp.wide ? (const void*)kz.z_wide : (const void*)kz.z, fp.threads, fp.lds_bytes,
                                  cus, 2);
    fp.blocks = std::min(fp.resident, fp.tiles);
    c->pass_z = fp;
  }
}

namespace {
// mode bit 0: forward transforms (x, y, z); bit 3: the projection; bit 1: backward transforms
// (z, y, x). The z pass runs whatever of its three steps the mode names.
// nb grids of the context's shape are transformed together (consecutive in the force, spectrum
// and velocity buffers): for the x and y passes they are nb * Mz planes; the z pass decodes
// the grid index from its tile. Persistent grids: min(resident blocks, tiles of all nb grids).
template <typename T>
hipError_t run_fft(ffcm_ctx* c, int mode, T* out, int nb) {
  const Scalars<T>& s = *(const Scalars<T>*)c->scalars;
  const int Mx = c->host.M[0], My = c->host.M[1], Mz = c->host.M[2];
  const int Mh = c->Mh, Mzb = Mz * nb;
  const FftPass& xf = c->pass_xf;
  const FftPass& py = c->pass_y;
  const FftPass& pz = c->pass_z;
  const FftPass& xi = c->pass_xi;
  const FftKernels<T> kx = kernels_for<T>(Mx), ky = kernels_for<T>(My), kz = kernels_for<T>(Mz);
  auto grid = [nb](const FftPass& p) {
    const long long t = (long long)p.tiles * nb;
    return (unsigned)(t < p.resident ? t : p.resident);
  };
  C2<T>* spec = (C2<T>*)c->d.spectrum;
  const C2<T>* tw[3] = {(const C2<T>*)c->d.twiddle[0], (const C2<T>*)c->d.twiddle[1], (const C2<T>*)c->d.twiddle[2]};
  // persistent passes: the grid is the resident block count; the tiles behind it are the work
  auto work = [nb](const FftPass& p) { return (int64_t)p.tiles * nb; };
  hipError_t e = hipSuccess;
  if (mode & 1) {
    const FftXArgs<T> xa{(T*)c->d.force, spec, c->fft[0], My, Mzb, xf.width, xf.tp, c->Mhp, tw[0]};
    e = submit(c, kTagFft + 0, kx.x_fwd, kx.x_fwd_g, grid(xf), xf.threads, xf.lds_bytes, xa, work(xf));
    const FftYArgs<T> ya{spec, c->fft[1], 3 * Mh, 3 * c->Mhp, py.width, py.log2w, py.tp, Mzb, tw[1]};
    if (e == hipSuccess) e = submit(c, kTagFft + 1, ky.y_fwd, ky.y_fwd_g, grid(py), py.threads, py.lds_bytes, ya, work(py));
  }
  const FftZArgs<T> za{spec, c->fft[2], My, Mh, c->Mhp, pz.width, pz.log2w, pz.tp, tw[2], (const T*)c->d.k_x,
                       (const T*)c->d.k_y, (const T*)c->d.k_z, s.inv_mu_M, mode, nb};
  if (e == hipSuccess)
    e = pz.wide ? submit(c, kTagFft + 2, kz.z_wide, kz.z_wide_g, grid(pz), pz.threads, pz.lds_bytes, za, work(pz))
                : submit(c, kTagFft + 2, kz.z, kz.z_g, grid(pz), pz.threads, pz.lds_bytes, za, work(pz));
  if (mode & 2) {
    const FftYArgs<T> ya{spec, c->fft[1], 3 * Mh, 3 * c->Mhp, py.width, py.log2w, py.tp, Mzb, tw[1]};
    if (e == hipSuccess) e = submit(c, kTagFft + 3, ky.y_inv, ky.y_inv_g, grid(py), py.threads, py.lds_bytes, ya, work(py));
    const FftXArgs<T> xa{out ? out : (T*)c->d.velocity, spec, c->fft[0], My, Mzb, xi.width, xi.tp, c->Mhp, tw[0]};
    if (e == hipSuccess) e = submit(c, kTagFft + 4, kx.x_inv, kx.x_inv_g, grid(xi), xi.threads, xi.lds_bytes, xa, work(xi));
  }
  return e;
}
}  // namespace

// Full Stokes solve: force grid -> velocity grid.
template <typename T>
hipError_t launch_fft_stokes(ffcm_ctx* c, T* out, int nb) {
  return run_fft<T>(c, kFftForward | kFftProject | kFftInverse, out, nb);
}

// Forward transforms + projection only; the spectrum is left projected, in the state the
// reference's backward transforms read (src/stokes_solve.jl:33-40).
template <typename T>
hipError_t launch_fft_forward_project(ffcm_ctx* c) {
  return run_fft<T>(c, kFftForward | kFftProject, (T*)nullptr, 1);
}

// Any of the three steps alone or together (kFft* bits), on the context's own grids and
// spectrum: the reference's separate mul! and projection calls (src/stokes_solve.jl:24-40).
template <typename T>
hipError_t launch_fft_steps(ffcm_ctx* c, int steps) {
  return run_fft<T>(c, steps, (T*)nullptr, 1);
}

// Backward transforms alone, from the spectrum buffer (src/stokes_solve.jl:38-40).
template <typename T>
hipError_t launch_fft_inverse(ffcm_ctx* c) {
  return run_fft<T>(c, kFftInverse, (T*)nullptr, 1);
}

template void plan_passes<float>(ffcm_ctx*, int);
template void plan_passes<double>(ffcm_ctx*, int);
template hipError_t launch_fft_stokes<float>(ffcm_ctx*, float*, int);
template hipError_t launch_fft_stokes<double>(ffcm_ctx*, double*, int);
template hipError_t launch_fft_forward_project<float>(ffcm_ctx*);
template hipError_t launch_fft_forward_project<double>(ffcm_ctx*);
template hipError_t launch_fft_inverse<float>(ffcm_ctx*);
template hipError_t launch_fft_inverse<double>(ffcm_ctx*);
template hipError_t launch_fft_steps<float>(ffcm_ctx*, int);
template hipError_t launch_fft_steps<double>(ffcm_ctx*, int);

}  // namespace ffcm
