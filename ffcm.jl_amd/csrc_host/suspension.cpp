// Synthetic suspensions for the benchmark and test workloads (host code, no GPU).
//
// Positions: random sequential addition (RSA) of unit spheres in a cubic periodic box: a
// candidate centre is rejected when its minimum-image distance to an accepted centre is below
// 2 (= 2a). The first N - n_clustered centres are uniform; the rest are drawn around
// n_clusters Gaussian cluster centres (uniform in the box, isotropic standard deviation
// cluster_std, wrapped into the box), filled one after another: clustered particle j of n
// goes to cluster floor(j * n_clusters / n). A clustered
// particle that is rejected max_cluster_tries times is drawn uniformly instead (counted).
// Forces: i.i.d. N(0, 1) per component with the per-component mean subtracted.
//
// Random numbers: SplitMix64 seeding xoshiro256**, doubles from the top 53 bits, normals by
// Box-Muller (both outputs used), so results are identical on every platform. The cluster
// draws come from a counter-based stream instead: draw t of particle i is a hash of (seed, i,
// t), so the draws of one particle can be tested in parallel and the first accepted one (in
// draw order) is the same for any thread count. Dense clusters reject thousands of draws per
// particle; with max_cluster_tries = 10,000 the 2^20-particle suspension needs about 2e9 tests.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; ++i) {
      x += 0x9E3779B97F4A7C15ull;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
  bool have_spare = false;
  double spare = 0;
  double normal() {
    if (have_spare) {
      have_spare = false;
      return spare;
    }
    double u1 = uniform();
    double u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double t = 6.283185307179586476925286766559 * u2;
    spare = r * std::sin(t);
    have_spare = true;
    return r * std::cos(t);
  }
};

uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Three N(0, 1) values of draw t of particle i: Marsaglia's polar method on pairs of
// uniforms in (-1, 1) from the counter stream of (i, t) (a pair outside the unit disc is
// skipped; the stream of one draw is 2^16 values long, far more than it ever uses).
void counter_normal3(uint64_t key, int64_t i, int64_t t, double out[3]) {
  uint64_t c = mix64(key + (uint64_t)i * 0xD1B54A32D192ED03ull) + ((uint64_t)t << 16);
  int got = 0;
  while (got < 3) {
    const double x = (double)(int64_t)mix64(c++) * 0x1.0p-63;
    const double y = (double)(int64_t)mix64(c++) * 0x1.0p-63;
    const double s = x * x + y * y;
    if (s >= 1.0 || s == 0.0) continue;
    const double f = std::sqrt(-2.0 * (double)std::log((float)s) / s);
    out[got++] = x * f;
    if (got < 3) out[got++] = y * f;
  }
}

double wrap(double x, double L) {
  double r = std::fmod(x, L);
  if (r < 0) r += L;
  if (r >= L) r -= L;
  return r;
}

// Accepted centres binned into cells of side cs >= 2 (n per axis), stored inline per cell: a
// candidate is tested against the 27 cells around its own. Pairwise distances of at least 2
// bound a cell's count: 8 when cs < 4/sqrt(3) (two of nine points would share one of the
// eight half-side subcubes, whose diagonal is below 2), else 27 (thirds, cs < 3 when n >= 3).
struct Grid {
  int n, cap;
  double cs, L;
  std::vector<int> count;
  std::vector<double> pos;  // [cell][cap][3]
  Grid(double L_, int64_t) : L(L_) {
    n = (int)std::floor(L / 2.0);
    if (n < 1) n = 1;
    cs = L / n;
    cap = cs < 2.3094010767585 ? 8 : 27;
    if (n < 3) cap = 64;  // tiny boxes: each cell spans most of the box
    count.assign((size_t)n * n * n, 0);
    pos.assign((size_t)n * n * n * cap * 3, 0.0);
  }
  int cell(double x) const {
    int c = (int)(x / cs);
    return c >= n ? n - 1 : (c < 0 ? 0 : c);
  }
  bool free_at(const double* p) const {
    const int c[3] = {cell(p[0]), cell(p[1]), cell(p[2])};
    int nb[3][3];
    int cnt = n >= 3 ? 3 : n;  // distinct neighbour cells per axis
    for (int a = 0; a < 3; ++a) {
      if (n >= 3) {
        nb[a][0] = c[a];
        nb[a][1] = c[a] == 0 ? n - 1 : c[a] - 1;
        nb[a][2] = c[a] == n - 1 ? 0 : c[a] + 1;
      } else {
        for (int k = 0; k < n; ++k) nb[a][k] = k;
      }
    }
    const double hL = 0.5 * L;
    for (int iz = 0; iz < cnt; ++iz)
      for (int iy = 0; iy < cnt; ++iy) {
        const size_t row = (size_t)n * (nb[1][iy] + (size_t)n * nb[2][iz]);
        for (int ix = 0; ix < cnt; ++ix) {
          const size_t cc = row + nb[0][ix];
          const double* q = &pos[cc * cap * 3];
          for (int j = 0, m = count[cc]; j < m; ++j, q += 3) {
            double r2 = 0;
            for (int a = 0; a < 3; ++a) {
              double d = p[a] - q[a];
              if (d > hL)
                d -= L;
              else if (d < -hL)
                d += L;
              r2 += d * d;
            }
            if (r2 < 4.0) return false;
          }
        }
      }
    return true;
  }
  bool insert(const double* p) {
    const size_t c = (size_t)cell(p[0]) + (size_t)n * (cell(p[1]) + (size_t)n * cell(p[2]));
    if (count[c] >= cap) return false;
    double* q = &pos[(c * cap + count[c]) * 3];
    for (int a = 0; a < 3; ++a) q[a] = p[a];
    ++count[c];
    return true;
  }
};

// Sense-counting spin barrier for the team of one parallel region (every thread of the
// region runs on its own core for the whole clustered phase, and a wait is a few hundred
// nanoseconds, against microseconds for a sleeping barrier; there are about four per
// clustered particle).
struct SpinBarrier {
  std::atomic<int> count{0};
  std::atomic<int> gen{0};
  int n = 1;
  void wait() {
    const int g0 = gen.load(std::memory_order_acquire);
    if (count.fetch_add(1, std::memory_order_acq_rel) == n - 1) {
      count.store(0, std::memory_order_relaxed);
      gen.store(g0 + 1, std::memory_order_release);
    } else {
      while (gen.load(std::memory_order_acquire) == g0) __builtin_ia32_pause();
    }
  }
};

// Fine cells (side about 1/3) that lie entirely within distance 2 of an accepted centre,
// one bit each: a candidate in such a cell is rejected without the neighbour search. This
// only skips work; the accepted draws are the same as with the search alone. A cell is
// marked when its farthest corner is closer than 2 - 1e-9, which also absorbs the rounding of
// the candidate's cell index.
struct Cover {
  int nf;
  double fs, L;
  std::vector<uint64_t> bits;
  explicit Cover(double L_) : L(L_) {
    nf = std::max(1, (int)std::floor(L * 3.0));
    fs = L / nf;
    bits.assign(((size_t)nf * nf * nf + 63) / 64, 0);
  }
  size_t index(int ix, int iy, int iz) const { return (size_t)ix + (size_t)nf * (iy + (size_t)nf * iz); }
  int cell(double x) const {
    int c = (int)(x / fs);
    return c >= nf ? nf - 1 : (c < 0 ? 0 : c);
  }
  bool covered(const double* q) const {
    const size_t i = index(cell(q[0]), cell(q[1]), cell(q[2]));
    return (bits[i >> 6] >> (i & 63)) & 1;
  }
  void mark(const double* p) {
    const double lim = (2.0 - 1e-9) * (2.0 - 1e-9);
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = (int)std::floor((p[a] - 2.0) / fs);
      hi[a] = (int)std::floor((p[a] + 2.0) / fs);
    }
    auto far2 = [&](int a, int i) {
      const double c0 = i * fs, c1 = c0 + fs;
      const double d = std::max(std::fabs(p[a] - c0), std::fabs(p[a] - c1));
      return d * d;
    };
    auto wrapi = [&](int i) { return ((i % nf) + nf) % nf; };
    for (int iz = lo[2]; iz <= hi[2]; ++iz) {
      const double dz = far2(2, iz);
      if (dz >= lim) continue;
      for (int iy = lo[1]; iy <= hi[1]; ++iy) {
        const double dy = dz + far2(1, iy);
        if (dy >= lim) continue;
        for (int ix = lo[0]; ix <= hi[0]; ++ix) {
          if (dy + far2(0, ix) >= lim) continue;
          const size_t i = index(wrapi(ix), wrapi(iy), wrapi(iz));
          bits[i >> 6] |= uint64_t(1) << (i & 63);
        }
      }
    }
  }
};

}  // namespace

extern "C" {

// Returns 0 on success, 1 when a uniform draw could not be placed after 10^7 tries (box too
// full). Y receives 3N doubles (x, y, z per particle).
// Returns 0 on success, 1 when a uniform draw could not be placed after 10^7 tries (box too
// full). Y receives 3N doubles (x, y, z per particle).
int ffcm_rsa_suspension(int64_t N, double L, int n_clusters, int64_t n_clustered, double cluster_std,
                        uint64_t seed, int max_cluster_tries, double* Y, int64_t* fallbacks) {
  Rng rng(seed);
  Grid g(L, N);
  std::vector<double> centres(3 * (size_t)(n_clusters > 0 ? n_clusters : 0));
  for (int c = 0; c < n_clusters; ++c)
    for (int a = 0; a < 3; ++a) centres[3 * c + a] = rng.uniform() * L;
  const int64_t n_uniform = n_clusters > 0 ? N - n_clustered : N;
  const uint64_t cluster_key = mix64(seed ^ 0xC1B57E2D0F3A6C49ull);
  auto cluster_draw = [&](int64_t i, int64_t t, double q[3]) {
    const int c = (int)((i - n_uniform) * n_clusters / (N - n_uniform));
    double z[3];
    counter_normal3(cluster_key, i, t, z);
    for (int a = 0; a < 3; ++a) {
      double x = centres[3 * c + a] + cluster_std * z[a];
      if (x < 0 && x >= -L)
        x += L;
      else if (x >= L && x < 2 * L)
        x -= L;
      else if (x < 0 || x >= L)
        x = wrap(x, L);
      q[a] = x < L ? x : 0.0;
    }
  };
  int status = 0;
  int64_t fb = 0;
  // one placement: the given point, else uniform draws from the sequential stream
  auto place = [&](int64_t i, const double* first) -> bool {
    double p[3];
    bool placed = first != nullptr;
    if (placed)
      for (int a = 0; a < 3; ++a) p[a] = first[a];
    for (int64_t t = 0; !placed; ++t) {
      if (t > 10000000) return false;
      for (int a = 0; a < 3; ++a) p[a] = rng.uniform() * L;
      placed = g.free_at(p);
    }
    for (int a = 0; a < 3; ++a) Y[3 * i + a] = p[a];
    return g.insert(p);
  };
  for (int64_t i = 0; i < n_uniform; ++i)
    if (!place(i, nullptr)) return 1;
  if (n_uniform < N) {
    Cover cov(L);
    for (int64_t i = 0; i < n_uniform; ++i) cov.mark(&Y[3 * i]);
    // Clustered particles: all threads test draws of the current particle in rounds; thread
    // k of T tests draws t0 + k, t0 + k + T, ... (up to `per` of them) and stops at its first
    // free one or once a smaller free draw is known. The smallest free draw wins, which is
    // the first free draw in draw order whatever T is. Rounds grow from 4 draws per thread to
    // 1024.
    // one result slot per parity of i: thread 0 resets the next particle's slot while the
    // other threads may still read this particle's
    std::atomic<int64_t> best_slot[2];
    best_slot[0].store(max_cluster_tries);
    best_slot[1].store(max_cluster_tries);
    SpinBarrier bar;
#pragma omp parallel
    {
      const int T = omp_get_num_threads(), k = omp_get_thread_num();
#pragma omp single
      bar.n = T;  // implicit barrier
      for (int64_t i = n_uniform; i < N; ++i) {
        std::atomic<int64_t>& best = best_slot[i & 1];
        int64_t t0 = 0;
        for (int per = 4; t0 < max_cluster_tries; per = per < 1024 ? 8 * per : per) {
          const int64_t t1 = std::min<int64_t>(t0 + (int64_t)per * T, max_cluster_tries);
          for (int64_t t = t0 + k; t < t1; t += T) {
            if (t > best.load(std::memory_order_relaxed)) break;
            double q[3];
            cluster_draw(i, t, q);
            if (!cov.covered(q) && g.free_at(q)) {
              int64_t b = best.load(std::memory_order_relaxed);
              while (t < b && !best.compare_exchange_weak(b, t, std::memory_order_relaxed)) {
              }
              break;
            }
          }
          bar.wait();
          // every store to best of this round precedes the barrier; the next stores are of
          // later draws (>= t1), and its reset comes one particle later
          const bool done = best.load(std::memory_order_relaxed) < t1;
          t0 = t1;
          if (done) break;
        }
        if (k == 0) {
          double p[3];
          const int64_t b = best.load(std::memory_order_relaxed);
          const bool hit = b < max_cluster_tries;
          if (hit) cluster_draw(i, b, p);
          if (!hit) ++fb;
          if (!place(i, hit ? p : nullptr)) status = 1;
          cov.mark(&Y[3 * i]);
          best_slot[(i + 1) & 1].store(max_cluster_tries, std::memory_order_relaxed);
        }
        bar.wait();  // the grid holds particle i before any thread tests i + 1
        if (status) break;
      }
    }
  }
  if (fallbacks) *fallbacks = fb;
  return status;
}

// n_rhs sets of i.i.d. N(0,1) forces (3N each, x, y, z per particle), each with its
// per-component mean subtracted.
int ffcm_normal_forces(int64_t N, int n_rhs, uint64_t seed, double* F) {
  Rng rng(seed);
  for (int r = 0; r < n_rhs; ++r) {
    double* f = F + (size_t)r * 3 * N;
    double mean[3] = {0, 0, 0};
    for (int64_t i = 0; i < 3 * N; ++i) {
      f[i] = rng.normal();
      mean[i % 3] += f[i];
    }
    for (int a = 0; a < 3; ++a) mean[a] /= (double)N;
    for (int64_t i = 0; i < 3 * N; ++i) f[i] -= mean[i % 3];
  }
  return 0;
}

// Threads of the clustered draws (the default is OpenMP's, OMP_NUM_THREADS); the positions do
// not depend on it.
int ffcm_suspension_set_threads(int n) {
  omp_set_num_threads(n < 1 ? 1 : n);
  return 0;
}

// Uniform doubles in [lo, hi) (used for log-uniform system sizes of the ensemble workload).
int ffcm_uniform(int64_t n, uint64_t seed, double lo, double hi, double* out) {
  Rng rng(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * rng.uniform();
  return 0;
}
}
