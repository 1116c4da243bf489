// Synthetic suspensions for the benchmark and test workloads (host code, no GPU).
//
// Positions: random sequential addition (RSA) of unit spheres in a cubic periodic box: a
// candidate centre is rejected when its minimum-image distance to an accepted centre is below
// 2 (= 2a). The first N - n_clustered centres are uniform; the rest are drawn around
// n_clusters Gaussian cluster centres (uniform in the box, isotropic standard deviation
// cluster_std, wrapped into the box), cluster i % n_clusters for particle i. A clustered
// particle that is rejected max_cluster_tries times is drawn uniformly instead (counted).
// Forces: i.i.d. N(0, 1) per component with the per-component mean subtracted.
//
// Random numbers: SplitMix64 seeding xoshiro256**, doubles from the top 53 bits, normals by
// Box-Muller (both outputs used), so results are identical on every platform.
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

double wrap(double x, double L) {
  double r = std::fmod(x, L);
  if (r < 0) r += L;
  if (r >= L) r -= L;
  return r;
}

struct Grid {
  int n;
  double cs, L;
  std::vector<int64_t> head, next;
  Grid(double L_, int64_t N) : L(L_) {
    n = (int)std::floor(L / 2.0);
    if (n < 1) n = 1;
    cs = L / n;
    head.assign((size_t)n * n * n, -1);
    next.assign((size_t)N, -1);
  }
  int cell(double x) const {
    int c = (int)(x / cs);
    return c >= n ? n - 1 : (c < 0 ? 0 : c);
  }
  bool free_at(const double* Y, const double* p) const {
    const int c[3] = {cell(p[0]), cell(p[1]), cell(p[2])};
    int nb[3][3];
    int cnt = n >= 3 ? 3 : n;  // distinct neighbour cells per axis
    for (int a = 0; a < 3; ++a) {
      if (n >= 3) {
        nb[a][0] = c[a] == 0 ? n - 1 : c[a] - 1;
        nb[a][1] = c[a];
        nb[a][2] = c[a] == n - 1 ? 0 : c[a] + 1;
      } else {
        for (int k = 0; k < n; ++k) nb[a][k] = k;
      }
    }
    for (int iz = 0; iz < cnt; ++iz)
      for (int iy = 0; iy < cnt; ++iy) {
        const size_t row = (size_t)n * (nb[1][iy] + (size_t)n * nb[2][iz]);
        for (int ix = 0; ix < cnt; ++ix) {
          for (int64_t j = head[row + nb[0][ix]]; j >= 0; j = next[j]) {
            double r2 = 0;
            for (int a = 0; a < 3; ++a) {
              double d = p[a] - Y[3 * j + a];
              if (d > 0.5 * L)
                d -= L;
              else if (d < -0.5 * L)
                d += L;
              r2 += d * d;
            }
            if (r2 < 4.0) return false;
          }
        }
      }
    return true;
  }
  void insert(int64_t i, const double* p) {
    const size_t c = (size_t)cell(p[0]) + (size_t)n * (cell(p[1]) + (size_t)n * cell(p[2]));
    next[i] = head[c];
    head[c] = i;
  }
};

}  // namespace

extern "C" {

// Returns 0 on success, 1 when a uniform draw could not be placed after 10^7 tries (box too
// full). Y receives 3N doubles (x, y, z per particle).
int ffcm_rsa_suspension(int64_t N, double L, int n_clusters, int64_t n_clustered, double cluster_std,
                        uint64_t seed, int max_cluster_tries, double* Y, int64_t* fallbacks) {
  Rng rng(seed);
  Grid g(L, N);
  std::vector<double> centres(3 * (size_t)(n_clusters > 0 ? n_clusters : 0));
  for (int c = 0; c < n_clusters; ++c)
    for (int a = 0; a < 3; ++a) centres[3 * c + a] = rng.uniform() * L;
  int64_t fb = 0;
  const int64_t n_uniform = N - n_clustered;
  double p[3];
  for (int64_t i = 0; i < N; ++i) {
    bool placed = false;
    if (i >= n_uniform && n_clusters > 0) {
      const int c = (int)((i - n_uniform) % n_clusters);
      for (int t = 0; t < max_cluster_tries && !placed; ++t) {
        for (int a = 0; a < 3; ++a) p[a] = wrap(centres[3 * c + a] + cluster_std * rng.normal(), L);
        placed = g.free_at(Y, p);
      }
      if (!placed) ++fb;
    }
    for (int64_t t = 0; !placed; ++t) {
      if (t > 10000000) return 1;
      for (int a = 0; a < 3; ++a) p[a] = rng.uniform() * L;
      placed = g.free_at(Y, p);
    }
    for (int a = 0; a < 3; ++a) Y[3 * i + a] = p[a];
    g.insert(i, p);
  }
  if (fallbacks) *fallbacks = fb;
  return 0;
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

// Uniform doubles in [lo, hi) (used for log-uniform system sizes of the ensemble workload).
int ffcm_uniform(int64_t n, uint64_t seed, double lo, double hi, double* out) {
  Rng rng(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = lo + (hi - lo) * rng.uniform();
  return 0;
}
}
