// datagen.cpp -- deterministic synthetic particle snapshots.
//
// Clustered generator: a fraction `frac` of the N points is drawn from
// `clusters` isotropic Gaussian clusters (centres uniform in [0, box)^d,
// sigma either fixed or lognormal(mu, s)); the rest are uniform in the box.
// Points are stored cluster by cluster (cluster sizes differ by at most one),
// uniform points last.  Every coordinate is a pure function of
// (seed, stream, index) through a splitmix64 counter hash, so the output is
// identical for any thread count and on any host with the same libm.
// Normals come from the Box-Muller transform.

#include <cmath>
#include <cstdint>
#include <cstring>

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t hash3(uint64_t seed, uint64_t stream, uint64_t idx) {
  return mix64(mix64(mix64(seed) ^ (stream * 0xD1B54A32D192ED03ull)) ^ idx);
}
inline double u01(uint64_t h) { return (h >> 11) * 0x1.0p-53; }
inline double normal(uint64_t h1, uint64_t h2) {
  const double u1 = ((h1 >> 11) + 1) * 0x1.0p-53;  // (0, 1]
  const double u2 = u01(h2);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

struct Spec {
  uint64_t n, clusters, seed;
  int dim;
  double box, frac, sigma, mu, s;  // sigma > 0: fixed; else lognormal(mu, s)
};

template <typename T>
void generate(const Spec& sp, T* x, T* y, T* z) {
  const uint64_t ncl = sp.clusters ? static_cast<uint64_t>(sp.n * sp.frac) : 0;
  const uint64_t C = sp.clusters ? sp.clusters : 1;
  const uint64_t q = ncl / C, r = ncl % C;
  T* P[3] = {x, y, z};
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < static_cast<int64_t>(sp.n); ++ii) {
    const uint64_t i = static_cast<uint64_t>(ii);
    if (i < ncl) {
      const uint64_t c = i < r * (q + 1) ? i / (q + 1) : r + (i - r * (q + 1)) / q;
      double sig = sp.sigma;
      if (!(sig > 0))
        sig = std::exp(sp.mu + sp.s * normal(hash3(sp.seed, 2, 2 * c),
                                            hash3(sp.seed, 2, 2 * c + 1)));
      for (int a = 0; a < sp.dim; ++a) {
        const double centre = sp.box * u01(hash3(sp.seed, 1, 3 * c + a));
        const uint64_t k = 2 * (3 * i + a);
        P[a][i] = static_cast<T>(centre + sig * normal(hash3(sp.seed, 3, k),
                                                       hash3(sp.seed, 3, k + 1)));
      }
    } else {
      for (int a = 0; a < sp.dim; ++a)
        P[a][i] = static_cast<T>(sp.box * u01(hash3(sp.seed, 4, 3 * i + a)));
    }
  }
}

}  // namespace

extern "C" {

// float_bits 32 or 64; z may be null when dim == 2.  sigma > 0 selects a
// fixed cluster sigma, otherwise sigma ~ exp(mu + s * N(0,1)) per cluster.
int gpz_gen_clustered(int float_bits, int dim, uint64_t n, uint64_t clusters,
                      double box, double frac, double sigma, double mu,
                      double s, uint64_t seed, void* x, void* y, void* z) {
  if ((dim != 2 && dim != 3) || (dim == 3 && !z) || !x || !y) return -1;
  Spec sp{n, clusters, seed, dim, box, frac, sigma, mu, s};
  if (float_bits == 64)
    generate<double>(sp, static_cast<double*>(x), static_cast<double*>(y),
                     static_cast<double*>(z));
  else if (float_bits == 32)
    generate<float>(sp, static_cast<float*>(x), static_cast<float*>(y),
                    static_cast<float*>(z));
  else
    return -1;
  return 0;
}

// Uniform double in [0,1) and standard normal from the same hash, exposed so
// that batch/serving workload mixes derive their parameters reproducibly.
double gpz_gen_uniform(uint64_t seed, uint64_t stream, uint64_t idx) {
  return u01(hash3(seed, stream, idx));
}

}  // extern "C"
