// Random sequential addition of spherocylinders into a periodic cubic box
// (host only): the initial state of large synthetic workloads.  Candidate
// rods get a uniform centre and a uniform axis; a candidate is kept when its
// segment distance to every rod already placed is at least dmin.  Rods are
// binned into cells at least one reach (longest rod plus dmin) wide, so
// a candidate is tested against the 27 surrounding cells only, and only
// against rods whose centres lie within (L_i + L_j)/2 + dmin.  Should the
// acceptance rate collapse (dense boxes), axes are drawn ever closer to
// +-z.  Deterministic for a given seed: xoshiro256** seeded through
// splitmix64, Box-Muller normals.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/msc.h"

namespace {

struct Xoshiro {
  uint64_t s[4];
  explicit Xoshiro(uint64_t seed) {
    for (auto &w : s) {  // splitmix64
      seed += 0x9e3779b97f4a7c15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      w = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
  double normal() {
    double u1;
    do u1 = uniform(); while (u1 <= 0.0);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * uniform());
  }
};

double wrap(double d, double box) { return d - box * std::floor(d / box + 0.5); }

// squared minimum distance of two segments (closest points, clamped)
double seg2(const double rA[3], const double uA[3], double LA, const double rB[3], const double uB[3], double LB,
            double box) {
  double d[3], d1[3], d2[3], r[3];
  for (int a = 0; a < 3; ++a) {
    d[a] = wrap(rA[a] - rB[a], box);
    d1[a] = uA[a] * LA;
    d2[a] = uB[a] * LB;
    r[a] = d[a] - 0.5 * d1[a] + 0.5 * d2[a];
  }
  auto dot = [](const double *x, const double *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
  const double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r), b = dot(d1, d2), c = dot(d1, r);
  const double den = a * e - b * b;
  double s = den > 1e-12 ? std::clamp((b * f - c * e) / den, 0.0, 1.0) : 0.0;
  double t = e > 0 ? (b * s + f) / e : 0.0;
  if (t < 0) {
    t = 0;
    s = a > 0 ? std::clamp(-c / a, 0.0, 1.0) : 0.0;
  } else if (t > 1) {
    t = 1;
    s = a > 0 ? std::clamp((b - c) / a, 0.0, 1.0) : 0.0;
  }
  double v2 = 0;
  for (int q = 0; q < 3; ++q) {
    const double v = r[q] + s * d1[q] - t * d2[q];
    v2 += v * v;
  }
  return v2;
}

}  // namespace

extern "C" int msc_place_rods(int N, double box, const double *len, double dmin, unsigned long long seed,
                              double *r, double *u) {
  if (N < 1 || !(box > 0) || !len || !(dmin > 0) || !r || !u) return MSC_E_ARG;
  double lmax = 0;
  for (int i = 0; i < N; ++i) lmax = std::max(lmax, len[i]);
  const double reach = lmax + dmin;
  int nc = (int)std::floor(box / reach);
  if (nc < 3) nc = 1;
  nc = std::min(nc, 512);
  const double cw = box / nc;
  std::vector<std::vector<int>> cells((size_t)nc * nc * nc);
  auto cell_of = [&](const double x[3], int c[3]) {
    for (int a = 0; a < 3; ++a) c[a] = std::min(nc - 1, std::max(0, (int)std::floor((x[a] + 0.5 * box) / cw)));
  };
  Xoshiro rng(seed);
  double kappa = 0.0;
  long tried = 0, got = 0, total = 0;
  const long max_total = 4000L * N + 100000000L;
  for (int placed = 0; placed < N;) {
    if (++total > max_total) return MSC_E_CAPACITY;
    double x[3], v[3];
    for (int a = 0; a < 3; ++a) x[a] = (rng.uniform() - 0.5) * box;
    if (kappa == 0.0) {
      double n2 = 0;
      do {
        n2 = 0;
        for (int a = 0; a < 3; ++a) {
          v[a] = rng.normal();
          n2 += v[a] * v[a];
        }
      } while (n2 < 1e-20);
      for (int a = 0; a < 3; ++a) v[a] /= std::sqrt(n2);
    } else {  // von Mises-Fisher about +z, random polarity
      const double w = 1.0 + std::log(std::exp(-2 * kappa) + (1 - std::exp(-2 * kappa)) * rng.uniform()) / kappa;
      const double ph = 2 * M_PI * rng.uniform(), sw = std::sqrt(std::max(0.0, 1 - w * w));
      const double sg = rng.uniform() < 0.5 ? -1.0 : 1.0;
      v[0] = sg * sw * std::cos(ph);
      v[1] = sg * sw * std::sin(ph);
      v[2] = sg * w;
    }
    const double L = len[placed];
    int c[3];
    cell_of(x, c);
    bool ok = true;
    const int span = nc == 1 ? 0 : 1;
    for (int dz = -span; dz <= span && ok; ++dz)
      for (int dy = -span; dy <= span && ok; ++dy)
        for (int dx = -span; dx <= span && ok; ++dx) {
          const int cx = (c[0] + dx + nc) % nc, cy = (c[1] + dy + nc) % nc, cz = (c[2] + dz + nc) % nc;
          for (int j : cells[((size_t)cz * nc + cy) * nc + cx]) {
            double d2 = 0;
            for (int a = 0; a < 3; ++a) {
              const double d = wrap(x[a] - r[3 * j + a], box);
              d2 += d * d;
            }
            const double lim = 0.5 * (L + len[j]) + dmin;
            if (d2 >= lim * lim) continue;
            if (seg2(x, v, L, r + 3 * j, u + 3 * j, len[j], box) < dmin * dmin) {
              ok = false;
              break;
            }
          }
        }
    ++tried;
    if (ok) {
      for (int a = 0; a < 3; ++a) {
        r[3 * placed + a] = x[a];
        u[3 * placed + a] = v[a];
      }
      cells[((size_t)c[2] * nc + c[1]) * nc + c[0]].push_back(placed);
      ++placed;
      ++got;
    }
    if (tried >= 100000 && got < tried / 100) {  // acceptance below 1 %: narrow the axes
      kappa = kappa == 0.0 ? 2.0 : 2 * kappa;
      tried = got = 0;
    }
  }
  return MSC_OK;
}
