// Seeded list-mode events for the strip scanner from an elliptical phantom.
//
// Follows the reference phantom path (2d_strip_phantom):
//   region choice by CDF of intensity*area   src/2d/geometry/phantom.h:104-122,183-191
//   uniform point in a rotated ellipse       src/2d/geometry/ellipse.h:74-94
//   non-additive rejection (first on top)    src/2d/geometry/phantom.h:193-203
//   angle phi = pi/2 * U(-1,1), tan = tan(phi) phantom.h:210-214, strip/event.h:22-27
//   projection + detection |z_u|,|z_d| < L/2 src/2d/strip/scanner.h:107-113,195-214
//   Gaussian noise on z_u, z_d (sigma_z), dl (sigma_dl)  scanner.h:256-267
// but draws from a counter-based generator so the output does not depend on
// the thread count: slot i uses Philox4x32-10 (Salmon et al., SC'11) with
// key = seed and counter = (i, draw), retrying until an emission is
// detected.  Each slot holds one detected event, so exactly n are produced.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/petstrip.h"

namespace petstrip {
int set_error(int code, const char* msg);
}

namespace {

struct Philox {
  uint32_t k0, k1;
  static inline void mulhilo(uint32_t a, uint32_t b, uint32_t& hi,
                             uint32_t& lo) {
    uint64_t p = (uint64_t)a * b;
    hi = (uint32_t)(p >> 32);
    lo = (uint32_t)p;
  }
  // Philox4x32 with 10 rounds: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
  // increments 0x9E3779B9 / 0xBB67AE85.
  void operator()(const uint32_t in[4], uint32_t out[4]) const {
    uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint32_t a = k0, b = k1;
    for (int r = 0; r < 10; ++r) {
      uint32_t hi0, lo0, hi1, lo1;
      mulhilo(0xD2511F53u, c0, hi0, lo0);
      mulhilo(0xCD9E8D57u, c2, hi1, lo1);
      uint32_t n0 = hi1 ^ c1 ^ a, n1 = lo1, n2 = hi0 ^ c3 ^ b, n3 = lo0;
      c0 = n0; c1 = n1; c2 = n2; c3 = n3;
      a += 0x9E3779B9u;
      b += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
  }
};

inline double u01(uint32_t x) { return (x >> 8) * (1.0 / 16777216.0); }      // [0,1)
inline double u01_open(uint32_t x) { return ((x >> 8) + 1) * (1.0 / 16777216.0); }  // (0,1]

struct Region {
  double x, y, a, b, s, c;  // centre, semi-axes, sin/cos of rotation
  double A, B, C;           // A rx^2 + 2 C rx ry + B ry^2 <= 1 (ellipse.h:30-33,46-52)
  double cdf;
  bool contains(double px, double py) const {
    double rx = px - x, ry = py - y;
    return A * rx * rx + 2 * C * rx * ry + B * ry * ry <= 1;
  }
};

}  // namespace

extern "C" int petstrip_phantom_events(const petstrip_scanner* sc,
                                       const double* regions, int32_t n_regions,
                                       uint64_t seed, int64_t n, float* out,
                                       int n_threads) {
  using petstrip::set_error;
  if (!sc || n < 0 || (n > 0 && !out) || n_regions <= 0 || !regions)
    return set_error(PETSTRIP_EINVAL, "bad phantom arguments");
  std::vector<Region> reg((size_t)n_regions);
  double total = 0;
  for (int i = 0; i < n_regions; ++i) {
    const double* r = regions + 6 * i;
    Region& g = reg[(size_t)i];
    g.x = r[0];
    g.y = r[1];
    g.a = r[2];
    g.b = r[3];
    if (!(g.a > 0) || !(g.b > 0) || !(r[5] >= 0))
      return set_error(PETSTRIP_EINVAL,
                       "phantom region needs positive axes and intensity >= 0");
    double ang = r[4] * M_PI / 180;
    g.s = sin(ang);
    g.c = cos(ang);
    g.A = g.c * g.c / (g.a * g.a) + g.s * g.s / (g.b * g.b);
    g.B = g.s * g.s / (g.a * g.a) + g.c * g.c / (g.b * g.b);
    g.C = g.s * g.c * (1 / (g.a * g.a) - 1 / (g.b * g.b));
    total += r[5] * M_PI * g.a * g.b;  // weight = intensity * area
    g.cdf = total;
  }
  if (!(total > 0))
    return set_error(PETSTRIP_EINVAL, "phantom has zero total intensity");
  for (auto& g : reg) g.cdf /= total;
  const double R = sc->radius, half_L = sc->scintillator_length / 2;
  const double sz = sc->sigma_z, sdl = sc->sigma_dl;
  const Philox rng{(uint32_t)seed, (uint32_t)(seed >> 32)};

  if (n_threads < 1) n_threads = 1;
  n_threads = (int)std::min<int64_t>(n_threads, std::max<int64_t>(n / 4096, 1));
  std::vector<int> status((size_t)n_threads, 0);
  auto work = [&](int tid) {
    int64_t lo = n * tid / n_threads, hi = n * (tid + 1) / n_threads;
    for (int64_t i = lo; i < hi; ++i) {
      uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0, 0};
      uint32_t r[4];
      bool done = false;
      for (uint32_t draw = 0; draw < (1u << 24) && !done; ++draw) {
        ctr[2] = 2 * draw;
        rng(ctr, r);
        // region by CDF (first with u <= cdf)
        double u = u01(r[0]);
        size_t k = 0;
        while (k + 1 < reg.size() && u > reg[k].cdf) ++k;
        const Region& g = reg[k];
        double ang = 2 * M_PI * u01(r[1]);
        double rad = sqrt(u01(r[2]));
        double ex = g.a * rad * cos(ang), ey = g.b * rad * sin(ang);
        double pz = g.c * ex - g.s * ey + g.x;
        double py = g.s * ex + g.c * ey + g.y;
        bool covered = false;
        for (size_t j = 0; j < k && !covered; ++j)
          covered = reg[j].contains(pz, py);
        if (covered) continue;
        double phi = M_PI_2 * (2 * u01(r[3]) - 1);
        double t = tan(phi);
        double zu = pz + (R - py) * t;
        double zd = pz - (R + py) * t;
        double dl = -2 * py * sqrt(t * t + 1);
        if (!(fabs(zu) < half_L && fabs(zd) < half_L)) continue;
        ctr[2] = 2 * draw + 1;
        rng(ctr, r);
        // Box-Muller: two normals from (r0, r1), one from (r2, r3)
        double m0 = sqrt(-2 * log(u01_open(r[0])));
        double m1 = sqrt(-2 * log(u01_open(r[2])));
        double n0 = m0 * cos(2 * M_PI * u01(r[1]));
        double n1 = m0 * sin(2 * M_PI * u01(r[1]));
        double n2 = m1 * cos(2 * M_PI * u01(r[3]));
        out[3 * i] = (float)(zu + sz * n0);
        out[3 * i + 1] = (float)(zd + sz * n1);
        out[3 * i + 2] = (float)(dl + sdl * n2);
        done = true;
      }
      if (!done) {
        status[(size_t)tid] = 1;
        return;
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < n_threads; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& t : th) t.join();
  for (int s : status)
    if (s)
      return set_error(PETSTRIP_ESTATE,
                       "no detected emission after 2^24 draws in a slot");
  return PETSTRIP_OK;
}
