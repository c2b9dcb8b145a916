// Phantom emissions and detected strip events, shared by the host generator
// (threads) and the device generator (one lane per slot).
//
// Follows the reference phantom path (2d_strip_phantom):
//   region choice by CDF of the weights       src/2d/geometry/phantom.h:118-133,183-191
//     (ellipse: intensity*area, rectangle: intensity*4ab, point: intensity)
//   uniform point in a rotated ellipse        src/2d/geometry/ellipse.h:74-94
//   uniform point in a rectangle              src/2d/geometry/rectangle.h:34-52
//   non-additive rejection (first on top)     src/2d/geometry/phantom.h:193-203
//   angle phi = pi/2 * U(-1,1), tan = tan(phi) phantom.h:210-214, strip/event.h:22-27
//   projection + detection |z_u|,|z_d| < L/2  src/2d/strip/scanner.h:107-113,195-214
//   Gaussian noise on z_u, z_d (sigma_z), dl (sigma_dl)  scanner.h:256-267
// but draws from a counter-based generator so the output does not depend on
// the thread count or the device: slot i uses Philox4x32-10 (Salmon et al.,
// SC'11) with key = seed and counter = (i, 2 draw), the noise of a detected
// emission counter (i, 2 draw + 1).  An emission is the first draw whose
// point is not covered by an earlier region (every draw when additive).  In
// the detected mode a slot makes emissions until one is detected (one event
// per slot, exactly n events); in the emission mode a slot makes one
// emission, detected or not.  The arithmetic is double precision; host and
// device differ only where their libm results differ in the last bit.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PH_HD __host__ __device__ __forceinline__
#else
#define PH_HD static inline
#endif

namespace petstrip {

// Largest number of draws a slot makes before giving up.
constexpr uint32_t kPhantomMaxDraws = 1u << 16;
constexpr int kPhantomMaxRegions = 64;

enum { kRegionEllipse = 0, kRegionRectangle = 1, kRegionPoint = 2 };

struct PhantomRegion {
  double x, y, a, b, s, c;  // centre, semi-axes (half-sides), sin/cos of rotation
  double A, B, C;           // ellipse: A rx^2 + 2 C rx ry + B ry^2 <= 1 (ellipse.h:30-33,46-52)
  double cdf;               // cumulative weight, normalised
  int kind;                 // kRegion*
};

struct PhantomSpec {
  PhantomRegion reg[kPhantomMaxRegions];
  int n_regions;
  int additive;  // regions overlap additively (no rejection)
  double R, half_L, sz, sdl;
  uint32_t k0, k1;  // Philox key (seed)
};

// Philox4x32 with 10 rounds: multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key
// increments 0x9E3779B9 / 0xBB67AE85.
PH_HD void philox4x32_10(uint32_t k0, uint32_t k1, const uint32_t in[4], uint32_t out[4]) {
  uint32_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
  uint32_t a = k0, b = k1;
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c1 ^ a, n1 = lo1, n2 = hi0 ^ c3 ^ b, n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
    a += 0x9E3779B9u;
    b += 0xBB67AE85u;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

PH_HD double ph_u01(uint32_t x) { return (x >> 8) * (1.0 / 16777216.0); }  // [0,1)
PH_HD double ph_u01_open(uint32_t x) {                                      // (0,1]
  return ((x >> 8) + 1) * (1.0 / 16777216.0);
}

PH_HD bool ph_contains(const PhantomRegion& g, double px, double py) {
  const double rx = px - g.x, ry = py - g.y;
  if (g.kind == kRegionRectangle) return fabs(rx) <= g.a && fabs(ry) <= g.b;
  if (g.kind == kRegionPoint) return px == g.x && py == g.y;
  return g.A * rx * rx + 2 * g.C * rx * ry + g.B * ry * ry <= 1;
}

// The next emission of the slot whose counter is ctr, from draw *draw on:
// point (px, py) and angle phi; *draw is left at the emission's draw.  False
// when kPhantomMaxDraws draws found none.
PH_HD bool ph_next_emission(const PhantomSpec& p, uint32_t ctr[4], uint32_t* draw,
                            double* px, double* py, double* phi) {
  const double kPi = 3.14159265358979323846;
  uint32_t r[4];
  for (; *draw < kPhantomMaxDraws; ++*draw) {
    ctr[2] = 2 * *draw;
    philox4x32_10(p.k0, p.k1, ctr, r);
    // region by CDF (first with u <= cdf)
    const double u = ph_u01(r[0]);
    int k = 0;
    while (k + 1 < p.n_regions && u > p.reg[k].cdf) ++k;
    const PhantomRegion& g = p.reg[k];
    double x, y;
    if (g.kind == kRegionRectangle) {
      x = g.x + (2 * ph_u01(r[1]) - 1) * g.a;
      y = g.y + (2 * ph_u01(r[2]) - 1) * g.b;
    } else if (g.kind == kRegionPoint) {
      x = g.x;
      y = g.y;
    } else {
      const double ang = 2 * kPi * ph_u01(r[1]);
      const double rad = sqrt(ph_u01(r[2]));
      const double ex = g.a * rad * cos(ang), ey = g.b * rad * sin(ang);
      x = g.c * ex - g.s * ey + g.x;
      y = g.s * ex + g.c * ey + g.y;
    }
    bool covered = false;
    if (!p.additive)
      for (int j = 0; j < k && !covered; ++j) covered = ph_contains(p.reg[j], x, y);
    if (covered) continue;
    *px = x;
    *py = y;
    *phi = (kPi / 2) * (2 * ph_u01(r[3]) - 1);
    return true;
  }
  return false;
}

// Projection of an emission onto the strips (scanner.h:107-113): the exact
// response and whether both ends lie on the strips.
PH_HD bool ph_detect(const PhantomSpec& p, double px, double py, double phi,
                     double exact[3]) {
  const double t = tan(phi);
  exact[0] = px + (p.R - py) * t;
  exact[1] = px - (p.R + py) * t;
  exact[2] = -2 * py * sqrt(t * t + 1);
  return fabs(exact[0]) < p.half_L && fabs(exact[1]) < p.half_L;
}

// Noise of the emission at draw `draw` (scanner.h:256-267), Box-Muller: two
// normals from (r0, r1), one from (r2, r3).
PH_HD void ph_noise(const PhantomSpec& p, uint32_t ctr[4], uint32_t draw,
                    const double exact[3], float out[3]) {
  const double kPi = 3.14159265358979323846;
  uint32_t r[4];
  ctr[2] = 2 * draw + 1;
  philox4x32_10(p.k0, p.k1, ctr, r);
  const double m0 = sqrt(-2 * log(ph_u01_open(r[0])));
  const double m1 = sqrt(-2 * log(ph_u01_open(r[2])));
  const double n0 = m0 * cos(2 * kPi * ph_u01(r[1]));
  const double n1 = m0 * sin(2 * kPi * ph_u01(r[1]));
  const double n2 = m1 * cos(2 * kPi * ph_u01(r[3]));
  out[0] = (float)(exact[0] + p.sz * n0);
  out[1] = (float)(exact[1] + p.sz * n1);
  out[2] = (float)(exact[2] + p.sdl * n2);
}

// Slot i's event (z_u, z_d, dl) in the detected mode; false if no emission
// was detected within kPhantomMaxDraws draws.
PH_HD bool phantom_slot(const PhantomSpec& p, int64_t i, float out[3]) {
  uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0, 0};
  double px, py, phi, exact[3];
  for (uint32_t draw = 0; ph_next_emission(p, ctr, &draw, &px, &py, &phi); ++draw) {
    if (!ph_detect(p, px, py, phi, exact)) continue;
    ph_noise(p, ctr, draw, exact, out);
    return true;
  }
  return false;
}

// Image grid of the generator's histograms: PixelGrid semantics
// (src/2d/geometry/pixel_grid.h: column = floor((x - lower_left_x) / size),
// row = floor((y - lower_left_y) / size), row 0 at the bottom), and the tan
// planes of the tan-bins map (src/3d/geometry/voxel_grid.h:78-121,
// plane = floor((tan + max_tan) / (2 max_tan / tan_bins))).
struct PhantomGrid {
  int n_columns, n_rows;
  double pixel_size, llx, lly;
  int tan_bins;
  double max_tan;
};

PH_HD int ph_pixel(const PhantomGrid& g, double x, double y) {
  const double cx = floor((x - g.llx) / g.pixel_size);
  const double cy = floor((y - g.lly) / g.pixel_size);
  if (!(cx >= 0 && cx < g.n_columns && cy >= 0 && cy < g.n_rows)) return -1;
  return (int)cy * g.n_columns + (int)cx;
}

}  // namespace petstrip
