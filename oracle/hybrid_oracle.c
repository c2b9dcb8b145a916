/* CPU oracle of the 3D hybrid reconstruction; see hybrid_oracle.h.
 * Test infrastructure only.  Each function cites the reference lines it
 * restates (grey277/j-pet-mlem, src/...). */
#include "hybrid_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* LineSegment (2d/geometry/line_segment.h:21-31,62-84), in floats and in
 * doubles. */
typedef struct {
  float sx, sy, mx, my, dx, dy, nx, ny, len, dfo;
} seg32;
typedef struct {
  double sx, sy, mx, my, dx, dy, nx, ny, len, dfo;
} seg64;

static seg32 seg_f(const float* s) {
  seg32 g;
  g.sx = s[0];
  g.sy = s[1];
  g.mx = (s[0] + s[2]) / 2.0f;
  g.my = (s[1] + s[3]) / 2.0f;
  const float vx = s[2] - s[0], vy = s[3] - s[1];
  const float l = sqrtf(vx * vx + vy * vy); /* Vector::length */
  g.dx = vx / l;                            /* Vector::normalize divides */
  g.dy = vy / l;
  g.nx = -g.dy; /* perpendicular: (-y, x) */
  g.ny = g.dx;
  g.len = l;
  g.dfo = g.sx * g.nx + g.sy * g.ny;
  return g;
}
static seg64 seg_d(const float* s) {
  seg64 g;
  g.sx = s[0];
  g.sy = s[1];
  g.mx = ((double)s[0] + s[2]) / 2.0;
  g.my = ((double)s[1] + s[3]) / 2.0;
  const double vx = (double)s[2] - s[0], vy = (double)s[3] - s[1];
  const double l = sqrt(vx * vx + vy * vy);
  g.dx = vx / l;
  g.dy = vy / l;
  g.nx = -g.dy;
  g.ny = g.dx;
  g.len = l;
  g.dfo = g.sx * g.nx + g.sy * g.ny;
  return g;
}

/* PixelGrid::center_at (2d/geometry/pixel_grid.h:22,54-57) */
typedef struct {
  float llcx, llcy, ps;
  int32_t w, h;
} grid32;
static grid32 grid_f(const oracle_hybrid_geometry* g) {
  grid32 r;
  r.ps = (float)g->pixel_size;
  r.llcx = (float)g->lower_left_x + r.ps / 2.0f;
  r.llcy = (float)g->lower_left_y + r.ps / 2.0f;
  r.w = g->n_columns;
  r.h = g->n_rows;
  return r;
}
static inline void center_f(const grid32* r, int16_t x, int16_t y, float* cx, float* cy) {
  *cx = r->llcx + (float)x * r->ps;
  *cy = r->llcy + (float)y * r->ps;
}
static inline void center_d(const oracle_hybrid_geometry* g, int16_t x, int16_t y,
                            double* cx, double* cy) {
  *cx = g->lower_left_x + g->pixel_size / 2 + x * g->pixel_size;
  *cy = g->lower_left_y + g->pixel_size / 2 + y * g->pixel_size;
}

static inline int64_t lor_index(int16_t first, int16_t second) {
  return (int64_t)first * (first + 1) / 2 + second;
}

/* ------------------------------------------------------------------------ */
/* calculate_weight, calculate_sensitivity, normalize_geometry_weights
 * (3d/hybrid/reconstruction.h:212-265). */
void oracle_hybrid_weights_f32(const oracle_hybrid_geometry* g, int sensitivity_one,
                               float* weights, float* sensitivity) {
  const grid32 r = grid_f(g);
  const int64_t n_lors = (int64_t)(g->n_detectors + 1) * g->n_detectors / 2;
  for (int64_t l = 0; l < n_lors; ++l) {
    const seg32 s = seg_f(g->segments + 4 * l);
    const float sw = 0.3f * g->widths[l];
    /* a double normalisation times a float exponential, stored as float */
    const double norm = 1 / (sw * sqrt(2 * M_PI));
    const float inv2 = 1 / (2 * sw * sw);
    for (int64_t i = g->info_begin[l]; i < g->info_end[l]; ++i) {
      float cx, cy;
      center_f(&r, g->pixels[2 * i], g->pixels[2 * i + 1], &cx, &cy);
      const float d = (cx * s.nx + cy * s.ny) - s.dfo;
      weights[i] = (float)(norm * expf(-d * d * inv2));
    }
  }
  const size_t np = (size_t)g->n_columns * g->n_rows;
  /* sensitivity_one = 2: the caller's map (an external volume's plane 0,
   * reconstruction_cmd.cpp:505-519) normalises the weights as it is */
  if (sensitivity_one != 2)
    for (size_t p = 0; p < np; ++p) sensitivity[p] = sensitivity_one ? 1.0f : 0.0f;
  if (!sensitivity_one)
    for (int64_t i = 0; i < g->n_pixel_infos; ++i)
      sensitivity[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1]] +=
          weights[i];
  for (int64_t i = 0; i < g->n_pixel_infos; ++i)
    weights[i] /= sensitivity[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1]];
}

void oracle_hybrid_weights_f64(const oracle_hybrid_geometry* g, int sensitivity_one,
                               double* weights, double* sensitivity) {
  const int64_t n_lors = (int64_t)(g->n_detectors + 1) * g->n_detectors / 2;
  for (int64_t l = 0; l < n_lors; ++l) {
    const seg64 s = seg_d(g->segments + 4 * l);
    const double sw = 0.3 * g->widths[l];
    const double norm = 1 / (sw * sqrt(2 * M_PI));
    const double inv2 = 1 / (2 * sw * sw);
    for (int64_t i = g->info_begin[l]; i < g->info_end[l]; ++i) {
      double cx, cy;
      center_d(g, g->pixels[2 * i], g->pixels[2 * i + 1], &cx, &cy);
      const double d = (cx * s.nx + cy * s.ny) - s.dfo;
      weights[i] = norm * exp(-d * d * inv2);
    }
  }
  const size_t np = (size_t)g->n_columns * g->n_rows;
  /* sensitivity_one = 2: the caller's map (an external volume's plane 0,
   * reconstruction_cmd.cpp:505-519) normalises the weights as it is */
  if (sensitivity_one != 2)
    for (size_t p = 0; p < np; ++p) sensitivity[p] = sensitivity_one ? 1.0 : 0.0;
  if (!sensitivity_one)
    for (int64_t i = 0; i < g->n_pixel_infos; ++i)
      sensitivity[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1]] +=
          weights[i];
  for (int64_t i = 0; i < g->n_pixel_infos; ++i)
    weights[i] /= sensitivity[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1]];
}

/* ------------------------------------------------------------------------ */
/* The strip kernel (2d/strip/gaussian_kernel.h:27-33,51-79,93-105,140-149)
 * without contraction. */
typedef struct {
  float isz, isdl, c;
} k32;
static k32 kernel_consts_f(double sigma_z, double sigma_dl) {
  k32 k;
  const float sz = (float)sigma_z, sdl = (float)sigma_dl;
  k.isz = 1 / (sz * sz);
  k.isdl = 1 / (sdl * sdl);
  k.c = (float)(1 / (2 * M_PI * M_PI) * (double)sqrtf(k.isz * k.isz * k.isdl));
  return k;
}

static inline float kernel_f(const k32* k, float y, float t, float sec, float R, float dz,
                             float dy) {
  const float sec_sq = sec * sec;
  const float yp = dy + y;
  const float ap = -(yp - R) * sec_sq, aq = -(yp + R) * sec_sq;
  const float ar = -2 * yp * sec * t;
  const float pq = dz - dy * t;
  const float br = -2 * dy * sec;
  const float a_a = ap * k->isz * ap + aq * k->isz * aq + ar * k->isdl * ar;
  const float b_b = pq * k->isz * pq + pq * k->isz * pq + br * k->isdl * br;
  return (k->c / sqrtf(a_a)) * expf(-0.5f * b_b);
}
static inline double kernel_d(double isz, double isdl, double c, double y, double t,
                              double sec, double R, double dz, double dy) {
  const double sec_sq = sec * sec;
  const double yp = dy + y;
  const double ap = -(yp - R) * sec_sq, aq = -(yp + R) * sec_sq;
  const double ar = -2 * yp * sec * t;
  const double pq = dz - dy * t;
  const double br = -2 * dy * sec;
  const double a_a = ap * isz * ap + aq * isz * aq + ar * isdl * ar;
  const double b_b = pq * isz * pq + pq * isz * pq + br * isdl * br;
  return (c / sqrt(a_a)) * exp(-0.5 * b_b);
}

static inline float frame_sens_f(float z, float y, float R, float L) {
  const float L2 = L / 2;
  const float a = (L2 - z) / (R - y), b = (L2 + z) / (R + y);
  const float c = -(L2 + z) / (R - y), d = (-L2 + z) / (R + y);
  /* std::min and std::max (the host branch of util/cuda/compat.h:76-77) */
  const float tmax = atanf(b < a ? b : a);
  const float tmin = atanf(c < d ? d : c);
  return (tmax - tmin) / (float)M_PI;
}
static inline double frame_sens_d(double z, double y, double R, double L) {
  const double L2 = L / 2;
  const double a = (L2 - z) / (R - y), b = (L2 + z) / (R + y);
  const double c = -(L2 + z) / (R - y), d = (-L2 + z) / (R + y);
  return (atan(b < a ? b : a) - atan(c < d ? d : c)) / M_PI;
}

/* ------------------------------------------------------------------------ */
/* translate_to_frame and bb_intersects_grid_with_positive_weight
 * (reconstruction.h:97-161), Response::calculate_tan_y_z (2d/strip/
 * response.h:22-37), GaussianKernel::ellipse_bb (gaussian_kernel.h:108-128,
 * 159-160). */
static int64_t lower_bound_f(const float* a, int64_t lo, int64_t hi, float v) {
  while (lo < hi) {
    const int64_t m = lo + (hi - lo) / 2;
    if (a[m] < v)
      lo = m + 1;
    else
      hi = m;
  }
  return lo;
}
static int64_t upper_bound_f(const float* a, int64_t lo, int64_t hi, float v) {
  while (lo < hi) {
    const int64_t m = lo + (hi - lo) / 2;
    if (!(v < a[m]))
      lo = m + 1;
    else
      hi = m;
  }
  return lo;
}

int64_t oracle_hybrid_frames_f32(const oracle_hybrid_geometry* g, double sigma_z,
                                 double sigma_dl, const float* weights, const int16_t* lors,
                                 const float* zzd, int64_t n, oracle_hybrid_event* out) {
  const k32 k = kernel_consts_f(sigma_z, sigma_dl);
  const float ps = (float)g->pixel_size, z_left = (float)g->z_left;
  int64_t kept = 0;
  for (int64_t e = 0; e < n; ++e) {
    oracle_hybrid_event ev;
    ev.first = lors[2 * e];
    ev.second = lors[2 * e + 1];
    const int64_t l = lor_index(ev.first, ev.second);
    const seg32 s = seg_f(g->segments + 4 * l);
    const float R = s.len / 2;
    const float zu = zzd[3 * e], zd = zzd[3 * e + 1], dl = zzd[3 * e + 2];
    ev.tan = (zu - zd) / (2 * R);
    ev.up = -0.5f * (dl / sqrtf(1 + ev.tan * ev.tan));
    ev.right = 0.5f * (zu + zd + (2 * ev.up * ev.tan));
    const float tsq = ev.tan * ev.tan;
    ev.sec = sqrtf(1 + ev.tan * ev.tan);
    const float A = 4 * ev.sec * ev.sec * k.isdl + 2 * tsq * k.isz;
    const float B = -4 * ev.tan * k.isz;
    const float C = 2 * k.isz;
    const float B_2 = (B / 2) * (B / 2);
    const float half_up = 3 / sqrtf(A - (B_2 / C));
    const float half_right = 3 / sqrtf(C - (B_2 / A));
    const float zl = ev.right - half_right, zr = ev.right + half_right;
    /* plane(z): truncation toward zero to short */
    const int16_t pl = (int16_t)((zl - z_left) / ps);
    const int16_t pr = (int16_t)((zr - z_left) / ps);
    ev.plane_begin = pl > 0 ? pl : 0;
    ev.plane_end = (int16_t)(pr + 1) < g->n_planes ? (int16_t)(pr + 1) : (int16_t)g->n_planes;
    const float y_up = ev.up + half_up, y_dn = ev.up - half_up;
    const float t_up = (y_up + R) / (2 * R), t_dn = (y_dn + R) / (2 * R);
    /* one past the first position above t_up: may reach the next LOR's
     * first pixel info (reconstruction.h:133-138) */
    ev.end = upper_bound_f(g->positions, g->info_begin[l], g->info_end[l], t_up) + 1;
    ev.begin = lower_bound_f(g->positions, g->info_begin[l], g->info_end[l], t_dn);
    /* the grid filter */
    int keep = 0;
    if (!(ev.plane_end <= 0 || ev.plane_begin >= g->n_planes)) {
      for (int64_t i = ev.begin; i < ev.end && i < g->n_pixel_infos; ++i) {
        const int16_t x = g->pixels[2 * i], y = g->pixels[2 * i + 1];
        if (x >= 0 && x < g->n_columns && y >= 0 && y < g->n_rows && weights[i] > 0) {
          keep = 1;
          break;
        }
      }
    }
    if (keep) out[kept++] = ev;
  }
  return kept;
}

/* ------------------------------------------------------------------------ */
/* Index of the sensitivity factor of voxel (pix, iz): the pixel's entry of
 * the W x H map, or plane |iz - D| of an external volume of D planes
 * (reconstruction.h:357-362); |iz - D| = D, past the volume, reads D - 1. */
static inline size_t sens_index(size_t pix, size_t np, int iz, int depth) {
  if (depth <= 0) return pix;
  int p = iz < depth ? depth - iz : iz - depth;
  if (p > depth - 1) p = depth - 1;
  return pix + np * (size_t)p;
}

/* The iteration (reconstruction.h:267-445) with one thread: per event the
 * denominator over (pixel info, plane) in that order, each voxel's weight
 * kept in a voxel-indexed cache, then every (pixel info, plane) adds its
 * voxel's cached weight over the denominator; rho becomes the sum. */
/* weight_planes > 1: per-plane weights (weight_planes x n_pixel_infos),
 * voxel plane iz reading plane |iz - D| of them as of the sensitivity
 * volume (D = sens_depth = weight_planes, reconstruction.h:345-356). */
void oracle_hybrid_mlem_planes_f32(const oracle_hybrid_geometry* g, double sigma_z,
                                   double sigma_dl, double length, const float* weights,
                                   int weight_planes, const float* sensitivity,
                                   int sens_depth, const oracle_hybrid_event* ev, int64_t n,
                                   int n_iter, const float* rho0, float* out) {
  const k32 k = kernel_consts_f(sigma_z, sigma_dl);
  const grid32 r = grid_f(g);
  const float L = (float)length, z_left = (float)g->z_left;
  const size_t np = (size_t)g->n_columns * g->n_rows;
  const size_t nv = np * g->n_planes;
  float* rho = malloc(nv * sizeof(float));
  float* next = malloc(nv * sizeof(float));
  float* cache = calloc(nv, sizeof(float));
  memcpy(rho, rho0, nv * sizeof(float));
  for (int it = 0; it < n_iter; ++it) {
    memset(next, 0, nv * sizeof(float));
    for (int64_t e = 0; e < n; ++e) {
      const oracle_hybrid_event* v = &ev[e];
      const seg32 s = seg_f(g->segments + 4 * lor_index(v->first, v->second));
      const float R = s.len / 2;
      float denom = 0;
      for (int64_t i = v->begin; i < v->end; ++i) {
        const int16_t px = g->pixels[2 * i], py = g->pixels[2 * i + 1];
        const size_t pix = (size_t)px + (size_t)g->n_columns * py;
        float cx, cy;
        center_f(&r, px, py, &cx, &cy);
        const float up = (cx - s.mx) * s.dx + (cy - s.my) * s.dy;
        for (int iz = v->plane_begin; iz < v->plane_end; ++iz) {
          const size_t vox = pix + (size_t)iz * np;
          const float z = ((float)iz + 0.5f) * r.ps + z_left;
          const float kk = kernel_f(&k, v->up, v->tan, v->sec, R, z - v->right, up - v->up);
          const float k2 = kk / frame_sens_f(z, up, R, L);
          if (k2 <= 0) {
            cache[vox] = 0;
            continue;
          }
          const size_t si = sens_index(pix, np, iz, sens_depth);
          const float w =
              weight_planes > 1 ? weights[(si / np) * (size_t)g->n_pixel_infos + i] : weights[i];
          const float wt = k2 * w * rho[vox];
          denom += wt * sensitivity[si];
          cache[vox] = wt;
        }
      }
      if (!(denom > 0)) continue;
      const float inv = 1 / denom;
      for (int64_t i = v->begin; i < v->end; ++i) {
        const size_t pix =
            (size_t)g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1];
        for (int iz = v->plane_begin; iz < v->plane_end; ++iz) {
          const size_t vox = pix + (size_t)iz * np;
          next[vox] += cache[vox] * inv;
        }
      }
    }
    memcpy(rho, next, nv * sizeof(float));
    memcpy(out + (size_t)it * nv, rho, nv * sizeof(float));
  }
  free(rho);
  free(next);
  free(cache);
}

void oracle_hybrid_mlem_f32(const oracle_hybrid_geometry* g, double sigma_z,
                            double sigma_dl, double length, const float* weights,
                            const float* sensitivity, int sens_depth,
                            const oracle_hybrid_event* ev, int64_t n, int n_iter,
                            const float* rho0, float* out) {
  oracle_hybrid_mlem_planes_f32(g, sigma_z, sigma_dl, length, weights, 1, sensitivity,
                                sens_depth, ev, n, n_iter, rho0, out);
}

void oracle_hybrid_mlem_planes_f64(const oracle_hybrid_geometry* g, double sigma_z,
                                   double sigma_dl, double length, const double* weights,
                                   int weight_planes, const double* sensitivity,
                                   int sens_depth, const oracle_hybrid_event* ev, int64_t n,
                                   int n_iter, const double* rho0, double* out) {
  const double isz = 1 / (sigma_z * sigma_z), isdl = 1 / (sigma_dl * sigma_dl);
  const double c = 1 / (2 * M_PI * M_PI) * sqrt(isz * isz * isdl);
  const size_t np = (size_t)g->n_columns * g->n_rows;
  const size_t nv = np * g->n_planes;
  double* rho = malloc(nv * sizeof(double));
  double* next = malloc(nv * sizeof(double));
  double* cache = calloc(nv, sizeof(double));
  memcpy(rho, rho0, nv * sizeof(double));
  for (int it = 0; it < n_iter; ++it) {
    memset(next, 0, nv * sizeof(double));
    for (int64_t e = 0; e < n; ++e) {
      const oracle_hybrid_event* v = &ev[e];
      const seg64 s = seg_d(g->segments + 4 * lor_index(v->first, v->second));
      const double R = s.len / 2;
      double denom = 0;
      for (int64_t i = v->begin; i < v->end; ++i) {
        const int16_t px = g->pixels[2 * i], py = g->pixels[2 * i + 1];
        const size_t pix = (size_t)px + (size_t)g->n_columns * py;
        double cx, cy;
        center_d(g, px, py, &cx, &cy);
        const double up = (cx - s.mx) * s.dx + (cy - s.my) * s.dy;
        for (int iz = v->plane_begin; iz < v->plane_end; ++iz) {
          const size_t vox = pix + (size_t)iz * np;
          const double z = (iz + 0.5) * g->pixel_size + g->z_left;
          const double kk = kernel_d(isz, isdl, c, v->up, v->tan, v->sec, R, z - v->right,
                                     up - v->up);
          const double k2 = kk / frame_sens_d(z, up, R, length);
          if (k2 <= 0) {
            cache[vox] = 0;
            continue;
          }
          const size_t si = sens_index(pix, np, iz, sens_depth);
          const double w =
              weight_planes > 1 ? weights[(si / np) * (size_t)g->n_pixel_infos + i] : weights[i];
          const double wt = k2 * w * rho[vox];
          denom += wt * sensitivity[si];
          cache[vox] = wt;
        }
      }
      if (!(denom > 0)) continue;
      for (int64_t i = v->begin; i < v->end; ++i) {
        const size_t pix =
            (size_t)g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1];
        for (int iz = v->plane_begin; iz < v->plane_end; ++iz) {
          const size_t vox = pix + (size_t)iz * np;
          next[vox] += cache[vox] / denom;
        }
      }
    }
    memcpy(rho, next, nv * sizeof(double));
    memcpy(out + (size_t)it * nv, rho, nv * sizeof(double));
  }
  free(rho);
  free(next);
  free(cache);
}

void oracle_hybrid_mlem_f64(const oracle_hybrid_geometry* g, double sigma_z,
                            double sigma_dl, double length, const double* weights,
                            const double* sensitivity, int sens_depth,
                            const oracle_hybrid_event* ev, int64_t n, int n_iter,
                            const double* rho0, double* out) {
  oracle_hybrid_mlem_planes_f64(g, sigma_z, sigma_dl, length, weights, 1, sensitivity,
                                sens_depth, ev, n, n_iter, rho0, out);
}

/* Per-plane weights (reconstruction.h:236-271 over n_planes_half planes):
 * the sensitivity volume per (pixel, plane) as the sum of the plane's
 * weights in pixel-info order (sens_mode 0), one (1) or given (2, left in
 * sensitivity as is), then each plane's weights divided by it. */
void oracle_hybrid_plane_weights_f32(const oracle_hybrid_geometry* g, const float* raw,
                                     int planes, int sens_mode, float* weights,
                                     float* sensitivity) {
  const size_t np = (size_t)g->n_columns * g->n_rows, ni = (size_t)g->n_pixel_infos;
  if (sens_mode != 2)
    for (size_t k = 0; k < np * (size_t)planes; ++k) sensitivity[k] = sens_mode ? 1.0f : 0.0f;
  if (sens_mode == 0)
    for (size_t i = 0; i < ni; ++i)
      for (int p = 0; p < planes; ++p)
        sensitivity[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1] + np * p] +=
            raw[p * ni + i];
  for (size_t i = 0; i < ni; ++i)
    for (int p = 0; p < planes; ++p)
      weights[p * ni + i] =
          raw[p * ni + i] /
          sensitivity[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1] + np * p];
}

void oracle_hybrid_plane_weights_f64(const oracle_hybrid_geometry* g, const double* raw,
                                     int planes, int sens_mode, double* weights,
                                     double* sensitivity) {
  const size_t np = (size_t)g->n_columns * g->n_rows, ni = (size_t)g->n_pixel_infos;
  if (sens_mode != 2)
    for (size_t k = 0; k < np * (size_t)planes; ++k) sensitivity[k] = sens_mode ? 1.0 : 0.0;
  if (sens_mode == 0)
    for (size_t i = 0; i < ni; ++i)
      for (int p = 0; p < planes; ++p)
        sensitivity[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1] + np * p] +=
            raw[p * ni + i];
  for (size_t i = 0; i < ni; ++i)
    for (int p = 0; p < planes; ++p)
      weights[p * ni + i] =
          raw[p * ni + i] /
          sensitivity[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1] + np * p];
}
