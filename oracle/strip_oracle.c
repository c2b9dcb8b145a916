/*
 * CPU oracle for the 2D strip-PET list-mode MLEM path.
 * TEST INFRASTRUCTURE ONLY: see strip_oracle.h for what it restates, how it
 * is pinned and who may load it.
 *
 * Compiled with -ffp-contract=off: every fused multiply-add below is an
 * explicit fmaf()/fma() placed where the reference build fuses ("fused").
 */
#include "strip_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#ifndef M_1_PI
#define M_1_PI 0.31830988618379067154
#endif

/* ------------------------------------------------------------------------ */
/* Float->int conversions with x86 cvttss2si/cvttsd2si semantics (what the
 * reference's (int) and float->short conversions compile to): truncation
 * toward zero, out-of-range and NaN give INT32_MIN; a float->short goes
 * through that int and keeps the low 16 bits. */
static inline int32_t cvt_i32_f(float x) {
  if (!(x > -2147483904.0f && x < 2147483648.0f)) return INT32_MIN;
  return (int32_t)x;
}
static inline int32_t cvt_i32_d(double x) {
  if (!(x > -2147483649.0 && x < 2147483648.0)) return INT32_MIN;
  return (int32_t)x;
}
static inline int16_t wrap_i16(int32_t v) { return (int16_t)(uint16_t)(uint32_t)v; }
static inline int16_t clamp_i16(int16_t v, int16_t lo, int16_t hi) {
  /* Pixel::clamp: min(br, max(tl, x)) (pixel.h:51-54) */
  int16_t m = (lo < v) ? v : lo;
  return (m < hi) ? m : hi;
}

/* ------------------------------------------------------------------------ */
/* FP32 scanner + kernel constants (scanner.h:50-67, gaussian_kernel.h:27-33)*/
typedef struct {
  float R, L, ph, pw, sz, sdl, cy, cz;
  int16_t n_y, n_z;
  float size_y, size_z, tl_y, tl_z, tl_y_half_h, tl_z_half_w, half_L;
  float isz, isdl, c; /* inv_pow_sigma_z, inv_pow_sigma_dl, normalisation */
} s32;

static void make_s32(const oracle_scanner* s, s32* o) {
  o->R = (float)s->radius;
  o->L = (float)s->scintillator_length;
  o->n_y = (int16_t)s->n_y_pixels;
  o->n_z = (int16_t)s->n_z_pixels;
  o->ph = (float)s->pixel_height;
  o->pw = (float)s->pixel_width;
  o->sz = (float)s->sigma_z;
  o->sdl = (float)s->sigma_dl;
  o->cy = (float)s->center_y;
  o->cz = (float)s->center_z;
  o->size_y = (float)o->n_y * o->ph;
  o->size_z = (float)o->n_z * o->pw;
  o->tl_y = o->cy + o->size_y / 2;
  o->tl_z = o->cz - o->size_z / 2;
  o->tl_y_half_h = o->tl_y - o->ph / 2;
  o->tl_z_half_w = o->tl_z + o->pw / 2;
  o->half_L = o->L / 2;
  o->isz = 1 / (o->sz * o->sz);
  o->isdl = 1 / (o->sdl * o->sdl);
  /* double constant times a float sqrt, stored as float (gaussian_kernel.h:30-33) */
  o->c = (float)(1 / (2 * M_PI * M_PI) *
                 (double)sqrtf(o->isz * o->isz * o->isdl));
}

typedef struct {
  double R, L, ph, pw, sz, sdl, cy, cz;
  int16_t n_y, n_z;
  double tl_y, tl_z, tl_y_half_h, tl_z_half_w, half_L;
  double isz, isdl, c;
} s64;

static void make_s64(const oracle_scanner* s, s64* o, double sz, double sdl) {
  o->R = s->radius;
  o->L = s->scintillator_length;
  o->n_y = (int16_t)s->n_y_pixels;
  o->n_z = (int16_t)s->n_z_pixels;
  o->ph = s->pixel_height;
  o->pw = s->pixel_width;
  o->sz = sz;
  o->sdl = sdl;
  o->cy = s->center_y;
  o->cz = s->center_z;
  o->tl_y = o->cy + (o->n_y * o->ph) / 2;
  o->tl_z = o->cz - (o->n_z * o->pw) / 2;
  o->tl_y_half_h = o->tl_y - o->ph / 2;
  o->tl_z_half_w = o->tl_z + o->pw / 2;
  o->half_L = o->L / 2;
  o->isz = 1 / (sz * sz);
  o->isdl = 1 / (sdl * sdl);
  o->c = 1 / (2 * M_PI * M_PI) * sqrt(o->isz * o->isz * o->isdl);
}

/* ------------------------------------------------------------------------ */
/* Response::calculate_tan_y_z (response.h:22-37).                          */
static inline void tan_y_z_f(float R, float zu, float zd, float dl, float* t,
                             float* y, float* z) {
  *t = (zu - zd) / (R + R);
  *y = (dl / sqrtf(fmaf(*t, *t, 1.0f))) * -0.5f; /* fused: 1 + tan*tan */
  *z = fmaf(*y + *y, *t, zu + zd) * 0.5f;        /* fused: (zu+zd) + 2y*tan */
}
static inline void tan_y_z_d(double R, double zu, double zd, double dl,
                             double* t, double* y, double* z) {
  *t = (zu - zd) / (R + R);
  *y = (dl / sqrt(fma(*t, *t, 1.0))) * -0.5;
  *z = fma(*y + *y, *t, zu + zd) * 0.5;
}

/* GaussianKernel::ellipse_bb (gaussian_kernel.h:108-128, 159-160). */
static inline void ellipse_bb_f(float isz, float isdl, float t, float* sec,
                                float* A, float* B, float* C, float* bb_y,
                                float* bb_z) {
  float tsq = t * t;
  *sec = sqrtf(tsq + 1.0f);
  *A = fmaf((4.0f * *sec) * *sec, isdl, (tsq + tsq) * isz); /* fused */
  *B = (t * -4.0f) * isz;
  *C = isz + isz;
  float hb = *B * 0.5f;
  float B_2 = hb * hb;
  *bb_y = 3.0f / sqrtf(*A - B_2 / *C);
  *bb_z = 3.0f / sqrtf(*C - B_2 / *A);
}
static inline void ellipse_bb_d(double isz, double isdl, double t, double* sec,
                                double* A, double* B, double* C, double* bb_y,
                                double* bb_z) {
  double tsq = t * t;
  *sec = sqrt(tsq + 1.0);
  *A = fma((4.0 * *sec) * *sec, isdl, (tsq + tsq) * isz);
  *B = (t * -4.0) * isz;
  *C = isz + isz;
  double hb = *B * 0.5;
  double B_2 = hb * hb;
  *bb_y = 3.0 / sqrt(*A - B_2 / *C);
  *bb_z = 3.0 / sqrt(*C - B_2 / *A);
}

/* GaussianKernel::operator()(y, tan, sec, R, distance) (gaussian_kernel.h:51-79)
 * with the reference build's fusions. */
static inline float kernel_f(const s32* k, float y, float t, float sec,
                             float R, float dz, float dy) {
  float sec_sq = sec * sec;
  float yp = dy + y;
  float ap = -(yp - R) * sec_sq;
  float aq = -(R + yp) * sec_sq;
  float ar = ((yp * -2.0f) * sec) * t;
  float pq = fmaf(-dy, t, dz); /* fused: dz - dy*tan */
  float br = (dy * -2.0f) * sec;
  float norm = fmaf(ar * k->isdl, ar, fmaf(ap * k->isz, ap, (aq * k->isz) * aq));
  float bcb = fmaf((pq * k->isz) * pq, 2.0f, (br * k->isdl) * br);
  return (k->c / sqrtf(norm)) * expf(bcb * -0.5f);
}
static inline double kernel_d(double isz, double isdl, double c, double y,
                              double t, double sec, double R, double dz,
                              double dy) {
  double sec_sq = sec * sec;
  double yp = dy + y;
  double ap = -(yp - R) * sec_sq;
  double aq = -(R + yp) * sec_sq;
  double ar = ((yp * -2.0) * sec) * t;
  double pq = fma(-dy, t, dz);
  double br = (dy * -2.0) * sec;
  double norm = fma(ar * isdl, ar, fma(ap * isz, ap, (aq * isz) * aq));
  double bcb = fma((pq * isz) * pq, 2.0, (br * isdl) * br);
  return (c / sqrt(norm)) * exp(bcb * -0.5);
}

/* AnalyticKernel::operator()(y, tan, sec, R, distance)
 * (analytic_kernel.h:28-62): a, o, b as written there, then
 *   norm = a.S.a + 2 o.S.b,  K = c / sqrt(norm) exp(-(b.S.b - (b.S.a)^2/norm)/2)
 * with u.S.v = (u_p v_p + u_q v_q) isz + u_r v_r isdl in the reference's
 * order (multiply_inv_cor_mat, gaussian_kernel.h:152-156). */
static inline double analytic_d(double isz, double isdl, double c, double y,
                                double t, double sec, double R, double dz,
                                double dy) {
  double sec_sq = sec * sec;
  double yp = dy + y;
  double ap = -(yp - R) * sec_sq, aq = -(yp + R) * sec_sq;
  double ar = -2 * yp * sec * t;
  double op = ap * t, oq = aq * t, orr = -yp * sec * (1 + 2 * t * t);
  double pq = dz - dy * t;
  double br = -2 * dy * sec;
  double aia = ap * isz * ap + aq * isz * aq + ar * isdl * ar;
  double bia = pq * isz * ap + pq * isz * aq + br * isdl * ar;
  double bib = pq * isz * pq + pq * isz * pq + br * isdl * br;
  double oib = op * isz * pq + oq * isz * pq + orr * isdl * br;
  double norm = aia + 2 * oib;
  return c / sqrt(norm) * exp(-0.5 * (bib - bia * bia / norm));
}
static inline float analytic_f(const s32* k, float y, float t, float sec,
                               float R, float dz, float dy) {
  float sec_sq = sec * sec;
  float yp = dy + y;
  float ap = -(yp - R) * sec_sq, aq = -(yp + R) * sec_sq;
  float ar = -2 * yp * sec * t;
  float op = ap * t, oq = aq * t, orr = -yp * sec * (1 + 2 * t * t);
  float pq = dz - dy * t;
  float br = -2 * dy * sec;
  float aia = ap * k->isz * ap + aq * k->isz * aq + ar * k->isdl * ar;
  float bia = pq * k->isz * ap + pq * k->isz * aq + br * k->isdl * ar;
  float bib = pq * k->isz * pq + pq * k->isz * pq + br * k->isdl * br;
  float oib = op * k->isz * pq + oq * k->isz * pq + orr * k->isdl * br;
  float norm = aia + 2 * oib;
  return k->c / sqrtf(norm) * expf(-0.5f * (bib - bia * bia / norm));
}

/* ------------------------------------------------------------------------ */
/* Event geometry: centre, ellipse and the clamped exclusive box
 * (reconstruction.h:222-246). */
typedef struct {
  float t, y, z, sec, A, B, C;
  int x0, x1, y0, y1; /* box [x0,x1) x [y0,y1) */
} ev32;

static inline void event_f(const s32* s, const float* e, ev32* o) {
  float bb_y, bb_z;
  tan_y_z_f(s->R, e[0], e[1], e[2], &o->t, &o->y, &o->z);
  ellipse_bb_f(s->isz, s->isdl, o->t, &o->sec, &o->A, &o->B, &o->C, &bb_y,
               &bb_z);
  /* Scanner::pixel_at (scanner.h:142-144): truncating conversion */
  int16_t cx = wrap_i16(cvt_i32_f((o->z - s->tl_z) / s->pw));
  int16_t cyp = wrap_i16(cvt_i32_f((s->tl_y - o->y) / s->ph));
  /* n_pixels_in_line (reconstruction.h:216-218) */
  int32_t hw = cvt_i32_f(bb_z / s->pw + 0.5f);
  int32_t hh = cvt_i32_f(bb_y / s->ph + 0.5f);
  int16_t x0 = wrap_i16((int32_t)((uint32_t)cx - (uint32_t)hw));
  int16_t y0 = wrap_i16((int32_t)((uint32_t)cyp - (uint32_t)hh));
  int16_t x1 = wrap_i16((int32_t)((uint32_t)cx + (uint32_t)hw));
  int16_t y1 = wrap_i16((int32_t)((uint32_t)cyp + (uint32_t)hh));
  int16_t zx = 0, zy = 0, mx = (int16_t)(s->n_z - 1), my = (int16_t)(s->n_y - 1);
  o->x0 = clamp_i16(x0, zx, mx);
  o->y0 = clamp_i16(y0, zy, my);
  o->x1 = clamp_i16(x1, zx, mx);
  o->y1 = clamp_i16(y1, zy, my);
}

typedef struct {
  double t, y, z, sec, A, B, C;
  int x0, x1, y0, y1;
} ev64;

static inline void event_d(const s64* s, const float* e, ev64* o) {
  double bb_y, bb_z;
  tan_y_z_d(s->R, e[0], e[1], e[2], &o->t, &o->y, &o->z);
  ellipse_bb_d(s->isz, s->isdl, o->t, &o->sec, &o->A, &o->B, &o->C, &bb_y,
               &bb_z);
  int16_t cx = wrap_i16(cvt_i32_d((o->z - s->tl_z) / s->pw));
  int16_t cyp = wrap_i16(cvt_i32_d((s->tl_y - o->y) / s->ph));
  int32_t hw = cvt_i32_d(bb_z / s->pw + 0.5);
  int32_t hh = cvt_i32_d(bb_y / s->ph + 0.5);
  int16_t x0 = wrap_i16((int32_t)((uint32_t)cx - (uint32_t)hw));
  int16_t y0 = wrap_i16((int32_t)((uint32_t)cyp - (uint32_t)hh));
  int16_t x1 = wrap_i16((int32_t)((uint32_t)cx + (uint32_t)hw));
  int16_t y1 = wrap_i16((int32_t)((uint32_t)cyp + (uint32_t)hh));
  int16_t mx = (int16_t)(s->n_z - 1), my = (int16_t)(s->n_y - 1);
  o->x0 = clamp_i16(x0, 0, mx);
  o->y0 = clamp_i16(y0, 0, my);
  o->x1 = clamp_i16(x1, 0, mx);
  o->y1 = clamp_i16(y1, 0, my);
}

/* ------------------------------------------------------------------------ */
/* bb_update (reconstruction.h:220-298), FP32.  Scratch holds up to box-size
 * (index, K*rho) pairs.  Returns box and ellipse pixel counts. */
typedef struct {
  int32_t* idx;
  float* kr;
  size_t cap;
} scratch32;

static void ensure_cap32(scratch32* sc, size_t need) {
  if (need <= sc->cap) return;
  free(sc->idx);
  free(sc->kr);
  sc->cap = need;
  sc->idx = (int32_t*)malloc(need * sizeof(int32_t));
  sc->kr = (float*)malloc(need * sizeof(float));
}

static void bb_update_f32(const s32* s, const float* e, const float* rho,
                          float* out, scratch32* sc, uint64_t* nbox,
                          uint64_t* nell, float* denom_out) {
  ev32 v;
  event_f(s, e, &v);
  size_t box = (v.x1 > v.x0 && v.y1 > v.y0)
                   ? (size_t)(v.x1 - v.x0) * (size_t)(v.y1 - v.y0)
                   : 0;
  ensure_cap32(sc, box + 1);
  size_t n_ell = 0;
  float denom = 0;
  for (int iy = v.y0; iy < v.y1; ++iy) {
    float py = fmaf(-(float)iy, s->ph, s->tl_y_half_h); /* fused pixel_center */
    float dy = py - v.y;
    float Ady2 = (dy * v.A) * dy;
    float Bdy = dy * v.B;
    for (int iz = v.x0; iz < v.x1; ++iz) {
      float px = fmaf((float)iz, s->pw, s->tl_z_half_w); /* fused */
      float dz = px - v.z;
      /* in_ellipse (gaussian_kernel.h:131-137), fused twice */
      float q = fmaf(v.C * dz, dz, fmaf(Bdy, dz, Ady2));
      if (!(q <= 9.0f)) continue;
      int32_t p = iy * s->n_z + iz;
      float K = kernel_f(s, v.y, v.t, v.sec, s->R, dz, dy);
      float kr = K * rho[p];
      denom = kr + denom;
      sc->idx[n_ell] = p;
      sc->kr[n_ell] = kr;
      ++n_ell;
    }
  }
  if (denom_out) *denom_out = denom;
  if (nbox) *nbox += box;
  if (nell) *nell += n_ell;
  if (!out) return;
  float inv = (denom > 0) ? 1 / denom : 0; /* reconstruction.h:290 */
  for (size_t i = 0; i < n_ell; ++i) {
    float* o = &out[sc->idx[i]];
    *o = fmaf(sc->kr[i], inv, *o); /* fused: out += kr*inv */
  }
}

int oracle_mlem_f32(const oracle_scanner* sc_in, const float* events,
                    int64_t n, float* rho, int n_iter, int n_threads,
                    float* rho_iters, uint64_t* counts) {
  s32 s;
  make_s32(sc_in, &s);
  if (n_threads < 1) n_threads = 1;
  const size_t np = (size_t)s.n_y * (size_t)s.n_z;
  float* priv = (float*)malloc((size_t)n_threads * np * sizeof(float));
  if (!priv) return -1;
  uint64_t tot_box = 0, tot_ell = 0;
  for (int it = 0; it < n_iter; ++it) {
    memset(priv, 0, (size_t)n_threads * np * sizeof(float));
    uint64_t it_box = 0, it_ell = 0;
#pragma omp parallel num_threads(n_threads) reduction(+ : it_box, it_ell)
    {
      scratch32 sc = {0, 0, 0};
      float* mine = priv + (size_t)omp_get_thread_num() * np;
#pragma omp for schedule(dynamic)
      for (int64_t e = 0; e < n; ++e)
        bb_update_f32(&s, events + 3 * e, rho, mine, &sc, &it_box, &it_ell, 0);
      free(sc.idx);
      free(sc.kr);
    }
    /* rho.assign(0); rho += thread images in thread order (:143-149) */
    for (size_t i = 0; i < np; ++i) rho[i] = 0;
    for (int t = 0; t < n_threads; ++t)
      for (size_t i = 0; i < np; ++i) rho[i] += priv[(size_t)t * np + i];
    if (rho_iters) memcpy(rho_iters + (size_t)it * np, rho, np * sizeof(float));
    tot_box = it_box;
    tot_ell = it_ell;
  }
  if (counts) {
    counts[0] = tot_box;
    counts[1] = tot_ell;
  }
  free(priv);
  return 0;
}

int oracle_denominators_f32(const oracle_scanner* sc_in, const float* events,
                            int64_t n, const float* rho, float* denom,
                            int n_threads) {
  s32 s;
  make_s32(sc_in, &s);
  if (n_threads < 1) n_threads = 1;
#pragma omp parallel num_threads(n_threads)
  {
    scratch32 sc = {0, 0, 0};
#pragma omp for schedule(dynamic, 64)
    for (int64_t e = 0; e < n; ++e)
      bb_update_f32(&s, events + 3 * e, rho, 0, &sc, 0, 0, &denom[e]);
    free(sc.idx);
    free(sc.kr);
  }
  return 0;
}

int oracle_footprint(const oracle_scanner* sc_in, const float* events,
                     int64_t n, int32_t* n_box, int32_t* n_ell, int n_threads) {
  s32 s;
  make_s32(sc_in, &s);
  if (n_threads < 1) n_threads = 1;
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 256)
  for (int64_t e = 0; e < n; ++e) {
    ev32 v;
    event_f(&s, events + 3 * e, &v);
    int32_t box = (v.x1 > v.x0 && v.y1 > v.y0) ? (v.x1 - v.x0) * (v.y1 - v.y0) : 0;
    int32_t ell = 0;
    for (int iy = v.y0; iy < v.y1; ++iy) {
      float py = fmaf(-(float)iy, s.ph, s.tl_y_half_h);
      float dy = py - v.y;
      float Ady2 = (dy * v.A) * dy;
      float Bdy = dy * v.B;
      for (int iz = v.x0; iz < v.x1; ++iz) {
        float px = fmaf((float)iz, s.pw, s.tl_z_half_w);
        float dz = px - v.z;
        float q = fmaf(v.C * dz, dz, fmaf(Bdy, dz, Ady2));
        ell += (q <= 9.0f);
      }
    }
    n_box[e] = box;
    n_ell[e] = ell;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 twin of bb_update / operator(). */
typedef struct {
  int32_t* idx;
  double* kr;
  size_t cap;
} scratch64;

static void ensure_cap64(scratch64* sc, size_t need) {
  if (need <= sc->cap) return;
  free(sc->idx);
  free(sc->kr);
  sc->cap = need;
  sc->idx = (int32_t*)malloc(need * sizeof(int32_t));
  sc->kr = (double*)malloc(need * sizeof(double));
}

static void bb_update_f64(const s64* s, const float* e, const double* rho,
                          double* out, scratch64* sc) {
  ev64 v;
  event_d(s, e, &v);
  size_t box = (v.x1 > v.x0 && v.y1 > v.y0)
                   ? (size_t)(v.x1 - v.x0) * (size_t)(v.y1 - v.y0)
                   : 0;
  ensure_cap64(sc, box + 1);
  size_t n_ell = 0;
  double denom = 0;
  for (int iy = v.y0; iy < v.y1; ++iy) {
    double py = fma(-(double)iy, s->ph, s->tl_y_half_h);
    double dy = py - v.y;
    double Ady2 = (dy * v.A) * dy;
    double Bdy = dy * v.B;
    for (int iz = v.x0; iz < v.x1; ++iz) {
      double px = fma((double)iz, s->pw, s->tl_z_half_w);
      double dz = px - v.z;
      double q = fma(v.C * dz, dz, fma(Bdy, dz, Ady2));
      if (!(q <= 9.0)) continue;
      int32_t p = iy * s->n_z + iz;
      double K = kernel_d(s->isz, s->isdl, s->c, v.y, v.t, v.sec, s->R, dz, dy);
      double kr = K * rho[p];
      denom += kr;
      sc->idx[n_ell] = p;
      sc->kr[n_ell] = kr;
      ++n_ell;
    }
  }
  double inv = (denom > 0) ? 1 / denom : 0;
  for (size_t i = 0; i < n_ell; ++i) out[sc->idx[i]] += sc->kr[i] * inv;
}

int oracle_mlem_f64(const oracle_scanner* sc_in, const float* events,
                    int64_t n, double* rho, int n_iter, int n_threads,
                    double* rho_iters) {
  s64 s;
  make_s64(sc_in, &s, sc_in->sigma_z, sc_in->sigma_dl);
  if (n_threads < 1) n_threads = 1;
  const size_t np = (size_t)s.n_y * (size_t)s.n_z;
  double* priv = (double*)malloc((size_t)n_threads * np * sizeof(double));
  if (!priv) return -1;
  for (int it = 0; it < n_iter; ++it) {
    memset(priv, 0, (size_t)n_threads * np * sizeof(double));
#pragma omp parallel num_threads(n_threads)
    {
      scratch64 sc = {0, 0, 0};
      double* mine = priv + (size_t)omp_get_thread_num() * np;
#pragma omp for schedule(dynamic)
      for (int64_t e = 0; e < n; ++e)
        bb_update_f64(&s, events + 3 * e, rho, mine, &sc);
      free(sc.idx);
      free(sc.kr);
    }
    for (size_t i = 0; i < np; ++i) rho[i] = 0;
    for (int t = 0; t < n_threads; ++t)
      for (size_t i = 0; i < np; ++i) rho[i] += priv[(size_t)t * np + i];
    if (rho_iters)
      memcpy(rho_iters + (size_t)it * np, rho, np * sizeof(double));
  }
  free(priv);
  return 0;
}

/* The reference --gpu iteration (cuda/reconstruction.cu:59-113,
 * cuda/reconstruction/warp_granularity.cuh:42-153) in FP64: per event the
 * denominator D = sum_p K rho_p s_p over the same box and 3-sigma ellipse,
 * then rho'_p += K rho_p / D, with the Gaussian (kernel 0) or the analytic
 * kernel (kernel 1) and the sensitivity s (NULL: 1).  D = 0 contributes
 * nothing (the reference divides unguarded). */
static void bb_update_variant_f64(const s64* s, const float* e, const double* rho,
                                  int kernel, const double* sens, double* out,
                                  scratch64* sc) {
  ev64 v;
  event_d(s, e, &v);
  size_t box = (v.x1 > v.x0 && v.y1 > v.y0)
                   ? (size_t)(v.x1 - v.x0) * (size_t)(v.y1 - v.y0)
                   : 0;
  ensure_cap64(sc, box + 1);
  size_t n_ell = 0;
  double denom = 0;
  for (int iy = v.y0; iy < v.y1; ++iy) {
    double py = fma(-(double)iy, s->ph, s->tl_y_half_h);
    double dy = py - v.y;
    double Ady2 = (dy * v.A) * dy;
    double Bdy = dy * v.B;
    for (int iz = v.x0; iz < v.x1; ++iz) {
      double px = fma((double)iz, s->pw, s->tl_z_half_w);
      double dz = px - v.z;
      double q = fma(v.C * dz, dz, fma(Bdy, dz, Ady2));
      if (!(q <= 9.0)) continue;
      int32_t p = iy * s->n_z + iz;
      double K = kernel ? analytic_d(s->isz, s->isdl, s->c, v.y, v.t, v.sec, s->R, dz, dy)
                        : kernel_d(s->isz, s->isdl, s->c, v.y, v.t, v.sec, s->R, dz, dy);
      double kr = K * rho[p];
      denom += kr * (sens ? sens[p] : 1.0);
      sc->idx[n_ell] = p;
      sc->kr[n_ell] = kr;
      ++n_ell;
    }
  }
  double inv = (denom > 0) ? 1 / denom : 0;
  for (size_t i = 0; i < n_ell; ++i) out[sc->idx[i]] += sc->kr[i] * inv;
}

int oracle_mlem_variant_f64(const oracle_scanner* sc_in, const float* events,
                            int64_t n, double* rho, int n_iter, int kernel,
                            const double* sens, int n_threads, double* rho_iters) {
  s64 s;
  make_s64(sc_in, &s, sc_in->sigma_z, sc_in->sigma_dl);
  if (n_threads < 1) n_threads = 1;
  const size_t np = (size_t)s.n_y * (size_t)s.n_z;
  double* priv = (double*)malloc((size_t)n_threads * np * sizeof(double));
  if (!priv) return -1;
  for (int it = 0; it < n_iter; ++it) {
    memset(priv, 0, (size_t)n_threads * np * sizeof(double));
#pragma omp parallel num_threads(n_threads)
    {
      scratch64 sc = {0, 0, 0};
      double* mine = priv + (size_t)omp_get_thread_num() * np;
#pragma omp for schedule(dynamic)
      for (int64_t e = 0; e < n; ++e)
        bb_update_variant_f64(&s, events + 3 * e, rho, kernel, sens, mine, &sc);
      free(sc.idx);
      free(sc.kr);
    }
    for (size_t i = 0; i < np; ++i) rho[i] = 0;
    for (int t = 0; t < n_threads; ++t)
      for (size_t i = 0; i < np; ++i) rho[i] += priv[(size_t)t * np + i];
    if (rho_iters)
      memcpy(rho_iters + (size_t)it * np, rho, np * sizeof(double));
  }
  free(priv);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Point helpers and kernel probes for the unit-test goldens. */
float oracle_analytic_kernel_f32(double sigma_z, double sigma_dl, float y, float tan,
                                 float sec, float R, float dz, float dy) {
  oracle_scanner sc = {1, 1, 1, 1, 1, 1, sigma_z, sigma_dl, 0, 0};
  s32 s;
  make_s32(&sc, &s);
  return analytic_f(&s, y, tan, sec, R, dz, dy);
}

double oracle_analytic_kernel_f64(double sigma_z, double sigma_dl, double y,
                                  double tan, double sec, double R, double dz,
                                  double dy) {
  double isz = 1 / (sigma_z * sigma_z), isdl = 1 / (sigma_dl * sigma_dl);
  double c = 1 / (2 * M_PI * M_PI) * sqrt(isz * isz * isdl);
  return analytic_d(isz, isdl, c, y, tan, sec, R, dz, dy);
}

float oracle_kernel_f32(double sigma_z, double sigma_dl, float y, float tan,
                        float sec, float R, float dz, float dy) {
  oracle_scanner sc = {1, 1, 1, 1, 1, 1, sigma_z, sigma_dl, 0, 0};
  s32 s;
  make_s32(&sc, &s);
  return kernel_f(&s, y, tan, sec, R, dz, dy);
}

double oracle_kernel_f64(double sigma_z, double sigma_dl, double y, double tan,
                         double sec, double R, double dz, double dy) {
  double isz = 1 / (sigma_z * sigma_z), isdl = 1 / (sigma_dl * sigma_dl);
  double c = 1 / (2 * M_PI * M_PI) * sqrt(isz * isz * isdl);
  return kernel_d(isz, isdl, c, y, tan, sec, R, dz, dy);
}

void oracle_ellipse_bb_f32(double sigma_z, double sigma_dl, float tan,
                           float out[6]) {
  float sz = (float)sigma_z, sdl = (float)sigma_dl;
  float isz = 1 / (sz * sz), isdl = 1 / (sdl * sdl);
  ellipse_bb_f(isz, isdl, tan, &out[0], &out[1], &out[2], &out[3], &out[4],
               &out[5]);
}

void oracle_ellipse_bb_f64(double sigma_z, double sigma_dl, double tan,
                           double out[6]) {
  double isz = 1 / (sigma_z * sigma_z), isdl = 1 / (sigma_dl * sigma_dl);
  ellipse_bb_d(isz, isdl, tan, &out[0], &out[1], &out[2], &out[3], &out[4],
               &out[5]);
}

void oracle_tan_y_z_f32(double radius, const float ev[3], float out[3]) {
  tan_y_z_f((float)radius, ev[0], ev[1], ev[2], &out[0], &out[1], &out[2]);
}

void oracle_tan_y_z_f64(double radius, const double ev[3], double out[3]) {
  tan_y_z_d(radius, ev[0], ev[1], ev[2], &out[0], &out[1], &out[2]);
}

void oracle_pixel_at_f32(const oracle_scanner* sc, float z, float y,
                         int32_t out[2]) {
  s32 s;
  make_s32(sc, &s);
  out[0] = wrap_i16(cvt_i32_f((z - s.tl_z) / s.pw));
  out[1] = wrap_i16(cvt_i32_f((s.tl_y - y) / s.ph));
}

void oracle_pixel_center_f32(const oracle_scanner* sc, int32_t x, int32_t y,
                             float out[2]) {
  s32 s;
  make_s32(sc, &s);
  out[0] = fmaf((float)x, s.pw, s.tl_z_half_w);
  out[1] = fmaf(-(float)y, s.ph, s.tl_y_half_h);
}

void oracle_pixel_at_f64(const oracle_scanner* sc, double z, double y,
                         int32_t out[2]) {
  s64 s;
  make_s64(sc, &s, sc->sigma_z, sc->sigma_dl);
  out[0] = wrap_i16(cvt_i32_d((z - s.tl_z) / s.pw));
  out[1] = wrap_i16(cvt_i32_d((s.tl_y - y) / s.ph));
}

void oracle_pixel_center_f64(const oracle_scanner* sc, int32_t x, int32_t y,
                             double out[2]) {
  s64 s;
  make_s64(sc, &s, sc->sigma_z, sc->sigma_dl);
  out[0] = fma((double)x, s.pw, s.tl_z_half_w);
  out[1] = fma(-(double)y, s.ph, s.tl_y_half_h);
}

/* Scanner::sensitivity (scanner.h:147-171). */
static float sens_point_f(const s32* s, float z, float y) {
  float L_plus = s->half_L + z;
  float L_minus = s->half_L - z;
  float R_plus = s->R + y;
  float R_minus = s->R - y;
  if (R_plus <= 0) R_plus = 1;
  if (R_minus <= 0) R_minus = 1;
  float a = L_minus / R_minus, b = L_plus / R_plus;
  float mn = (b < a) ? b : a; /* std::min */
  float c = -L_plus / R_minus, d = -L_minus / R_plus;
  float mx = (c < d) ? d : c; /* std::max */
  return (float)M_1_PI * (atanf(mn) - atanf(mx));
}

double oracle_sensitivity_point_f64(double R, double L, double z, double y) {
  double hl = L / 2;
  double L_plus = hl + z, L_minus = hl - z, R_plus = R + y, R_minus = R - y;
  if (R_plus <= 0) R_plus = 1;
  if (R_minus <= 0) R_minus = 1;
  double a = L_minus / R_minus, b = L_plus / R_plus;
  double mn = (b < a) ? b : a;
  double c = -L_plus / R_minus, d = -L_minus / R_plus;
  double mx = (c < d) ? d : c;
  return M_1_PI * (atan(mn) - atan(mx));
}

int oracle_sensitivity_f32(const oracle_scanner* sc, int mode, float* out) {
  s32 s;
  make_s32(sc, &s);
  for (int y = 0; y < s.n_y; ++y)
    for (int x = 0; x < s.n_z; ++x) {
      float pz = fmaf((float)x, s.pw, s.tl_z_half_w);
      float py = fmaf(-(float)y, s.ph, s.tl_y_half_h);
      float v;
      if (mode == 0) {
        v = sens_point_f(&s, pz, py);
      } else {
        float hw = s.pw / 2, hh = s.ph / 2;
        v = sens_point_f(&s, pz, py) / 3 + sens_point_f(&s, pz + hw, py + hh) / 6 +
            sens_point_f(&s, pz - hw, py - hh) / 6 +
            sens_point_f(&s, pz - hw, py + hh) / 6 +
            sens_point_f(&s, pz + hw, py - hh) / 6;
      }
      out[(size_t)y * s.n_z + x] = v;
    }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 log-likelihood and its gradient with frozen support.
 *   ln K = ln c - 1/2 ln(norm) - 1/2 bSb,  c = 1/(2 pi^2 sz^2 sdl)
 *   norm = isz (ap^2 + aq^2) + isdl ar^2,  bSb = 2 isz pq^2 + isdl br^2
 *   d lnK/d sz  = -2/sz + isz (ap^2+aq^2)/(sz norm) + 2 isz pq^2 / sz
 *   d lnK/d sdl = -1/sdl + isdl ar^2/(sdl norm) + isdl br^2 / sdl        */
int oracle_loglik_f64(const oracle_scanner* sc_in, const float* events,
                      int64_t n, const double* rho, const double* sens,
                      double szk, double sdlk, double szs, double sdls,
                      double* loglik, double* lambda, double* grad_rho,
                      double grad_sigma[2], int n_threads) {
  s64 sk, ss;
  make_s64(sc_in, &sk, szk, sdlk);
  make_s64(sc_in, &ss, szs, sdls);
  if (n_threads < 1) n_threads = 1;
  const size_t np = (size_t)sk.n_y * (size_t)sk.n_z;
  double* priv = 0;
  if (grad_rho) {
    priv = (double*)calloc((size_t)n_threads * np, sizeof(double));
    if (!priv) return -1;
  }
  double ll = 0, gsz = 0, gsdl = 0;
#pragma omp parallel num_threads(n_threads) reduction(+ : ll, gsz, gsdl)
  {
    scratch64 sc = {0, 0, 0};
    double* mine = priv ? priv + (size_t)omp_get_thread_num() * np : 0;
#pragma omp for schedule(dynamic, 64)
    for (int64_t e = 0; e < n; ++e) {
      ev64 v;
      event_d(&ss, events + 3 * e, &v); /* support at sigma_support */
      size_t box = (v.x1 > v.x0 && v.y1 > v.y0)
                       ? (size_t)(v.x1 - v.x0) * (size_t)(v.y1 - v.y0)
                       : 0;
      ensure_cap64(&sc, box + 1);
      size_t n_ell = 0;
      double lam = 0, dsz = 0, dsdl = 0;
      for (int iy = v.y0; iy < v.y1; ++iy) {
        double py = fma(-(double)iy, sk.ph, sk.tl_y_half_h);
        double dy = py - v.y;
        double Ady2 = (dy * v.A) * dy, Bdy = dy * v.B;
        for (int iz = v.x0; iz < v.x1; ++iz) {
          double px = fma((double)iz, sk.pw, sk.tl_z_half_w);
          double dz = px - v.z;
          double q = fma(v.C * dz, dz, fma(Bdy, dz, Ady2));
          if (!(q <= 9.0)) continue;
          int32_t p = iy * sk.n_z + iz;
          double sec_sq = v.sec * v.sec, yp = dy + v.y;
          double ap = -(yp - sk.R) * sec_sq, aq = -(sk.R + yp) * sec_sq;
          double ar = -2 * yp * v.sec * v.t;
          double pq = dz - dy * v.t, br = -2 * dy * v.sec;
          double apq = ap * ap + aq * aq;
          double norm = sk.isz * apq + sk.isdl * ar * ar;
          double bsb = 2 * sk.isz * pq * pq + sk.isdl * br * br;
          double K = sk.c / sqrt(norm) * exp(-0.5 * bsb);
          double kr = K * rho[p];
          lam += kr;
          dsz += kr * (-2 / szk + sk.isz * apq / (szk * norm) +
                       2 * sk.isz * pq * pq / szk);
          dsdl += kr * (-1 / sdlk + sk.isdl * ar * ar / (sdlk * norm) +
                        sk.isdl * br * br / sdlk);
          sc.idx[n_ell] = p;
          sc.kr[n_ell] = K;
          ++n_ell;
        }
      }
      if (lambda) lambda[e] = lam;
      if (lam > 0) {
        ll += log(lam);
        gsz += dsz / lam;
        gsdl += dsdl / lam;
        if (mine)
          for (size_t i = 0; i < n_ell; ++i) mine[sc.idx[i]] += sc.kr[i] / lam;
      }
    }
    free(sc.idx);
    free(sc.kr);
  }
  for (size_t i = 0; i < np; ++i) ll -= (sens ? sens[i] : 1.0) * rho[i];
  if (grad_rho) {
    for (size_t i = 0; i < np; ++i) {
      double g = -(sens ? sens[i] : 1.0);
      for (int t = 0; t < n_threads; ++t) g += priv[(size_t)t * np + i];
      grad_rho[i] = g;
    }
    free(priv);
  }
  if (loglik) *loglik = ll;
  if (grad_sigma) {
    grad_sigma[0] = gsz;
    grad_sigma[1] = gsdl;
  }
  return 0;
}
