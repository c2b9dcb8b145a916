// Strip-PET event geometry shared by host and device code.
//
// Everything that decides WHICH pixels an event touches is computed here
// with the exact float operation sequence the reference CPU build executes,
// so the GPU visits bit-for-bit the same pixels:
//   Response::calculate_tan_y_z        src/2d/strip/response.h:22-37
//   GaussianKernel::ellipse_bb         src/2d/strip/gaussian_kernel.h:108-128
//   GaussianKernel::in_ellipse         src/2d/strip/gaussian_kernel.h:131-137
//   Scanner::pixel_at / pixel_center   src/2d/strip/scanner.h:136-144
//   n_pixels_in_line + clamped box     src/2d/strip/reconstruction.h:216-246
// The reference is built by GCC with -O3 -march=native; in C++ mode GCC
// contracts a*b+c into FMA, so on x86 hosts with FMA the reference fuses a
// fixed set of sites.  Those sites are written as explicit fmaf() here and
// the file is compiled with -ffp-contract=off, so no other site is fused.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PS_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define PS_HD static inline
#endif

namespace petstrip {

// Scanner constants in the float representation Scanner<float, short> and
// GaussianKernel<float> hold (scanner.h:50-67, gaussian_kernel.h:27-33).
struct StripConst {
  float R, L, ph, pw;
  float tl_y, tl_z, tl_y_half_h, tl_z_half_w;
  float half_L, isz, isdl, c;  // 1/sz^2, 1/sdl^2, 1/(2 pi^2) sqrt(isz^2 isdl)
  float sigma_z, sigma_dl;
  int32_t W, H;  // n_z_pixels (columns), n_y_pixels (rows)
};

// Per-event geometry after the one-time setup pass.
struct EventGeom {
  float z, y, t, sec;   // ellipse centre (z, y), tan, sec = sqrt(tan^2 + 1)
  float A, B, C, pad;   // 3-sigma ellipse  A dy^2 + B dy dz + C dz^2 <= 9
  int32_t x0, x1, y0, y1;  // clamped exclusive box [x0,x1) x [y0,y1)
};

// x86 cvttss2si: truncation, INT32_MIN for NaN and out-of-range values.
PS_HD int32_t cvt_trunc_i32(float x) {
  if (!(x > -2147483904.0f && x < 2147483648.0f)) return INT32_MIN;
  return (int32_t)x;
}
PS_HD int32_t wrap_i16(int32_t v) {
  return (int32_t)(int16_t)(uint16_t)(uint32_t)v;
}
PS_HD int32_t clamp_i(int32_t v, int32_t lo, int32_t hi) {
  int32_t m = (lo < v) ? v : lo;
  return (m < hi) ? m : hi;
}

// Response (z_u, z_d, dl) -> event geometry.
PS_HD void event_geometry(const StripConst& s, float zu, float zd, float dl,
                          EventGeom& g) {
  // calculate_tan_y_z
  float t = (zu - zd) / (s.R + s.R);
  float y = (dl / sqrtf(fmaf(t, t, 1.0f))) * -0.5f;  // fused 1 + tan*tan
  float z = fmaf(y + y, t, zu + zd) * 0.5f;          // fused (zu+zd) + 2y tan
  // ellipse_bb
  float tsq = t * t;
  float sec = sqrtf(tsq + 1.0f);
  float A = fmaf((4.0f * sec) * sec, s.isdl, (tsq + tsq) * s.isz);  // fused
  float B = (t * -4.0f) * s.isz;
  float C = s.isz + s.isz;
  float hb = B * 0.5f;
  float B_2 = hb * hb;
  float bb_y = 3.0f / sqrtf(A - B_2 / C);
  float bb_z = 3.0f / sqrtf(C - B_2 / A);
  // pixel_at (truncating) and n_pixels_in_line (round half up)
  int32_t cx = wrap_i16(cvt_trunc_i32((z - s.tl_z) / s.pw));
  int32_t cy = wrap_i16(cvt_trunc_i32((s.tl_y - y) / s.ph));
  int32_t hw = cvt_trunc_i32(bb_z / s.pw + 0.5f);
  int32_t hh = cvt_trunc_i32(bb_y / s.ph + 0.5f);
  // Pixel(short) corners, both clamped to [0, n-1]; iteration is exclusive
  int32_t x0 = wrap_i16((int32_t)((uint32_t)cx - (uint32_t)hw));
  int32_t x1 = wrap_i16((int32_t)((uint32_t)cx + (uint32_t)hw));
  int32_t y0 = wrap_i16((int32_t)((uint32_t)cy - (uint32_t)hh));
  int32_t y1 = wrap_i16((int32_t)((uint32_t)cy + (uint32_t)hh));
  g.z = z;
  g.y = y;
  g.t = t;
  g.sec = sec;
  g.A = A;
  g.B = B;
  g.C = C;
  g.pad = 0.0f;
  g.x0 = clamp_i(x0, 0, s.W - 1);
  g.x1 = clamp_i(x1, 0, s.W - 1);
  g.y0 = clamp_i(y0, 0, s.H - 1);
  g.y1 = clamp_i(y1, 0, s.H - 1);
}

// Row part of pixel_center and of the ellipse quadratic.
PS_HD float row_dy(const StripConst& s, const EventGeom& g, int32_t iy) {
  return fmaf(-(float)iy, s.ph, s.tl_y_half_h) - g.y;  // fused pixel_center
}
PS_HD float col_dz(const StripConst& s, const EventGeom& g, int32_t iz) {
  return fmaf((float)iz, s.pw, s.tl_z_half_w) - g.z;  // fused pixel_center
}
// in_ellipse with the reference build's two fusions.
PS_HD bool in_ellipse(const EventGeom& g, float dy, float dz) {
  float Ady2 = (dy * g.A) * dy;
  float Bdy = dy * g.B;
  float q = fmaf(g.C * dz, dz, fmaf(Bdy, dz, Ady2));
  return q <= 9.0f;
}

}  // namespace petstrip
