// Test harness around the reference's own 2D strip MLEM (checker only).
//
// This file holds no reference code: it #includes the reference headers
// where they lie (the Makefile beside it passes -I <reference>/src) and
// exposes them behind a C ABI so the Python tests can drive them.
//   - Strip::Reconstruction::operator()   src/2d/strip/reconstruction.h:108-154
//   - Strip::GaussianKernel               src/2d/strip/gaussian_kernel.h:12-161
//   - Strip::Scanner                      src/2d/strip/scanner.h:26-291
//   - Strip::Response::calculate_tan_y_z  src/2d/strip/response.h:22-37
//   - Strip::AnalyticKernel               src/2d/strip/analytic_kernel.h:16-87
//   - util::text_parser                   src/util/text_parser.h:19-176
//   - Reconstruction::fast_load_txt_events src/2d/strip/reconstruction.h:169-196
// Only tests and the golden-vector script load the built library.
#include <cstdint>
#include <cstring>
#include <omp.h>

#include "2d/strip/reconstruction.h"
#include "2d/strip/gaussian_kernel.h"
#include "2d/strip/analytic_kernel.h"
#include "util/text_parser.h"

using F = float;
using S = short;
using Kernel = PET2D::Strip::GaussianKernel<F>;
using Recon = PET2D::Strip::Reconstruction<F, Kernel>;
using Scanner = PET2D::Strip::Scanner<F, S>;

namespace {
struct NoProgress {
  void operator()(int, bool = false) {}
};
}  // namespace

extern "C" {

// Runs n_iter MLEM iterations of the reference CPU path.  rho_inout is W*H
// (n_z columns, n_y rows, row 0 on top); rho_iters (optional) receives the
// image after every iteration.
int ref_strip_mlem(double radius, double length, int n_y, int n_z,
                   double pixel_size, double sigma_z, double sigma_dl,
                   const float* events, int64_t n_events, float* rho_inout,
                   int n_iter, int n_threads, float* rho_iters) {
  omp_set_num_threads(n_threads);
  Scanner scanner(radius, length, n_y, n_z, pixel_size, pixel_size, sigma_z,
                  sigma_dl);
  Recon recon(scanner);
  const size_t np = (size_t)n_y * (size_t)n_z;
  for (size_t i = 0; i < np; ++i) recon.rho[i] = rho_inout[i];
  recon.responses.reserve((size_t)n_events);
  for (int64_t e = 0; e < n_events; ++e)
    recon.responses.emplace_back(events[3 * e], events[3 * e + 1],
                                 events[3 * e + 2]);
  NoProgress progress;
  for (int it = 0; it < n_iter; ++it) {
    recon(progress, 1, it);
    if (rho_iters)
      std::memcpy(rho_iters + (size_t)it * np, recon.rho.data, np * sizeof(F));
  }
  std::memcpy(rho_inout, recon.rho.data, np * sizeof(F));
  return 0;
}

// The sensitivity map Reconstruction builds at construction
// (reconstruction.h:67-73, Scanner::sensitivity at the pixel centre).
int ref_strip_sensitivity(double radius, double length, int n_y, int n_z,
                          double pixel_size, double sigma_z, double sigma_dl,
                          float* out) {
  Scanner scanner(radius, length, n_y, n_z, pixel_size, pixel_size, sigma_z,
                  sigma_dl);
  Recon recon(scanner);
  std::memcpy(out, recon.sensitivity.data,
              (size_t)n_y * (size_t)n_z * sizeof(F));
  return 0;
}

// 5-point Scanner::pixel_sensitivity (scanner.h:174-185), as the reference
// --gpu path fills its sensitivity texture.
int ref_strip_pixel_sensitivity(double radius, double length, int n_y,
                                int n_z, double pixel_size, double sigma_z,
                                double sigma_dl, float* out) {
  Scanner scanner(radius, length, n_y, n_z, pixel_size, pixel_size, sigma_z,
                  sigma_dl);
  for (int y = 0; y < n_y; ++y)
    for (int x = 0; x < n_z; ++x)
      out[(size_t)y * n_z + x] =
          scanner.pixel_sensitivity(PET2D::Pixel<S>((S)x, (S)y));
  return 0;
}

// GaussianKernel::operator()(y, tan, sec, R, distance) in F=float or double.
double ref_strip_kernel_f32(double sigma_z, double sigma_dl, double y,
                            double tan, double sec, double R, double dz,
                            double dy) {
  Kernel k((F)sigma_z, (F)sigma_dl);
  return k((F)y, (F)tan, (F)sec, (F)R, PET2D::Vector<F>((F)dz, (F)dy));
}

double ref_strip_kernel_f64(double sigma_z, double sigma_dl, double y,
                            double tan, double sec, double R, double dz,
                            double dy) {
  PET2D::Strip::GaussianKernel<double> k(sigma_z, sigma_dl);
  return k(y, tan, sec, R, PET2D::Vector<double>(dz, dy));
}

// AnalyticKernel::operator()(y, tan, sec, R, distance) in float or double.
double ref_strip_analytic_kernel_f32(double sigma_z, double sigma_dl, double y,
                                     double tan, double sec, double R, double dz,
                                     double dy) {
  PET2D::Strip::AnalyticKernel<F> k((F)sigma_z, (F)sigma_dl);
  return k((F)y, (F)tan, (F)sec, (F)R, PET2D::Vector<F>((F)dz, (F)dy));
}

double ref_strip_analytic_kernel_f64(double sigma_z, double sigma_dl, double y,
                                     double tan, double sec, double R, double dz,
                                     double dy) {
  PET2D::Strip::AnalyticKernel<double> k(sigma_z, sigma_dl);
  return k(y, tan, sec, R, PET2D::Vector<double>(dz, dy));
}

// GaussianKernel::ellipse_bb (gaussian_kernel.h:108-128) in float.
void ref_strip_ellipse_bb(double sigma_z, double sigma_dl, double tan,
                          float out[6]) {
  Kernel k((F)sigma_z, (F)sigma_dl);
  F sec, A, B, C, bb_y, bb_z;
  k.ellipse_bb((F)tan, sec, A, B, C, bb_y, bb_z);
  out[0] = sec; out[1] = A; out[2] = B; out[3] = C; out[4] = bb_y;
  out[5] = bb_z;
}

// Response::calculate_tan_y_z (response.h:22-37) in float.
void ref_strip_tan_y_z(double radius, const float ev[3], float out[3]) {
  PET2D::Strip::Response<F> r(ev[0], ev[1], ev[2]);
  F t, y, z;
  r.calculate_tan_y_z((F)radius, t, y, z);
  out[0] = t; out[1] = y; out[2] = z;
}

// Scanner::pixel_at (scanner.h:142-144).
void ref_strip_pixel_at(double radius, double length, int n_y, int n_z,
                        double pixel_size, float z, float y, int out[2]) {
  Scanner scanner(radius, length, n_y, n_z, pixel_size, pixel_size, 0.01,
                  0.04);
  auto p = scanner.pixel_at(PET2D::Point<F>(z, y));
  out[0] = p.x; out[1] = p.y;
}

// Scanner::pixel_center (scanner.h:136-140) in float: {z, y}.
void ref_strip_pixel_center(double radius, double length, int n_y, int n_z,
                            double pixel_size, int x, int y, float out[2]) {
  Scanner scanner(radius, length, n_y, n_z, pixel_size, pixel_size, 0.01, 0.04);
  auto p = scanner.pixel_center(PET2D::Pixel<S>((S)x, (S)y));
  out[0] = p.x;
  out[1] = p.y;
}

// util::text_parser's readers on one line: n fields, each an int
// (kinds[k] = 0, operator>><int>) or a float (kinds[k] = 1,
// operator>><float>).  Returns -1 if the parser throws.
int ref_text_parse_fields(const char* text, int n, const int* kinds, float* out_f,
                          int* out_i) {
  util::text_parser p(text);
  try {
    for (int k = 0; k < n; ++k) {
      if (kinds[k] == 0) {
        int v;
        p >> v;
        out_i[k] = v;
      } else {
        float v;
        p >> v;
        out_f[k] = v;
      }
    }
  } catch (const char*) {
    return -1;
  }
  return 0;
}

// Strip::Reconstruction::fast_load_txt_events on a file: the responses it
// reads (up to cap triples into out).  Returns their number, or -1 if the
// reader throws.
int64_t ref_fast_load_txt_events(const char* fn, int is_3d, float* out, int64_t cap) {
  Scanner scanner(0.43, 0.5, 8, 8, 0.01, 0.01, 0.01, 0.04);
  Recon recon(scanner);
  try {
    recon.fast_load_txt_events(fn, is_3d != 0);
  } catch (...) {
    return -1;
  }
  const int64_t n = (int64_t)recon.responses.size();
  for (int64_t e = 0; e < n && e < cap; ++e) {
    out[3 * e] = recon.responses[e].z_u;
    out[3 * e + 1] = recon.responses[e].z_d;
    out[3 * e + 2] = recon.responses[e].dl;
  }
  return n;
}

}  // extern "C"
