/*
 * CPU oracle for the 2D strip-PET list-mode MLEM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the GPU path is
 * compared against and the timed CPU baseline; the product never links or
 * calls it.  Only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark load it.
 *
 * It restates the reference CPU algorithm (grey277/j-pet-mlem):
 *   Strip::Reconstruction::operator()   src/2d/strip/reconstruction.h:108-154
 *   Strip::Reconstruction::bb_update    src/2d/strip/reconstruction.h:216-298
 *   Strip::Response::calculate_tan_y_z  src/2d/strip/response.h:22-37
 *   Strip::GaussianKernel               src/2d/strip/gaussian_kernel.h:19-160
 *   Strip::Scanner                      src/2d/strip/scanner.h:39-67,136-185
 *   PET2D::Pixel::clamp/index           src/2d/geometry/pixel.h:29,51-54
 *   Strip::AnalyticKernel               src/2d/strip/analytic_kernel.h:28-62
 *   the --gpu iteration                 src/2d/strip/cuda/reconstruction.cu:59-113,
 *                                       cuda/reconstruction/warp_granularity.cuh:42-153
 *
 * Floating-point contraction: the reference is built by GCC with
 * -O3 -march=native (CMakeLists.txt:76-81).  In C++ mode GCC contracts
 * a*b+c into FMA even under -std=c++11, so on any FMA-capable x86 host the
 * reference computes a fixed set of fused operations.  The FP32 path below
 * spells out exactly those fusions with fmaf() and is compiled with
 * -ffp-contract=off so that nothing else is fused.  Each fused site is
 * marked "fused" in strip_oracle.c.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   - reference unit-test goldens (gaussian_kernel_test.cpp:15-51,
 *     kernel_test.cpp:31-49,81-113, response_test.cpp:11-55,
 *     scanner_test.cpp:15-66),
 *   - bitwise equality with the reference built from its own sources
 *     (oracle/_ref) for whole MLEM iterations with one thread,
 *   - invariants (sum of rho' = number of events with a positive
 *     denominator, last row/column zero, single-event closed form),
 *   - FP64 twin against the FP32 path.
 */
#ifndef STRIP_ORACLE_H
#define STRIP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order as the Scanner constructor (scanner.h:39-49). */
typedef struct oracle_scanner {
  double radius;              /* R: strips at y = +-R                      */
  double scintillator_length; /* L                                          */
  int32_t n_y_pixels;         /* image rows (H)                             */
  int32_t n_z_pixels;         /* image columns (W)                          */
  double pixel_height;
  double pixel_width;
  double sigma_z;
  double sigma_dl;
  double center_y;
  double center_z;
} oracle_scanner;

/* FP32 MLEM, n_iter iterations (reconstruction.h:108-154).  rho is W*H,
 * row 0 on top.  rho_iters (nullable) receives the image after each
 * iteration.  counts (nullable) receives {box pixels, ellipse pixels} summed
 * over one iteration.  n_threads threads, dynamic schedule, private images
 * summed serially in thread order, as the reference does. */
int oracle_mlem_f32(const oracle_scanner* s, const float* events, int64_t n,
                    float* rho, int n_iter, int n_threads, float* rho_iters,
                    uint64_t* counts);

/* The same algorithm with every quantity in double (the FP64 twin). */
int oracle_mlem_f64(const oracle_scanner* s, const float* events, int64_t n,
                    double* rho, int n_iter, int n_threads,
                    double* rho_iters);

/* Per-event footprint: bounding-box pixels and in-ellipse pixels. */
int oracle_footprint(const oracle_scanner* s, const float* events, int64_t n,
                     int32_t* n_box, int32_t* n_ell, int n_threads);

/* Per-event denominators sum_p K*rho (FP32 path, sequential order). */
int oracle_denominators_f32(const oracle_scanner* s, const float* events,
                            int64_t n, const float* rho, float* denom,
                            int n_threads);

/* GaussianKernel::operator()(y, tan, sec, R, distance). */
float oracle_kernel_f32(double sigma_z, double sigma_dl, float y, float tan,
                        float sec, float R, float dz, float dy);
double oracle_kernel_f64(double sigma_z, double sigma_dl, double y, double tan,
                         double sec, double R, double dz, double dy);

/* AnalyticKernel::operator()(y, tan, sec, R, distance)
 * (analytic_kernel.h:28-62). */
float oracle_analytic_kernel_f32(double sigma_z, double sigma_dl, float y,
                                 float tan, float sec, float R, float dz,
                                 float dy);
double oracle_analytic_kernel_f64(double sigma_z, double sigma_dl, double y,
                                  double tan, double sec, double R, double dz,
                                  double dy);

/* The reference --gpu iteration in FP64 (cuda/reconstruction.cu:59-113,
 * warp_granularity.cuh:42-153): denominator sum K rho sens, update
 * rho' = sum_e K rho / D_e; kernel 0 = Gaussian, 1 = analytic; sens (W*H,
 * nullable = 1) is the caller's map, normally the 5-point pixel
 * sensitivity.  Same box and ellipse as oracle_mlem_f64. */
int oracle_mlem_variant_f64(const oracle_scanner* s, const float* events,
                            int64_t n, double* rho, int n_iter, int kernel,
                            const double* sens, int n_threads,
                            double* rho_iters);

/* GaussianKernel::ellipse_bb -> {sec, A, B, C, bb_y, bb_z}. */
void oracle_ellipse_bb_f32(double sigma_z, double sigma_dl, float tan,
                           float out[6]);
void oracle_ellipse_bb_f64(double sigma_z, double sigma_dl, double tan,
                           double out[6]);

/* Response::calculate_tan_y_z -> {tan, y, z}. */
void oracle_tan_y_z_f32(double radius, const float ev[3], float out[3]);
void oracle_tan_y_z_f64(double radius, const double ev[3], double out[3]);

/* Scanner::pixel_at -> {x, y}; Scanner::pixel_center -> {z, y}. */
void oracle_pixel_at_f32(const oracle_scanner* s, float z, float y,
                         int32_t out[2]);
void oracle_pixel_center_f32(const oracle_scanner* s, int32_t x, int32_t y,
                             float out[2]);
void oracle_pixel_at_f64(const oracle_scanner* s, double z, double y,
                         int32_t out[2]);
void oracle_pixel_center_f64(const oracle_scanner* s, int32_t x, int32_t y,
                             double out[2]);

/* Sensitivity image: mode 0 = Scanner::sensitivity at the pixel centre
 * (reconstruction.h:67-73), mode 1 = Scanner::pixel_sensitivity 5-point
 * average (scanner.h:174-185, cuda/reconstruction.cu:115-126). */
int oracle_sensitivity_f32(const oracle_scanner* s, int mode, float* out);
/* Scanner::sensitivity at one point, double. */
double oracle_sensitivity_point_f64(double radius, double length, double z,
                                    double y);

/* Poisson list-mode log-likelihood in FP64,
 *   l(rho, sigma) = sum_{e: lambda_e>0} log(lambda_e) - sum_p sens_p rho_p,
 *   lambda_e = sum_{p in E_e} K_e(p; sigma_k) rho_p,
 * with the support E_e (3-sigma ellipse inside the clamped box) computed at
 * sigma_support and the kernel at sigma_kernel ("frozen support").  sens
 * may be NULL (then sens = 1, the implicit sensitivity of the reference CPU
 * update).  Outputs are nullable: lambda (n), grad_rho (W*H) = dl/drho,
 * grad_sigma[2] = dl/d(sigma_z, sigma_dl) at sigma_kernel. */
int oracle_loglik_f64(const oracle_scanner* s, const float* events, int64_t n,
                      const double* rho, const double* sens,
                      double sigma_z_kernel, double sigma_dl_kernel,
                      double sigma_z_support, double sigma_dl_support,
                      double* loglik, double* lambda, double* grad_rho,
                      double grad_sigma[2], int n_threads);

#ifdef __cplusplus
}
#endif

#endif
