// Launch wrappers of the strip-PET HIP kernels (strip_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "strip_geometry.h"

namespace petstrip {

hipError_t launch_event_setup(hipStream_t st, const StripConst& s,
                              const float* resp, int64_t n, EventGeom* geom);
hipError_t launch_footprint(hipStream_t st, const StripConst& s,
                            const EventGeom* geom, int64_t n, int32_t* nbox,
                            int32_t* nell);
// comp = 1: compensated (two-sum) denominators.
hipError_t launch_denominator(hipStream_t st, const StripConst& sg,
                              const StripConst& sk, const EventGeom* geom,
                              int64_t n, const float* rho, const float* weight,
                              float* lambda, float* inv, int analytic = 0, int comp = 0);
// out (FP64 sums), or out_u64 (non-negative values as integers in units
// 1 / fixed_scale) when given.
hipError_t launch_backproject(hipStream_t st, const StripConst& sg,
                              const StripConst& sk, const EventGeom* geom,
                              int64_t n, const float* rho, const float* w,
                              double* out, int analytic = 0,
                              unsigned long long* out_u64 = nullptr, double fixed_scale = 0.0);
// Analytic-kernel map of tan bins (w x h x planes floats into out).
hipError_t launch_tan_kernel_map(hipStream_t st, const StripConst& k, int w, int h,
                                 int planes, float pixel, float llx, float lly,
                                 float max_tan, float kpx, float kpy, float scale,
                                 float* out);
hipError_t launch_sigma_grad(hipStream_t st, const StripConst& sg,
                             const StripConst& sk, const EventGeom* geom,
                             int64_t n, const float* rho, const float* w,
                             double* out2);
hipError_t launch_fill(hipStream_t st, float* p, float v, int64_t n);
hipError_t launch_add_scalar(hipStream_t st, float* p, float v, int64_t n);
// *count = the number of entries of x[0, n) that are not > 0 (the counter
// is cleared by a kernel first, unless clear is false: then it must hold 0).
hipError_t launch_count_nonpositive(hipStream_t st, const float* x, int64_t n,
                                    unsigned long long* count, bool clear = true);
hipError_t launch_loglik_weights(hipStream_t st, const float* lambda,
                                 int64_t n, float* g, double* ll);
hipError_t launch_sensitivity(hipStream_t st, const StripConst& s, int mode,
                              float* out);

}  // namespace petstrip
