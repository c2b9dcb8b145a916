// HIP kernels of the strip-PET list-mode MLEM for gfx950 (wave64).
//
// Per iteration (reference: Reconstruction::operator(),
// src/2d/strip/reconstruction.h:108-154, and bb_update :220-298):
//   pass 1  k_denominator   one wave per event: lambda_e = sum K_e(p) rho[p]
//                           over the 3-sigma ellipse inside the clamped box,
//                           reduced across the wave with DPP; stores
//                           inv_e = lambda_e > 0 ? 1/lambda_e : 0.
//   pass 2  k_backproject   one wave per event: out[p] += K_e(p) rho[p] inv_e
//                           with hardware float atomics.
// Geometry (which pixels) is bit-identical to the reference CPU build
// (strip_geometry.h); the kernel value itself is evaluated in FP32 with the
// device exp/sqrt, within a few ulp of the host libm.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "strip_geometry.h"
#include "strip_kernels.h"
#include "wave_ops.h"

namespace petstrip {

// ---------------------------------------------------------------------------
// Wave64 helpers.

__device__ __forceinline__ int uniform_i(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// GaussianKernel::operator()(y, tan, sec, R, distance)
// (src/2d/strip/gaussian_kernel.h:51-79) in FP32.
__device__ __forceinline__ float kernel_value(const StripConst& k,
                                              const EventGeom& g, float dy,
                                              float dz) {
  float sec_sq = g.sec * g.sec;
  float yp = dy + g.y;
  float ap = -(yp - k.R) * sec_sq;
  float aq = -(k.R + yp) * sec_sq;
  float ar = ((yp * -2.0f) * g.sec) * g.t;
  float pq = fmaf(-dy, g.t, dz);
  float br = (dy * -2.0f) * g.sec;
  float norm = fmaf(ar * k.isdl, ar, fmaf(ap * k.isz, ap, (aq * k.isz) * aq));
  float bcb = fmaf((pq * k.isz) * pq, 2.0f, (br * k.isdl) * br);
  return (k.c / sqrtf(norm)) * expf(bcb * -0.5f);
}

// AnalyticKernel::operator()(y, tan, sec, R, distance)
// (src/2d/strip/analytic_kernel.h:28-62) in FP32: with a as for the Gaussian
// kernel, o = (a_p tan, a_q tan, -yp sec (1 + 2 tan^2)) and b, the kernel is
// c / sqrt(norm) exp(-(b S b - (b S a)^2 / norm) / 2), norm = a S a + 2 o S b.
// The support (box and 3-sigma ellipse) stays the Gaussian one, as in the
// reference.
__device__ __forceinline__ float analytic_value(const StripConst& k,
                                                const EventGeom& g, float dy,
                                                float dz) {
  const float sec_sq = g.sec * g.sec;
  const float yp = dy + g.y;
  const float ap = -(yp - k.R) * sec_sq, aq = -(yp + k.R) * sec_sq;
  const float ar = -2.0f * yp * g.sec * g.t;
  const float op = ap * g.t, oq = aq * g.t;
  const float orr = -yp * g.sec * (1.0f + 2.0f * g.t * g.t);
  const float pq = dz - dy * g.t;
  const float br = -2.0f * dy * g.sec;
  const float aia = ap * k.isz * ap + aq * k.isz * aq + ar * k.isdl * ar;
  const float bia = pq * k.isz * ap + pq * k.isz * aq + br * k.isdl * ar;
  const float bib = pq * k.isz * pq + pq * k.isz * pq + br * k.isdl * br;
  const float oib = op * k.isz * pq + oq * k.isz * pq + orr * k.isdl * br;
  const float norm = aia + 2.0f * oib;
  return k.c / sqrtf(norm) * expf(-0.5f * (bib - bia * bia / norm));
}

__device__ __forceinline__ float event_kernel(const StripConst& k, const EventGeom& g,
                                              float dy, float dz, int analytic) {
  return analytic ? analytic_value(k, g, dy, dz) : kernel_value(k, g, dy, dz);
}

__device__ __forceinline__ EventGeom load_geom(const EventGeom* geom, int64_t e) {
  const float4* p = reinterpret_cast<const float4*>(geom + e);
  float4 a = p[0], b = p[1];
  int4 c = reinterpret_cast<const int4*>(p)[2];
  EventGeom g;
  g.z = a.x; g.y = a.y; g.t = a.z; g.sec = a.w;
  g.A = b.x; g.B = b.y; g.C = b.z; g.pad = b.w;
  g.x0 = c.x; g.x1 = c.y; g.y0 = c.z; g.y1 = c.w;
  return g;
}

// ---------------------------------------------------------------------------
// Setup: responses -> geometry, footprint counts.

__global__ __launch_bounds__(256) void k_event_setup(StripConst s,
                                                     const float* __restrict__ resp,
                                                     int64_t n,
                                                     EventGeom* __restrict__ geom) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  EventGeom g;
  event_geometry(s, resp[3 * e], resp[3 * e + 1], resp[3 * e + 2], g);
  geom[e] = g;
}

// One wave per event: box and ellipse pixel counts.
__global__ __launch_bounds__(256) void k_footprint(StripConst s,
                                                   const EventGeom* __restrict__ geom,
                                                   int64_t n,
                                                   int32_t* __restrict__ n_box,
                                                   int32_t* __restrict__ n_ell) {
  const int lane = threadIdx.x & 63;
  int64_t e = (int64_t)blockIdx.x * (blockDim.x >> 6) + uniform_i(threadIdx.x >> 6);
  if (e >= n) return;
  EventGeom g = load_geom(geom, e);
  int bw = g.x1 - g.x0, bh = g.y1 - g.y0;
  int nbox = (bw > 0 && bh > 0) ? bw * bh : 0;
  int cnt = 0;
  for (int k = lane; k < nbox; k += 64) {
    int row = k / bw, col = k - row * bw;
    float dy = row_dy(s, g, g.y0 + row);
    float dz = col_dz(s, g, g.x0 + col);
    cnt += in_ellipse(g, dy, dz) ? 1 : 0;
  }
  float tot = wave_sum((float)cnt);
  if (lane == 0) {
    n_box[e] = nbox;
    n_ell[e] = (int32_t)tot;
  }
}

// ---------------------------------------------------------------------------
// Pass 1: denominators.  weight (nullable) multiplies rho (the reference
// --gpu variant weights the denominator by the pixel sensitivity).

// comp: each lane's terms and the wave's lanes summed with two-sums (the
// compensated accumulation mode); otherwise plain FP32 in a fixed order.
__global__ __launch_bounds__(256) void k_denominator(StripConst sg, StripConst sk,
                                                     const EventGeom* __restrict__ geom,
                                                     int64_t n,
                                                     const float* __restrict__ rho,
                                                     const float* __restrict__ weight,
                                                     float* __restrict__ lambda,
                                                     float* __restrict__ inv,
                                                     int analytic, int comp) {
  const int lane = threadIdx.x & 63;
  int64_t e = (int64_t)blockIdx.x * (blockDim.x >> 6) + uniform_i(threadIdx.x >> 6);
  if (e >= n) return;
  EventGeom g = load_geom(geom, e);
  int bw = g.x1 - g.x0, bh = g.y1 - g.y0;
  int nbox = (bw > 0 && bh > 0) ? bw * bh : 0;
  float acc = 0.0f, err = 0.0f;
  for (int k = lane; k < nbox; k += 64) {
    int row = k / bw, col = k - row * bw;
    int iy = g.y0 + row, iz = g.x0 + col;
    float dy = row_dy(sg, g, iy);
    float dz = col_dz(sg, g, iz);
    if (in_ellipse(g, dy, dz)) {
      int p = iy * sg.W + iz;
      float r = rho[p];
      if (weight) r *= weight[p];
      const float t = event_kernel(sk, g, dy, dz, analytic) * r;
      if (comp) {
        float s, d;
        two_sum(acc, t, s, d);
        acc = s;
        err += d;
      } else {
        acc += t;
      }
    }
  }
  float d = comp ? wave_sum_comp(acc, err) : wave_sum(acc);
  if (lane == 0) {
    lambda[e] = d;
    inv[e] = (d > 0.0f) ? 1.0f / d : 0.0f;
  }
}

// ---------------------------------------------------------------------------
// Pass 2: backprojection out[p] += K rho[p] * w_e (w_e = inv_e for MLEM,
// rho == nullptr gives out[p] += K * w_e, the adjoint of the projector).
// The sums are FP64 (a pixel can take millions of adds at 1e8 events), or,
// with out_u64, exact 64-bit integer sums of the non-negative values in
// units 1 / fixed_scale (run-to-run identical images).

__global__ __launch_bounds__(256) void k_backproject(StripConst sg, StripConst sk,
                                                     const EventGeom* __restrict__ geom,
                                                     int64_t n,
                                                     const float* __restrict__ rho,
                                                     const float* __restrict__ w,
                                                     double* __restrict__ out,
                                                     int analytic,
                                                     unsigned long long* __restrict__ out_u64,
                                                     double fixed_scale) {
  const int lane = threadIdx.x & 63;
  int64_t e = (int64_t)blockIdx.x * (blockDim.x >> 6) + uniform_i(threadIdx.x >> 6);
  if (e >= n) return;
  float we = w[e];
  if (we == 0.0f) return;
  EventGeom g = load_geom(geom, e);
  int bw = g.x1 - g.x0, bh = g.y1 - g.y0;
  int nbox = (bw > 0 && bh > 0) ? bw * bh : 0;
  for (int k = lane; k < nbox; k += 64) {
    int row = k / bw, col = k - row * bw;
    int iy = g.y0 + row, iz = g.x0 + col;
    float dy = row_dy(sg, g, iy);
    float dz = col_dz(sg, g, iz);
    if (in_ellipse(g, dy, dz)) {
      int p = iy * sg.W + iz;
      float kv = event_kernel(sk, g, dy, dz, analytic);
      if (rho) kv *= rho[p];
      if (out_u64) {
        const double v = (double)(kv * we);
        if (v > 0.0) atomicAdd(out_u64 + p, (unsigned long long)fma(v, fixed_scale, 0.5));
      } else
        unsafeAtomicAdd(&out[p], (double)(kv * we));
    }
  }
}

// ---------------------------------------------------------------------------
// Sigma derivative of lambda_e weighted by g_e, summed over events:
//   dlnK/dsz  = (-2 + isz (ap^2+aq^2)/norm + 2 isz pq^2) / sz
//   dlnK/dsdl = (-1 + isdl ar^2/norm + isdl br^2) / sdl
// accumulated in FP64 per wave, one atomic per wave.

__global__ __launch_bounds__(256) void k_sigma_grad(StripConst sg, StripConst sk,
                                                    const EventGeom* __restrict__ geom,
                                                    int64_t n,
                                                    const float* __restrict__ rho,
                                                    const float* __restrict__ w,
                                                    double* __restrict__ out2) {
  const int lane = threadIdx.x & 63;
  int64_t e = (int64_t)blockIdx.x * (blockDim.x >> 6) + uniform_i(threadIdx.x >> 6);
  if (e >= n) return;
  float we = w[e];
  if (we == 0.0f) return;
  EventGeom g = load_geom(geom, e);
  int bw = g.x1 - g.x0, bh = g.y1 - g.y0;
  int nbox = (bw > 0 && bh > 0) ? bw * bh : 0;
  float az = 0.0f, adl = 0.0f;
  for (int k = lane; k < nbox; k += 64) {
    int row = k / bw, col = k - row * bw;
    int iy = g.y0 + row, iz = g.x0 + col;
    float dy = row_dy(sg, g, iy);
    float dz = col_dz(sg, g, iz);
    if (in_ellipse(g, dy, dz)) {
      int p = iy * sg.W + iz;
      float sec_sq = g.sec * g.sec;
      float yp = dy + g.y;
      float ap = -(yp - sk.R) * sec_sq;
      float aq = -(sk.R + yp) * sec_sq;
      float ar = ((yp * -2.0f) * g.sec) * g.t;
      float pq = fmaf(-dy, g.t, dz);
      float br = (dy * -2.0f) * g.sec;
      float apq = ap * ap + aq * aq;
      float norm = sk.isz * apq + sk.isdl * (ar * ar);
      float bsb_z = 2.0f * sk.isz * (pq * pq);
      float bsb_dl = sk.isdl * (br * br);
      float K = (sk.c / sqrtf(norm)) * expf(-0.5f * (bsb_z + bsb_dl));
      float kr = K * rho[p];
      az += kr * (-2.0f + sk.isz * apq / norm + bsb_z);
      adl += kr * (-1.0f + sk.isdl * (ar * ar) / norm + bsb_dl);
    }
  }
  double sz = wave_sum_d((double)az) * (double)we / (double)sk.sigma_z;
  double sdl = wave_sum_d((double)adl) * (double)we / (double)sk.sigma_dl;
  if (lane == 0) {
    unsafeAtomicAdd(&out2[0], sz);
    unsafeAtomicAdd(&out2[1], sdl);
  }
}

// ---------------------------------------------------------------------------
// Analytic-kernel map of the phantom command's tan bins
// (src/2d/strip/phantom_cmd.cpp:396-440): voxel (x, y, plane) of a
// column x row x tan-plane grid (PixelGrid / VariableVoxelSizeVoxelGrid
// centres, row 0 at the bottom) holds
//   K(centre as the event point, tan, sec; kernel point) x 4 R pixel^2 dtan sec
//   x scale,
// with dtan = 2 max_tan / planes and K the analytic kernel.

__global__ __launch_bounds__(256) void k_tan_kernel_map(StripConst k, int w, int h,
                                                        int planes, float pixel,
                                                        float llx, float lly,
                                                        float max_tan, float kpx,
                                                        float kpy, float scale,
                                                        float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)w * h * planes;
  if (i >= n) return;
  const int x = (int)(i % w), y = (int)((i / w) % h), z = (int)(i / ((int64_t)w * h));
  const float dtan = max_tan / planes * 2;
  const float t = (z + 0.5f) * dtan + -max_tan;
  const float sec = sqrtf(1 + t * t);
  const float cx = llx + pixel / 2 + x * pixel, cy = lly + pixel / 2 + y * pixel;
  EventGeom g{};
  g.y = cy;
  g.t = t;
  g.sec = sec;
  const float kv = analytic_value(k, g, kpy - cy, kpx - cx);
  const float volume = 4 * k.R * (pixel * pixel * dtan) * sqrtf(1 + t * t);
  out[i] = kv * volume * scale;
}

// ---------------------------------------------------------------------------
// Elementwise helpers.

__global__ void k_fill(float* __restrict__ p, float v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_zero_u64(unsigned long long* p) { *p = 0ull; }

// Adds to *count the entries of x that are not > 0 (zero, negative or NaN
// denominators): a grid-stride count per lane, a wave sum, one atomic per
// wave.
__global__ void k_count_nonpositive(const float* __restrict__ x, int64_t n,
                                    unsigned long long* count) {
  long long c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    c += !(x[i] > 0.0f);
  for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
  if (__lane_id() == 0 && c) atomicAdd(count, (unsigned long long)c);
}

__global__ void k_add_scalar(float* __restrict__ p, float v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] += v;
}

// g_e = lambda_e > 0 ? 1/lambda_e : 0 and the log-likelihood data term.
__global__ __launch_bounds__(256) void k_loglik_weights(const float* __restrict__ lambda,
                                                        int64_t n,
                                                        float* __restrict__ g,
                                                        double* __restrict__ ll) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (e < n) {
    float l = lambda[e];
    g[e] = (l > 0.0f) ? 1.0f / l : 0.0f;
    if (l > 0.0f) v = log((double)l);
  }
  v = wave_sum_d(v);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(ll, v);
}

// Scanner::sensitivity (scanner.h:147-171) at a point.
__device__ float sens_point(const StripConst& s, float z, float y) {
  float L_plus = s.half_L + z;
  float L_minus = s.half_L - z;
  float R_plus = s.R + y;
  float R_minus = s.R - y;
  if (R_plus <= 0) R_plus = 1;
  if (R_minus <= 0) R_minus = 1;
  float a = L_minus / R_minus, b = L_plus / R_plus;
  float mn = (b < a) ? b : a;
  float c = -L_plus / R_minus, d = -L_minus / R_plus;
  float mx = (c < d) ? d : c;
  return 0.318309886183790671538f * (atanf(mn) - atanf(mx));
}

__global__ void k_sensitivity(StripConst s, int mode, float* __restrict__ out) {
  int x = blockIdx.x * blockDim.x + threadIdx.x;
  int y = blockIdx.y;
  if (x >= s.W) return;
  float pz = fmaf((float)x, s.pw, s.tl_z_half_w);
  float py = fmaf(-(float)y, s.ph, s.tl_y_half_h);
  float v;
  if (mode == 0) {
    v = sens_point(s, pz, py);
  } else {
    float hw = s.pw / 2, hh = s.ph / 2;
    v = sens_point(s, pz, py) / 3 + sens_point(s, pz + hw, py + hh) / 6 +
        sens_point(s, pz - hw, py - hh) / 6 + sens_point(s, pz - hw, py + hh) / 6 +
        sens_point(s, pz + hw, py - hh) / 6;
  }
  out[(size_t)y * s.W + x] = v;
}

// ---------------------------------------------------------------------------
// Launch wrappers (host side).

static inline unsigned blocks_for(int64_t n, int per_block) {
  return (unsigned)((n + per_block - 1) / per_block);
}

// A dispatch's grid is counted in work-items in 32 bits, so a launch of one
// wave per event covers at most 2^24 events (2^30 work-items); larger event
// sets go in slices, each kernel seeing its slice's pointers.
constexpr int64_t kWaveSlice = (int64_t)1 << 24;

hipError_t launch_event_setup(hipStream_t st, const StripConst& s,
                              const float* resp, int64_t n, EventGeom* geom) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_event_setup, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                     s, resp, n, geom);
  return hipGetLastError();
}

hipError_t launch_footprint(hipStream_t st, const StripConst& s,
                            const EventGeom* geom, int64_t n, int32_t* nbox,
                            int32_t* nell) {
  if (n <= 0) return hipSuccess;
  for (int64_t e0 = 0; e0 < n; e0 += kWaveSlice) {
    const int64_t m = std::min(kWaveSlice, n - e0);
    hipLaunchKernelGGL(k_footprint, dim3(blocks_for(m, 4)), dim3(256), 0, st, s,
                       geom + e0, m, nbox + e0, nell + e0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_denominator(hipStream_t st, const StripConst& sg,
                              const StripConst& sk, const EventGeom* geom,
                              int64_t n, const float* rho, const float* weight,
                              float* lambda, float* inv, int analytic, int comp) {
  if (n <= 0) return hipSuccess;
  for (int64_t e0 = 0; e0 < n; e0 += kWaveSlice) {
    const int64_t m = std::min(kWaveSlice, n - e0);
    hipLaunchKernelGGL(k_denominator, dim3(blocks_for(m, 4)), dim3(256), 0, st,
                       sg, sk, geom + e0, m, rho, weight, lambda + e0, inv + e0, analytic,
                       comp);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_backproject(hipStream_t st, const StripConst& sg,
                              const StripConst& sk, const EventGeom* geom,
                              int64_t n, const float* rho, const float* w,
                              double* out, int analytic, unsigned long long* out_u64,
                              double fixed_scale) {
  if (n <= 0) return hipSuccess;
  for (int64_t e0 = 0; e0 < n; e0 += kWaveSlice) {
    const int64_t m = std::min(kWaveSlice, n - e0);
    hipLaunchKernelGGL(k_backproject, dim3(blocks_for(m, 4)), dim3(256), 0, st,
                       sg, sk, geom + e0, m, rho, w + e0, out, analytic, out_u64,
                       fixed_scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_sigma_grad(hipStream_t st, const StripConst& sg,
                             const StripConst& sk, const EventGeom* geom,
                             int64_t n, const float* rho, const float* w,
                             double* out2) {
  if (n <= 0) return hipSuccess;
  for (int64_t e0 = 0; e0 < n; e0 += kWaveSlice) {
    const int64_t m = std::min(kWaveSlice, n - e0);
    hipLaunchKernelGGL(k_sigma_grad, dim3(blocks_for(m, 4)), dim3(256), 0, st,
                       sg, sk, geom + e0, m, rho, w + e0, out2);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_tan_kernel_map(hipStream_t st, const StripConst& k, int w, int h,
                                 int planes, float pixel, float llx, float lly,
                                 float max_tan, float kpx, float kpy, float scale,
                                 float* out) {
  const int64_t n = (int64_t)w * h * planes;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tan_kernel_map, dim3(blocks_for(n, 256)), dim3(256), 0, st, k, w, h,
                     planes, pixel, llx, lly, max_tan, kpx, kpy, scale, out);
  return hipGetLastError();
}

hipError_t launch_fill(hipStream_t st, float* p, float v, int64_t n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n, 256)), dim3(256), 0, st, p, v, n);
  return hipGetLastError();
}

hipError_t launch_count_nonpositive(hipStream_t st, const float* x, int64_t n,
                                    unsigned long long* count, bool clear) {
  if (clear) hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(1), 0, st, count);
  if (n > 0)
    hipLaunchKernelGGL(k_count_nonpositive, dim3(std::min<unsigned>(blocks_for(n, 256), 2048)),
                       dim3(256), 0, st, x, n, count);
  return hipGetLastError();
}

hipError_t launch_add_scalar(hipStream_t st, float* p, float v, int64_t n) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_add_scalar, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                     p, v, n);
  return hipGetLastError();
}

hipError_t launch_loglik_weights(hipStream_t st, const float* lambda, int64_t n,
                                 float* g, double* ll) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_loglik_weights, dim3(blocks_for(n, 256)), dim3(256), 0,
                     st, lambda, n, g, ll);
  return hipGetLastError();
}

hipError_t launch_sensitivity(hipStream_t st, const StripConst& s, int mode,
                              float* out) {
  hipLaunchKernelGGL(k_sensitivity, dim3(blocks_for(s.W, 128), s.H), dim3(128),
                     0, st, s, mode, out);
  return hipGetLastError();
}

}  // namespace petstrip
