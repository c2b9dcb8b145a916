// Fused MLEM iteration for gfx950: events sorted into spatial bins, one
// workgroup per chunk of a bin, rho and the backprojection held in LDS.
//
// Reference semantics: Reconstruction::bb_update
// (src/2d/strip/reconstruction.h:220-298) per event, Jacobi over events
// (reconstruction.h:121-149).  One wave handles one event at a time; its
// lanes take the event's box rows (64 at a time).  Each lane
//   1. finds its row's exact ellipse column interval (in_ellipse with the
//      reference's float operation order, gaussian_kernel.h:131-137) and
//      the row factor of the kernel;
//   2. pass 1: walks the interval summing K * rho read from the LDS window;
//      the DPP wave sum is the denominator lambda_e;
//   3. pass 2: walks it again adding K * rho / lambda_e into the LDS
//      accumulation window (ds_add), kernel values recomputed.
// The inner walks are branch-free and unrolled by four so that each lane
// keeps several LDS reads and exponentials in flight.  The window pitch is
// odd so that lanes on neighbouring rows of a tilted ellipse fall in
// different LDS banks.  At the end the window is added to the global image
// with one atomic per touched pixel.
//
// The kernel value follows GaussianKernel::operator()
// (gaussian_kernel.h:51-79) with its exponent split in a row part and a
// pixel part: K = [c / sqrt(a^T S a) * exp(-isdl br^2 / 2)] * exp(-isz pq^2),
// the pixel part evaluated as exp2(-(pq * sqrt(isz log2 e))^2) with pq
// computed from the pixel centre exactly as the reference does.
#include <hip/hip_runtime.h>

#include "strip_fused.h"
#include "strip_geometry.h"

namespace petstrip {

namespace {

__device__ __forceinline__ int uniform(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// Wave64 sum, fixed order: DPP within rows of 16, then the four row sums.
__device__ __forceinline__ float wave_sum(float v) {
  v += __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xb1, 0xf, 0xf, false));
  v += __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4e, 0xf, 0xf, false));
  v += __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xf, 0xf, false));
  v += __int_as_float(
      __builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xf, 0xf, false));
  float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
  float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
  float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
  float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
  return (r0 + r1) + (r2 + r3);
}

__device__ __forceinline__ int wave_max(int v) {
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return uniform(v);
}

__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Compensated wave sum of (hi, lo) pairs; partners combine in lane order so
// both compute the same value.
__device__ __forceinline__ float wave_sum_comp(float hi, float lo) {
  for (int off = 1; off < 64; off <<= 1) {
    float oh = __shfl_xor(hi, off, 64);
    float ol = __shfl_xor(lo, off, 64);
    bool upper = (__lane_id() & off) != 0;
    float a = upper ? oh : hi, b = upper ? hi : oh;
    float la = upper ? ol : lo, lb = upper ? lo : ol;
    float s, e;
    two_sum(a, b, s, e);
    hi = s;
    lo = e + (la + lb);
  }
  return hi + lo;
}

// Exact reference membership of column iz in the row with (Ady2, Bdy).
__device__ __forceinline__ bool exact_in(const StripConst& s, const EventGeom& g,
                                         float Ady2, float Bdy, int iz) {
  float dz = fmaf((float)iz, s.pw, s.tl_z_half_w) - g.z;
  float q = fmaf(g.C * dz, dz, fmaf(Bdy, dz, Ady2));
  return q <= 9.0f;
}

// One box row of an event: its ellipse interval and kernel row factor.
struct RowSeg {
  int lo, len;  // first ellipse column, ellipse pixels in the row
  float ndy;    // -(pixel_center.y - y_e)
  float scale;  // c / sqrt(a^T S a) * exp(-isdl br^2 / 2)
};

__device__ __forceinline__ RowSeg row_setup(const StripConst& sg, const StripConst& sk,
                                            const EventGeom& g, int iy) {
  RowSeg s;
  float dy = fmaf(-(float)iy, sg.ph, sg.tl_y_half_h) - g.y;
  float Ady2 = (dy * g.A) * dy;
  float Bdy = dy * g.B;
  s.ndy = -dy;
  s.len = 0;
  s.lo = g.x0;
  s.scale = 0.0f;
  // Superset of the interval from the real roots of C dz^2 + Bdy dz + Ady2 =
  // 9.05, then the exact test trims both ends.  Along a row the exact test
  // has no holes: adjacent columns differ in q by at least C pw^2 (checked
  // on the host), far above the rounding of q.
  float disc = Bdy * Bdy - 4.0f * g.C * (Ady2 - 9.05f);
  if (!(disc >= 0.0f)) return s;
  float sq = sqrtf(disc);
  float inv2c = 0.5f / g.C;
  float zlo = g.z + (-Bdy - sq) * inv2c;
  float zhi = g.z + (-Bdy + sq) * inv2c;
  float flo = floorf((zlo - sg.tl_z_half_w) / sg.pw) - 1.0f;
  float fhi = ceilf((zhi - sg.tl_z_half_w) / sg.pw) + 1.0f;
  int a = (flo > (float)g.x0) ? (int)flo : g.x0;
  int b = (fhi < (float)(g.x1 - 1)) ? (int)fhi : g.x1 - 1;
  while (a <= b && !exact_in(sg, g, Ady2, Bdy, a)) ++a;
  while (b >= a && !exact_in(sg, g, Ady2, Bdy, b)) --b;
  if (a > b) return s;
  s.lo = a;
  s.len = b - a + 1;
  // kernel row factor (gaussian_kernel.h:58-78)
  float sec_sq = g.sec * g.sec;
  float yp = dy + g.y;
  float ap = -(yp - sk.R) * sec_sq;
  float aq = -(sk.R + yp) * sec_sq;
  float ar = ((yp * -2.0f) * g.sec) * g.t;
  float br = (dy * -2.0f) * g.sec;
  float norm = fmaf(ar * sk.isdl, ar, fmaf(ap * sk.isz, ap, (aq * sk.isz) * aq));
  s.scale = (sk.c / sqrtf(norm)) * expf(-0.5f * ((br * sk.isdl) * br));
  return s;
}

template <int ACC>
struct AccType;
template <>
struct AccType<ACC_LDS_F32> {
  using T = float;
};
template <>
struct AccType<ACC_LDS_FIXED64> {
  using T = unsigned long long;
};

constexpr float kFixedScale = 1099511627776.0f;  // 2^40
constexpr int kUnroll = 4;

// exp(-isz pq^2) for the pixel at column izf of a row (reference pq).
__device__ __forceinline__ float pixel_gauss(const StripConst& sg, const EventGeom& g,
                                             float izf, float ndy, float s_pix) {
  const float dz = fmaf(izf, sg.pw, sg.tl_z_half_w) - g.z;
  const float u = fmaf(ndy, g.t, dz) * s_pix;
  return __builtin_amdgcn_exp2f(-(u * u));
}

}  // namespace

// One workgroup per work item (a chunk of events sharing a spatial bin).
template <int ACC, int WAVES, bool COMP>
__global__ __launch_bounds__(64 * WAVES) void k_mlem_fused(
    StripConst sg, StripConst sk, const EventGeom* __restrict__ sgeom,
    const int32_t* __restrict__ sidx, const WorkItem* __restrict__ items,
    const float* __restrict__ rho, float* __restrict__ lambda,
    void* __restrict__ acc_global, int acc_global_kind) {
  using A = typename AccType<ACC>::T;
  constexpr int kThreads = 64 * WAVES;
  extern __shared__ __align__(16) unsigned char smem[];
  const WorkItem it = items[blockIdx.x];
  const int ww = it.ww, wh = it.wh;
  const int pitch = ww | 1;  // odd pitch: rows of a tilted ellipse spread over banks
  const int npx = pitch * wh;
  float* rho_w = reinterpret_cast<float*>(smem);
  A* acc_w = reinterpret_cast<A*>(smem + (((size_t)npx * 4 + 15) & ~(size_t)15));
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = __lane_id();

  // load the rho window, clear the accumulation window
  for (int i = threadIdx.x; i < ww * wh; i += kThreads) {
    int r = i / ww, c = i - r * ww;
    rho_w[r * pitch + c] = rho[(size_t)(it.wy0 + r) * sg.W + (it.wx0 + c)];
  }
  for (int i = threadIdx.x; i < npx; i += kThreads) acc_w[i] = (A)0;
  __syncthreads();

  const float s_pix = sqrtf(sk.isz * 1.44269504088896341f);  // sqrt(isz log2 e)

  for (int e = it.begin + wave; e < it.end; e += WAVES) {
    const EventGeom g = sgeom[e];
    const int nrows = g.y1 - g.y0;
    const int row0 = (g.y0 - it.wy0) * pitch - it.wx0;  // window index of (g.y0, 0)

    // ---- pass 1: denominator -------------------------------------------------
    float acc = 0.0f, comp = 0.0f;
    for (int base = 0; base < nrows; base += 64) {
      const int r = base + lane;
      RowSeg s{0, 0, 0.0f, 0.0f};
      if (r < nrows) s = row_setup(sg, sk, g, g.y0 + r);
      const int w0 = s.len > 0 ? row0 + r * pitch + s.lo : 0;
      const int maxlen = wave_max(s.len);
      const float lof = (float)s.lo;
      for (int k = 0; k < maxlen; k += kUnroll) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int kk = k + u;
          const bool on = kk < s.len;
          const float gv = pixel_gauss(sg, g, lof + (float)kk, s.ndy, s_pix);
          const float kr = gv * s.scale * rho_w[on ? w0 + kk : w0];
          if constexpr (COMP) {
            float sum, err;
            two_sum(acc, on ? kr : 0.0f, sum, err);
            acc = sum;
            comp += err;
          } else {
            acc += on ? kr : 0.0f;
          }
        }
      }
    }
    const float lam = COMP ? wave_sum_comp(acc, comp) : wave_sum(acc);
    if (lane == 0) lambda[sidx[e]] = lam;
    if (!(lam > 0.0f)) continue;
    const float inv = 1.0f / lam;

    // ---- pass 2: backprojection into the LDS window --------------------------
    for (int base = 0; base < nrows; base += 64) {
      const int r = base + lane;
      RowSeg s{0, 0, 0.0f, 0.0f};
      if (r < nrows) s = row_setup(sg, sk, g, g.y0 + r);
      const int w0 = s.len > 0 ? row0 + r * pitch + s.lo : 0;
      const int maxlen = wave_max(s.len);
      const float lof = (float)s.lo;
      const float sc = s.scale * inv;
      for (int k = 0; k < maxlen; k += kUnroll) {
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int kk = k + u;
          if (kk < s.len) {
            const float gv = pixel_gauss(sg, g, lof + (float)kk, s.ndy, s_pix);
            const float v = gv * sc * rho_w[w0 + kk];
            if constexpr (ACC == ACC_LDS_F32) {
              atomicAdd(&acc_w[w0 + kk], v);
            } else {
              atomicAdd(&acc_w[w0 + kk],
                        (unsigned long long)__float2ull_rn(v * kFixedScale));
            }
          }
        }
      }
    }
  }
  __syncthreads();

  // ---- flush the window into the global accumulator ------------------------
  for (int i = threadIdx.x; i < ww * wh; i += kThreads) {
    int r = i / ww, c = i - r * ww;
    A v = acc_w[r * pitch + c];
    if (v == (A)0) continue;
    size_t p = (size_t)(it.wy0 + r) * sg.W + (it.wx0 + c);
    if constexpr (ACC == ACC_LDS_F32) {
      if (acc_global_kind == GLOBAL_F64)
        unsafeAtomicAdd(reinterpret_cast<double*>(acc_global) + p, (double)v);
      else
        unsafeAtomicAdd(reinterpret_cast<float*>(acc_global) + p, v);
    } else {
      // the window holds exact 2^-40 fixed-point sums
      if (acc_global_kind == GLOBAL_U64)
        atomicAdd(reinterpret_cast<unsigned long long*>(acc_global) + p, v);
      else if (acc_global_kind == GLOBAL_F64)
        unsafeAtomicAdd(reinterpret_cast<double*>(acc_global) + p,
                        (double)v * (1.0 / 1099511627776.0));
      else
        unsafeAtomicAdd(reinterpret_cast<float*>(acc_global) + p,
                        (float)((double)v * (1.0 / 1099511627776.0)));
    }
  }
}

// ---------------------------------------------------------------------------

size_t fused_lds_bytes(int max_window_px, int acc_mode, int max_rows, int waves) {
  (void)max_rows;
  (void)waves;
  size_t npx = (size_t)max_window_px;
  size_t accb = acc_mode == ACC_LDS_F32 ? 4 : 8;
  return ((npx * 4 + 15) & ~(size_t)15) + npx * accb;
}

namespace {

template <int ACC, int WAVES, bool COMP>
hipError_t launch_t(hipStream_t st, const StripConst& sg, const StripConst& sk,
                    const EventGeom* sgeom, const int32_t* sidx, const WorkItem* items,
                    int n_items, size_t lds_bytes, const float* rho, float* lambda,
                    void* acc_global, int acc_global_kind) {
  auto kern = k_mlem_fused<ACC, WAVES, COMP>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds_bytes);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(n_items), dim3(64 * WAVES), lds_bytes, st, sg, sk,
                     sgeom, sidx, items, rho, lambda, acc_global, acc_global_kind);
  return hipGetLastError();
}

template <int ACC, bool COMP>
hipError_t launch_w(int waves, hipStream_t st, const StripConst& sg,
                    const StripConst& sk, const EventGeom* sgeom, const int32_t* sidx,
                    const WorkItem* items, int n_items, size_t lds_bytes,
                    const float* rho, float* lambda, void* acc_global,
                    int acc_global_kind) {
  switch (waves) {
    case 16:
      return launch_t<ACC, 16, COMP>(st, sg, sk, sgeom, sidx, items, n_items, lds_bytes,
                                     rho, lambda, acc_global, acc_global_kind);
    case 8:
      return launch_t<ACC, 8, COMP>(st, sg, sk, sgeom, sidx, items, n_items, lds_bytes,
                                    rho, lambda, acc_global, acc_global_kind);
    default:
      return launch_t<ACC, 4, COMP>(st, sg, sk, sgeom, sidx, items, n_items, lds_bytes,
                                    rho, lambda, acc_global, acc_global_kind);
  }
}

}  // namespace

hipError_t launch_mlem_fused(hipStream_t st, const StripConst& sg, const StripConst& sk,
                             const EventGeom* sgeom, const int32_t* sidx,
                             const WorkItem* items, int n_items, size_t lds_bytes,
                             const float* rho, float* lambda, void* acc_global,
                             int acc_global_kind, int acc_mode, int max_rows,
                             int compensated, int waves) {
  (void)max_rows;
  if (n_items <= 0) return hipSuccess;
  if (acc_mode == ACC_LDS_F32) {
    if (compensated)
      return launch_w<ACC_LDS_F32, true>(waves, st, sg, sk, sgeom, sidx, items, n_items,
                                         lds_bytes, rho, lambda, acc_global,
                                         acc_global_kind);
    return launch_w<ACC_LDS_F32, false>(waves, st, sg, sk, sgeom, sidx, items, n_items,
                                        lds_bytes, rho, lambda, acc_global,
                                        acc_global_kind);
  }
  if (compensated)
    return launch_w<ACC_LDS_FIXED64, true>(waves, st, sg, sk, sgeom, sidx, items, n_items,
                                           lds_bytes, rho, lambda, acc_global,
                                           acc_global_kind);
  return launch_w<ACC_LDS_FIXED64, false>(waves, st, sg, sk, sgeom, sidx, items, n_items,
                                          lds_bytes, rho, lambda, acc_global,
                                          acc_global_kind);
}

}  // namespace petstrip
