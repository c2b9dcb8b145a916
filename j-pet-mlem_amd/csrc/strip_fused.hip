// Fused passes over spatially binned events for gfx950: one workgroup per
// chunk of a bin, the chunk's image window held in LDS.
//
// Reference semantics: Reconstruction::bb_update
// (src/2d/strip/reconstruction.h:220-298) per event, Jacobi over events
// (reconstruction.h:121-149).  One wave handles one event at a time; its
// lanes take the event's box rows (64 at a time).  Each lane
//   1. finds its row's exact ellipse column interval (in_ellipse with the
//      reference's float operation order, gaussian_kernel.h:131-137) and
//      the row factor of the kernel;
//   2. pass 1: walks the interval summing K * rho read from the LDS window;
//      the DPP wave sum is the denominator lambda_e (the forward projection);
//   3. pass 2: walks it again adding K * w_e (* rho) into the LDS
//      accumulation window (ds_add_u64 of fixed-point values).
// Modes:
//   MLEM      pass 1 + pass 2 with w_e = 1/lambda_e, values K rho / lambda_e
//             in 2^-40 fixed point (each <= 1)
//   FORWARD   pass 1 only, plus the sigma derivative of lambda_e and the
//             log-likelihood terms, reduced per workgroup into doubles
//   BACKWARD  pass 2 only with upstream weights g_e, values K g_e in signed
//             fixed point with a scale chosen from the largest K g_e
// The inner walks are branch-free where it matters and unrolled by four so
// that each lane keeps several LDS reads and exponentials in flight.  The
// window pitch is odd so that lanes on neighbouring rows of a tilted
// ellipse fall in different LDS banks.  At the end the window is added to
// the global accumulator with one atomic per touched pixel.
//
// Kernel value (GaussianKernel::operator(), gaussian_kernel.h:51-79): the
// exponent splits into a row part and a pixel part,
//   K = c/sqrt(a^T S a) exp(-isdl br^2/2) * exp(-isz pq^2)
//     = exp2(lsc_row - u^2),   u = pq sqrt(isz log2 e),
// with pq = dz - dy tan computed exactly as the reference at the row's first
// ellipse pixel and advanced by one pixel width per column (u = u0 + k du).
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

__device__ __forceinline__ double wave_sum_d(double v) {
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int wave_max(int v) {
  for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return uniform(v);
}

__device__ __forceinline__ float wave_max_f(float v) {
  for (int off = 32; off >= 1; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
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

constexpr float kLog2e = 1.44269504088896341f;

// One box row of an event as a lane walks it.
struct RowSeg {
  int len;        // ellipse pixels in the row (0: nothing to do)
  int w0;         // LDS window index of the first ellipse pixel
  float u0;       // u of the first ellipse pixel
  float lsc;      // log2 of the row factor
  float dz_row;   // row part of d lnK / d sigma_z (FORWARD)
  float ddl_row;  // d lnK / d sigma_dl (FORWARD)
};

template <bool SIGMA>
__device__ __forceinline__ RowSeg row_setup(const StripConst& sg, const StripConst& sk,
                                            const EventGeom& g, int iy, int row_w,
                                            float s_pix) {
  RowSeg s{0, 0, 0.0f, -1e30f, 0.0f, 0.0f};
  float dy = fmaf(-(float)iy, sg.ph, sg.tl_y_half_h) - g.y;
  float Ady2 = (dy * g.A) * dy;
  float Bdy = dy * g.B;
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
  s.len = b - a + 1;
  s.w0 = row_w + a;
  // pq at the first pixel, with the reference's rounding (gaussian_kernel.h:64)
  float dz = fmaf((float)a, sg.pw, sg.tl_z_half_w) - g.z;
  s.u0 = fmaf(-dy, g.t, dz) * s_pix;
  // kernel row factor (gaussian_kernel.h:58-78)
  float sec_sq = g.sec * g.sec;
  float yp = dy + g.y;
  float ap = -(yp - sk.R) * sec_sq;
  float aq = -(sk.R + yp) * sec_sq;
  float ar = ((yp * -2.0f) * g.sec) * g.t;
  float br = (dy * -2.0f) * g.sec;
  float norm = fmaf(ar * sk.isdl, ar, fmaf(ap * sk.isz, ap, (aq * sk.isz) * aq));
  float brb = (br * sk.isdl) * br;
  s.lsc = log2f(sk.c) - 0.5f * log2f(norm) - (0.5f * kLog2e) * brb;
  if constexpr (SIGMA) {
    // ln K = ln c - ln(norm)/2 - bSb/2 with c ~ 1/(sz^2 sdl), norm and bSb
    // linear in 1/sz^2 and 1/sdl^2:
    //   d lnK/d sz  = (-2 + isz (ap^2 + aq^2)/norm + 2 isz pq^2) / sz
    //   d lnK/d sdl = (-1 + isdl ar^2/norm + isdl br^2) / sdl
    float apq = fmaf(ap, ap, aq * aq);
    s.dz_row = (-2.0f + sk.isz * apq / norm) / sk.sigma_z;
    s.ddl_row = (-1.0f + sk.isdl * (ar * ar) / norm + brb) / sk.sigma_dl;
  }
  return s;
}

constexpr int kUnroll = 4;
constexpr int kCachedGroups = 4;  // row groups kept in registers between passes

// x in [0, 2^63): the two 32-bit halves of round-toward-zero(x), x given as
// w = x * 2^-32 (five VALU operations instead of a generic float->i64).
__device__ __forceinline__ unsigned long long split_u64(float w) {
  float hi = floorf(w);
  float lo = (w - hi) * 4294967296.0f;
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
}

}  // namespace

// One workgroup per work item (a chunk of events sharing a spatial bin).
template <int MODE, int WAVES, bool COMP>
__global__ __launch_bounds__(64 * WAVES) void k_strip_fused(FusedArgs a) {
  constexpr bool kPass1 = MODE != MODE_BACKWARD;
  constexpr bool kPass2 = MODE != MODE_FORWARD;
  constexpr bool kRho = MODE != MODE_BACKWARD;  // rho window in LDS
  constexpr bool kSigma = MODE == MODE_FORWARD;
  constexpr int kThreads = 64 * WAVES;
  using U = unsigned long long;
  extern __shared__ __align__(16) unsigned char smem[];
  const StripConst& sg = a.sg;
  const StripConst& sk = a.sk;
  const WorkItem it = a.items[blockIdx.x];
  const int ww = it.ww, wh = it.wh;
  const int pitch = ww | 1;  // odd pitch: rows of a tilted ellipse spread over banks
  const int npx = pitch * wh;
  // the rho window has spare space: unrolled lanes past their row's end read
  // up to maxlen + 3 elements beyond it (values discarded)
  const int nrho = npx + 2 * pitch + 8;
  float* rho_w = reinterpret_cast<float*>(smem);
  U* acc_w = reinterpret_cast<U*>(
      smem + (kRho ? (((size_t)nrho * 4 + 15) & ~(size_t)15) : 0));
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = __lane_id();

  if constexpr (kRho) {
    for (int i = threadIdx.x; i < nrho; i += kThreads) {
      int r = i / pitch, c = i - r * pitch;
      float v = 0.0f;
      if (r < wh && c < ww) v = a.rho[(size_t)(it.wy0 + r) * sg.W + (it.wx0 + c)];
      rho_w[i] = v;
    }
  }
  if constexpr (kPass2) {
    for (int i = threadIdx.x; i < npx; i += kThreads) acc_w[i] = 0ull;
  }
  __syncthreads();

  const float s_pix = sqrtf(sk.isz * kLog2e);  // sqrt(isz log2 e)
  const float du = sg.pw * s_pix;               // u step per column
  const float cz_pix = 2.0f / (kLog2e * sk.sigma_z);  // 2 isz pq^2 / sz = cz u^2
  double ll = 0.0, gsz = 0.0, gsdl = 0.0, mmax = 0.0;

  for (int e = it.begin + wave; e < it.end; e += WAVES) {
    const EventGeom g = a.sgeom[e];
    const int nrows = g.y1 - g.y0;
    const int ngroups = (nrows + 63) >> 6;
    const int row0 = (g.y0 - it.wy0) * pitch - it.wx0;  // window index of (g.y0, 0)
    const int orig = a.sidx[e];
    RowSeg cache[kCachedGroups];

    // ---- pass 1: denominator (and sigma sums) -------------------------------
    float lam = 0.0f;
    if constexpr (kPass1) {
      float acc = 0.0f, comp = 0.0f, sz = 0.0f, sdl = 0.0f, lmax = -1e30f;
#pragma unroll
      for (int gi = 0; gi < kCachedGroups + 1; ++gi) {
        // groups beyond the cached ones share the last iteration as a loop
        for (int gg = gi; gg < (gi < kCachedGroups ? gi + 1 : ngroups); ++gg) {
          if (gg >= ngroups) break;
          const int r = gg * 64 + lane;
          RowSeg s{0, 0, 0.0f, -1e30f, 0.0f, 0.0f};
          if (r < nrows) s = row_setup<kSigma>(sg, sk, g, g.y0 + r, row0 + r * pitch, s_pix);
          if (gi < kCachedGroups) cache[gi] = s;
          const int maxlen = wave_max(s.len);
          float s0 = 0.0f, s1 = 0.0f;
          for (int k = 0; k < maxlen; k += kUnroll) {
            const float kf = (float)k;
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
              const float uu = fmaf(kf + (float)u, du, s.u0);
              const float kv = __builtin_amdgcn_exp2f(fmaf(-uu, uu, s.lsc));
              float kr = kv * rho_w[s.w0 + k + u];
              kr = (k + u < s.len) ? kr : 0.0f;
              if constexpr (COMP) {
                float sum, err;
                two_sum(s0, kr, sum, err);
                s0 = sum;
                comp += err;
              } else {
                s0 += kr;
              }
              if constexpr (kSigma) s1 = fmaf(kr, uu * uu, s1);
            }
          }
          if constexpr (COMP) {
            float sum, err;
            two_sum(acc, s0, sum, err);
            acc = sum;
            comp += err;
          } else {
            acc += s0;
          }
          if constexpr (kSigma) {
            sz += fmaf(s.dz_row, s0, cz_pix * s1);
            sdl += s.ddl_row * s0;
            lmax = fmaxf(lmax, s.lsc);
          }
        }
      }
      lam = COMP ? wave_sum_comp(acc, comp) : wave_sum(acc);
      if (lane == 0) a.lambda[orig] = lam;
      if constexpr (kSigma) {
        // weight of this event: upstream g_e or 1/lambda_e (log-likelihood)
        float ge = a.g ? a.g[orig] : (lam > 0.0f ? 1.0f / lam : 0.0f);
        double wsz = wave_sum_d((double)sz), wsdl = wave_sum_d((double)sdl);
        float kmax = __builtin_amdgcn_exp2f(wave_max_f(lmax));
        if (lane == 0) {
          if (lam > 0.0f) ll += log((double)lam);
          gsz += (double)ge * wsz;
          gsdl += (double)ge * wsdl;
          mmax = fmax(mmax, (double)fabsf(ge) * (double)kmax);
        }
      }
    }

    // ---- pass 2: backprojection into the LDS window --------------------------
    if constexpr (kPass2) {
      // value multiplier, including the fixed-point unit / 2^32 (split_u64)
      float we;
      if constexpr (MODE == MODE_MLEM) {
        if (!(lam > 0.0f)) continue;
        we = 256.0f / lam;  // 2^40 / 2^32
      } else {
        float ge = a.g ? a.g[orig] : a.lambda[orig];
        if (!a.g) ge = ge > 0.0f ? 1.0f / ge : 0.0f;
        if (ge == 0.0f) continue;
        we = ge * a.fixed_scale * 2.3283064365386963e-10f;  // / 2^32
      }
      const bool neg = we < 0.0f;  // uniform: one sign per event
      const float wabs = fabsf(we);
#pragma unroll
      for (int gi = 0; gi < kCachedGroups + 1; ++gi) {
        for (int gg = gi; gg < (gi < kCachedGroups ? gi + 1 : ngroups); ++gg) {
          if (gg >= ngroups) break;
          RowSeg s;
          if (kPass1 && gi < kCachedGroups) {
            s = cache[gi];
          } else {
            const int r = gg * 64 + lane;
            s = RowSeg{0, 0, 0.0f, -1e30f, 0.0f, 0.0f};
            if (r < nrows) s = row_setup<false>(sg, sk, g, g.y0 + r, row0 + r * pitch, s_pix);
          }
          const int maxlen = wave_max(s.len);
          for (int k = 0; k < maxlen; k += kUnroll) {
            const float kf = (float)k;
            U v[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
              const float uu = fmaf(kf + (float)u, du, s.u0);
              float w = __builtin_amdgcn_exp2f(fmaf(-uu, uu, s.lsc)) * wabs;
              if constexpr (kRho) w *= rho_w[s.w0 + k + u];
              v[u] = split_u64(w);
              if (neg) v[u] = 0ull - v[u];  // two's complement: signed sums stay exact
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u)
              if (k + u < s.len) atomicAdd(&acc_w[s.w0 + k + u], v[u]);
          }
        }
      }
    }
  }

  if constexpr (kSigma) {
    // lane 0 of each wave holds the wave's scalars
    if (lane == 0) {
      unsafeAtomicAdd(&a.scalars[0], ll);
      unsafeAtomicAdd(&a.scalars[1], gsz);
      unsafeAtomicAdd(&a.scalars[2], gsdl);
      // non-negative doubles order like their bit patterns
      atomicMax(reinterpret_cast<U*>(&a.scalars[3]), (U)__double_as_longlong(mmax));
    }
  }

  if constexpr (kPass2) {
    __syncthreads();
    for (int i = threadIdx.x; i < ww * wh; i += kThreads) {
      int r = i / ww, c = i - r * ww;
      U v = acc_w[r * pitch + c];
      if (v == 0ull) continue;
      size_t p = (size_t)(it.wy0 + r) * sg.W + (it.wx0 + c);
      if (a.acc_kind == GLOBAL_U64)
        atomicAdd(reinterpret_cast<U*>(a.acc) + p, v);
      else if (a.acc_kind == GLOBAL_F64)
        unsafeAtomicAdd(reinterpret_cast<double*>(a.acc) + p,
                        (double)(long long)v * a.inv_fixed_scale);
      else
        unsafeAtomicAdd(reinterpret_cast<float*>(a.acc) + p,
                        (float)((double)(long long)v * a.inv_fixed_scale));
    }
  }
}

// ---------------------------------------------------------------------------

size_t fused_lds_bytes(int max_window_px, int max_pitch, int mode) {
  size_t npx = (size_t)max_window_px;
  size_t rho = mode == MODE_BACKWARD
                   ? 0
                   : (((npx + 2 * (size_t)max_pitch + 8) * 4 + 15) & ~(size_t)15);
  size_t acc = mode == MODE_FORWARD ? 0 : npx * 8;
  return rho + acc + 16;
}

namespace {

template <int MODE, int WAVES, bool COMP>
hipError_t launch_t(hipStream_t st, const FusedArgs& a, int n_items, size_t lds) {
  auto kern = k_strip_fused<MODE, WAVES, COMP>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(n_items), dim3(64 * WAVES), lds, st, a);
  return hipGetLastError();
}

template <int MODE, bool COMP>
hipError_t launch_w(int waves, hipStream_t st, const FusedArgs& a, int n_items, size_t lds) {
  switch (waves) {
    case 16: return launch_t<MODE, 16, COMP>(st, a, n_items, lds);
    case 8: return launch_t<MODE, 8, COMP>(st, a, n_items, lds);
    default: return launch_t<MODE, 4, COMP>(st, a, n_items, lds);
  }
}

}  // namespace

hipError_t launch_strip_fused(hipStream_t st, int mode, int waves, int compensated,
                              const FusedArgs& a, int n_items, size_t lds_bytes) {
  if (n_items <= 0) return hipSuccess;
  switch (mode) {
    case MODE_MLEM:
      return compensated ? launch_w<MODE_MLEM, true>(waves, st, a, n_items, lds_bytes)
                         : launch_w<MODE_MLEM, false>(waves, st, a, n_items, lds_bytes);
    case MODE_FORWARD:
      return compensated ? launch_w<MODE_FORWARD, true>(waves, st, a, n_items, lds_bytes)
                         : launch_w<MODE_FORWARD, false>(waves, st, a, n_items, lds_bytes);
    default:
      return launch_w<MODE_BACKWARD, false>(waves, st, a, n_items, lds_bytes);
  }
}

}  // namespace petstrip
