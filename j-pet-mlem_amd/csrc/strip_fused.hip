// Fused MLEM iteration for gfx950: events sorted into spatial bins, one
// workgroup per chunk of a bin, rho and the backprojection held in LDS.
//
// Reference semantics: Reconstruction::bb_update
// (src/2d/strip/reconstruction.h:220-298) per event, Jacobi over events
// (reconstruction.h:121-149).  For every event a wave
//   1. builds a row table: for each box row the exact column interval of
//      the 3-sigma ellipse (in_ellipse with the reference's float operation
//      order, gaussian_kernel.h:131-137) and the row's kernel factor;
//   2. pass 1: walks the ellipse pixels (each lane a contiguous run of the
//      row-major pixel list) summing K * rho from the LDS window; the wave
//      sum (DPP) is the denominator lambda_e;
//   3. pass 2: walks the same pixels again adding K * rho / lambda_e into the
//      LDS accumulation window (ds_add), kernel values recomputed.
// At the end the window is added to the global image with one atomic per
// touched pixel.
//
// The kernel value follows GaussianKernel::operator()
// (gaussian_kernel.h:51-79) with its exponent split in a row part and a
// pixel part: K = [c / sqrt(a^T S a) * exp(-isdl br^2 / 2)] * exp(-isz pq^2),
// the pixel part evaluated as exp2(-(pq * sqrt(isz log2 e))^2).
#include <hip/hip_runtime.h>

#include "strip_fused.h"
#include "strip_geometry.h"

namespace petstrip {

namespace {

constexpr int kWaves = 4;            // waves per workgroup
constexpr int kThreads = 64 * kWaves;

struct RowInfo {
  int32_t start;  // exclusive prefix of ellipse pixels before this row
  int16_t lo;     // first ellipse column (absolute)
  int16_t len;    // ellipse pixels in the row
  float ndy;      // -(pixel_center.y - y_e)
  float scale;    // c / sqrt(a^T S a) * exp(-isdl br^2 / 2)
};

__device__ __forceinline__ int uniform(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

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

// Compensated wave sum of (hi, lo) pairs: two-sum at every step.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ float wave_sum_comp(float hi, float lo) {
  for (int off = 1; off < 64; off <<= 1) {
    float oh = __shfl_xor(hi, off, 64);
    float ol = __shfl_xor(lo, off, 64);
    // symmetric order so both partners compute the identical result
    float a = (__lane_id() & off) ? oh : hi;
    float b = (__lane_id() & off) ? hi : oh;
    float la = (__lane_id() & off) ? ol : lo;
    float lb = (__lane_id() & off) ? lo : ol;
    float s, e;
    two_sum(a, b, s, e);
    hi = s;
    lo = e + (la + lb);
  }
  return hi + lo;
}

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = __lane_id();
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  return v;
}

// Exact reference membership of column iz in the row with (dy, Ady2, Bdy).
__device__ __forceinline__ bool exact_in(const StripConst& s, const EventGeom& g,
                                         float Ady2, float Bdy, int iz) {
  float dz = fmaf((float)iz, s.pw, s.tl_z_half_w) - g.z;
  float q = fmaf(g.C * dz, dz, fmaf(Bdy, dz, Ady2));
  return q <= 9.0f;
}

// Row setup for one box row iy: the exact ellipse column interval and the
// kernel's row factor.
__device__ __forceinline__ void row_setup(const StripConst& sg, const StripConst& sk,
                                          const EventGeom& g, int iy, int& lo,
                                          int& len, float& ndy, float& scale) {
  float dy = fmaf(-(float)iy, sg.ph, sg.tl_y_half_h) - g.y;
  float Ady2 = (dy * g.A) * dy;
  float Bdy = dy * g.B;
  ndy = -dy;
  len = 0;
  lo = g.x0;
  // Superset of the interval from the real roots of C dz^2 + Bdy dz + Ady2 = 9.05,
  // then exact tests move both ends inward.  Along a row the exact test has
  // no holes: adjacent columns differ in q by at least C pw^2 (checked on the
  // host), far above the rounding of q.
  float disc = Bdy * Bdy - 4.0f * g.C * (Ady2 - 9.05f);
  if (!(disc >= 0.0f)) return;
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
  if (a > b) return;
  lo = a;
  len = b - a + 1;
  // kernel row factor (gaussian_kernel.h:58-78)
  float sec_sq = g.sec * g.sec;
  float yp = dy + g.y;
  float ap = -(yp - sk.R) * sec_sq;
  float aq = -(sk.R + yp) * sec_sq;
  float ar = ((yp * -2.0f) * g.sec) * g.t;
  float br = (dy * -2.0f) * g.sec;
  float norm = fmaf(ar * sk.isdl, ar, fmaf(ap * sk.isz, ap, (aq * sk.isz) * aq));
  scale = (sk.c / sqrtf(norm)) * expf(-0.5f * ((br * sk.isdl) * br));
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

}  // namespace

// One workgroup per work item (a chunk of events sharing a spatial bin).
template <int ACC>
__global__ __launch_bounds__(kThreads) void k_mlem_fused(
    StripConst sg, StripConst sk, const EventGeom* __restrict__ sgeom,
    const int32_t* __restrict__ sidx, const WorkItem* __restrict__ items,
    const float* __restrict__ rho, float* __restrict__ lambda,
    void* __restrict__ acc_global, int acc_global_kind, int max_rows,
    int compensated) {
  using A = typename AccType<ACC>::T;
  extern __shared__ __align__(16) unsigned char smem[];
  const WorkItem it = items[blockIdx.x];
  const int ww = it.ww, wh = it.wh, npx = ww * wh;
  float* rho_w = reinterpret_cast<float*>(smem);
  size_t acc_off = ((size_t)npx * 4 + 15) & ~(size_t)15;
  A* acc_w = reinterpret_cast<A*>(smem + acc_off);
  size_t rows_off = (acc_off + (size_t)npx * sizeof(A) + 15) & ~(size_t)15;
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = __lane_id();
  RowInfo* rows = reinterpret_cast<RowInfo*>(smem + rows_off) + wave * max_rows;
  int2* lane_start =
      reinterpret_cast<int2*>(smem + rows_off + (size_t)kWaves * max_rows * sizeof(RowInfo)) +
      wave * 64;

  // load the rho window, clear the accumulation window
  for (int i = threadIdx.x; i < npx; i += kThreads) {
    int r = i / ww, c = i - r * ww;
    rho_w[i] = rho[(size_t)(it.wy0 + r) * sg.W + (it.wx0 + c)];
    acc_w[i] = (A)0;
  }
  __syncthreads();

  const float s_pix = sqrtf(sk.isz * 1.44269504088896341f);  // sqrt(isz log2 e)

  for (int e = it.begin + wave; e < it.end; e += kWaves) {
    const EventGeom g = sgeom[e];
    const int nrows = g.y1 - g.y0;
    // ---- row table -------------------------------------------------------
    int total = 0;
    for (int base = 0; base < nrows; base += 64) {
      int r = base + lane;
      int lo = 0, len = 0;
      float ndy = 0.0f, scale = 0.0f;
      if (r < nrows) row_setup(sg, sk, g, g.y0 + r, lo, len, ndy, scale);
      int incl = wave_incl_scan(len);
      if (r < nrows) {
        RowInfo ri;
        ri.start = total + incl - len;
        ri.lo = (int16_t)lo;
        ri.len = (int16_t)len;
        ri.ndy = ndy;
        ri.scale = scale;
        rows[r] = ri;
      }
      total += __shfl(incl, 63, 64);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // each lane walks the run [lane*T, min(lane*T+T, total)) of the
    // row-major ellipse pixel list; find the row holding each run start
    const int T = (total + 63) >> 6;
    if (T > 0) {
      for (int r = lane; r < nrows; r += 64) {
        RowInfo ri = rows[r];
        if (ri.len > 0) {
          int l0 = (ri.start + T - 1) / T;
          int l1 = (ri.start + ri.len - 1) / T;
          for (int l = l0; l <= l1 && l < 64; ++l)
            lane_start[l] = make_int2(r, l * T - ri.start);
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int f0 = lane * T;
    const int cnt = max(0, min(T, total - f0));
    int2 st = make_int2(0, 0);
    if (cnt > 0) st = lane_start[lane];

    // ---- pass 1: denominator -------------------------------------------------
    float acc = 0.0f, comp = 0.0f;
    {
      int r = st.x, k = st.y;
      RowInfo ri = rows[min(r, max(nrows - 1, 0))];
      for (int j = 0; j < cnt; ++j) {
        while (k >= ri.len) {
          ++r;
          k = 0;
          ri = rows[r];
        }
        const int iz = ri.lo + k;
        const float dz = fmaf((float)iz, sg.pw, sg.tl_z_half_w) - g.z;
        const float u = fmaf(ri.ndy, g.t, dz) * s_pix;
        const float K = __builtin_amdgcn_exp2f(-(u * u)) * ri.scale;
        const float rv = rho_w[(g.y0 + r - it.wy0) * ww + (iz - it.wx0)];
        const float kr = K * rv;
        if (compensated) {
          float s, err;
          two_sum(acc, kr, s, err);
          acc = s;
          comp += err;
        } else {
          acc += kr;
        }
        ++k;
      }
    }
    const float lam = compensated ? wave_sum_comp(acc, comp) : wave_sum(acc);
    if (lane == 0) lambda[sidx[e]] = lam;
    if (!(lam > 0.0f)) continue;
    const float inv = 1.0f / lam;

    // ---- pass 2: backprojection into the LDS window --------------------------
    {
      int r = st.x, k = st.y;
      RowInfo ri = rows[min(r, max(nrows - 1, 0))];
      float sc = ri.scale * inv;
      for (int j = 0; j < cnt; ++j) {
        while (k >= ri.len) {
          ++r;
          k = 0;
          ri = rows[r];
          sc = ri.scale * inv;
        }
        const int iz = ri.lo + k;
        const float dz = fmaf((float)iz, sg.pw, sg.tl_z_half_w) - g.z;
        const float u = fmaf(ri.ndy, g.t, dz) * s_pix;
        const int w = (g.y0 + r - it.wy0) * ww + (iz - it.wx0);
        const float v = __builtin_amdgcn_exp2f(-(u * u)) * sc * rho_w[w];
        if constexpr (ACC == ACC_LDS_F32) {
          atomicAdd(&acc_w[w], v);
        } else {
          atomicAdd(&acc_w[w], (unsigned long long)__float2ull_rn(v * kFixedScale));
        }
        ++k;
      }
    }
  }
  __syncthreads();

  // ---- flush the window into the global accumulator ------------------------
  for (int i = threadIdx.x; i < npx; i += kThreads) {
    A v = acc_w[i];
    if (v == (A)0) continue;
    int r = i / ww, c = i - r * ww;
    size_t p = (size_t)(it.wy0 + r) * sg.W + (it.wx0 + c);
    if constexpr (ACC == ACC_LDS_F32) {
      if (acc_global_kind == GLOBAL_F64)
        unsafeAtomicAdd(reinterpret_cast<double*>(acc_global) + p, (double)v);
      else
        unsafeAtomicAdd(reinterpret_cast<float*>(acc_global) + p, v);
    } else {
      atomicAdd(reinterpret_cast<unsigned long long*>(acc_global) + p, v);
    }
  }
}

// ---------------------------------------------------------------------------
// Setup kernels: spatial bins and work items.

// Bin of an event: the bin of its box centre; empty boxes get -1.
__global__ void k_bin_count(const EventGeom* __restrict__ geom, int64_t n,
                            int bin_w, int bin_h, int nbx, int32_t* __restrict__ bin,
                            int32_t* __restrict__ count) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  EventGeom g = geom[e];
  int b = -1;
  if (g.x1 > g.x0 && g.y1 > g.y0) {
    int cx = (g.x0 + g.x1) >> 1, cy = (g.y0 + g.y1) >> 1;
    b = (cy / bin_h) * nbx + (cx / bin_w);
    atomicAdd(&count[b], 1);
  }
  bin[e] = b;
}

__global__ void k_bin_scatter(const EventGeom* __restrict__ geom, int64_t n,
                              const int32_t* __restrict__ bin,
                              int32_t* __restrict__ cursor,
                              EventGeom* __restrict__ sgeom,
                              int32_t* __restrict__ sidx) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int b = bin[e];
  if (b < 0) return;
  int pos = atomicAdd(&cursor[b], 1);
  sgeom[pos] = geom[e];
  sidx[pos] = (int32_t)e;
}

// Window of each work item: the union of its events' boxes.
__global__ __launch_bounds__(256) void k_item_windows(const EventGeom* __restrict__ sgeom,
                                                      WorkItem* __restrict__ items) {
  __shared__ int red[4][5];
  WorkItem it = items[blockIdx.x];
  int x0 = 1 << 30, y0 = 1 << 30, x1 = -(1 << 30), y1 = -(1 << 30), mr = 0;
  for (int e = it.begin + threadIdx.x; e < it.end; e += blockDim.x) {
    EventGeom g = sgeom[e];
    x0 = min(x0, g.x0);
    y0 = min(y0, g.y0);
    x1 = max(x1, g.x1);
    y1 = max(y1, g.y1);
    mr = max(mr, g.y1 - g.y0);
  }
  for (int off = 32; off >= 1; off >>= 1) {
    x0 = min(x0, __shfl_xor(x0, off, 64));
    y0 = min(y0, __shfl_xor(y0, off, 64));
    x1 = max(x1, __shfl_xor(x1, off, 64));
    y1 = max(y1, __shfl_xor(y1, off, 64));
    mr = max(mr, __shfl_xor(mr, off, 64));
  }
  int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w][0] = x0;
    red[w][1] = y0;
    red[w][2] = x1;
    red[w][3] = y1;
    red[w][4] = mr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
      x0 = min(x0, red[i][0]);
      y0 = min(y0, red[i][1]);
      x1 = max(x1, red[i][2]);
      y1 = max(y1, red[i][3]);
      mr = max(mr, red[i][4]);
    }
    it.wx0 = x0;
    it.wy0 = y0;
    it.ww = x1 - x0;
    it.wh = y1 - y0;
    it.max_rows = mr;
    items[blockIdx.x] = it;
  }
}

// rho' from the accumulator; the accumulator is cleared for the next use.
__global__ void k_finalize(void* __restrict__ acc, int kind, float* __restrict__ out,
                           int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (kind == GLOBAL_F32) {
    float* a = reinterpret_cast<float*>(acc);
    out[i] = a[i];
    a[i] = 0.0f;
  } else if (kind == GLOBAL_F64) {
    double* a = reinterpret_cast<double*>(acc);
    out[i] = (float)a[i];
    a[i] = 0.0;
  } else {
    unsigned long long* a = reinterpret_cast<unsigned long long*>(acc);
    out[i] = (float)((double)a[i] * (1.0 / 1099511627776.0));
    a[i] = 0ull;
  }
}

// ---------------------------------------------------------------------------

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_bin_count(hipStream_t st, const EventGeom* geom, int64_t n,
                            int bin_w, int bin_h, int nbx, int32_t* bin,
                            int32_t* count) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin_count, dim3(nblk(n, 256)), dim3(256), 0, st, geom, n,
                     bin_w, bin_h, nbx, bin, count);
  return hipGetLastError();
}

hipError_t launch_bin_scatter(hipStream_t st, const EventGeom* geom, int64_t n,
                              const int32_t* bin, int32_t* cursor,
                              EventGeom* sgeom, int32_t* sidx) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin_scatter, dim3(nblk(n, 256)), dim3(256), 0, st, geom, n,
                     bin, cursor, sgeom, sidx);
  return hipGetLastError();
}

hipError_t launch_item_windows(hipStream_t st, const EventGeom* sgeom,
                               WorkItem* items, int n_items) {
  if (n_items <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_item_windows, dim3(n_items), dim3(256), 0, st, sgeom, items);
  return hipGetLastError();
}

size_t fused_lds_bytes(int max_window_px, int acc_mode, int max_rows) {
  size_t npx = (size_t)max_window_px;
  size_t acc_off = (npx * 4 + 15) & ~(size_t)15;
  size_t accb = acc_mode == ACC_LDS_F32 ? 4 : 8;
  size_t rows_off = (acc_off + npx * accb + 15) & ~(size_t)15;
  return rows_off + (size_t)kWaves * max_rows * sizeof(RowInfo) + kWaves * 64 * 8;
}

hipError_t launch_mlem_fused(hipStream_t st, const StripConst& sg, const StripConst& sk,
                             const EventGeom* sgeom, const int32_t* sidx,
                             const WorkItem* items, int n_items, size_t lds_bytes,
                             const float* rho, float* lambda, void* acc_global,
                             int acc_global_kind, int acc_mode, int max_rows,
                             int compensated) {
  if (n_items <= 0) return hipSuccess;
  if (acc_mode == ACC_LDS_F32) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_mlem_fused<ACC_LDS_F32>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlem_fused<ACC_LDS_F32>, dim3(n_items), dim3(kThreads),
                       lds_bytes, st, sg, sk, sgeom, sidx, items, rho, lambda,
                       acc_global, acc_global_kind, max_rows, compensated);
  } else {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&k_mlem_fused<ACC_LDS_FIXED64>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mlem_fused<ACC_LDS_FIXED64>, dim3(n_items), dim3(kThreads),
                       lds_bytes, st, sg, sk, sgeom, sidx, items, rho, lambda,
                       acc_global, acc_global_kind, max_rows, compensated);
  }
  return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t st, void* acc, int kind, float* out,
                           int64_t n) {
  hipLaunchKernelGGL(k_finalize, dim3(nblk(n, 256)), dim3(256), 0, st, acc, kind,
                     out, n);
  return hipGetLastError();
}

}  // namespace petstrip
