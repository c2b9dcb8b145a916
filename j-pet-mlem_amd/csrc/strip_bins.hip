// Event binning for the fused MLEM iteration: each event goes to the
// kBin x kBin pixel bin of its box centre, bins are cut into chunks, and
// each chunk gets the image window covering all its boxes.  Also the
// accumulator finalisation (rho' from the accumulator).
#include <hip/hip_runtime.h>

#include "strip_fused.h"
#include "strip_geometry.h"

namespace petstrip {


// Bin of an event: its box height class (taller than 32 rows, 17-32, at
// most 16) and width class (at most wide_cols columns, wider), then the bin
// of its box centre; empty boxes get -1.
__global__ void k_bin_count(const EventGeom* __restrict__ geom, int64_t n,
                            int bin_w, int bin_h, int nbx, int n_spatial, int wide_cols,
                            int32_t* __restrict__ bin, int32_t* __restrict__ count) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  EventGeom g = geom[e];
  int b = -1;
  if (g.x1 > g.x0 && g.y1 > g.y0) {
    int cx = (g.x0 + g.x1) >> 1, cy = (g.y0 + g.y1) >> 1;
    const int rows = g.y1 - g.y0;
    const int cls = rows <= 16 ? 2 : rows <= 32 ? 1 : 0;
    const int wcls = g.x1 - g.x0 > wide_cols ? 1 : 0;
    b = (cls * kWidthClasses + wcls) * n_spatial + (cy / bin_h) * nbx + (cx / bin_w);
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

// rho' from the accumulator (times mul[i], the image, for sums of K /
// lambda_e); the accumulator is cleared for the next use, and so is the
// counter *clear when given.
__global__ void k_finalize(void* __restrict__ acc, int kind, double unit, float bias,
                           int check_sign, int* __restrict__ overflow,
                           float* __restrict__ out, int64_t n, const float* __restrict__ mul,
                           unsigned long long* __restrict__ clear) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && clear) *clear = 0ull;
  if (i >= n) return;
  const double u = mul ? unit * (double)mul[i] : unit;
  if (kind == GLOBAL_F32) {
    float* a = reinterpret_cast<float*>(acc);
    out[i] = (float)((double)a[i] * u) + bias;
    a[i] = 0.0f;
  } else if (kind == GLOBAL_F64) {
    double* a = reinterpret_cast<double*>(acc);
    out[i] = (float)(a[i] * u) + bias;
    a[i] = 0.0;
  } else {
    long long* a = reinterpret_cast<long long*>(acc);
    long long v = a[i];
    if (check_sign && v < 0 && overflow) *overflow = 1;
    out[i] = (float)((double)v * u) + bias;
    a[i] = 0;
  }
}

// ---------------------------------------------------------------------------

static unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

hipError_t launch_bin_count(hipStream_t st, const EventGeom* geom, int64_t n,
                            int bin_w, int bin_h, int nbx, int n_spatial, int wide_cols,
                            int32_t* bin, int32_t* count) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bin_count, dim3(nblk(n, 256)), dim3(256), 0, st, geom, n,
                     bin_w, bin_h, nbx, n_spatial, wide_cols, bin, count);
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

hipError_t launch_finalize(hipStream_t st, void* acc, int kind, double unit,
                           float bias, int check_sign, int* overflow, float* out,
                           int64_t n, const float* mul, unsigned long long* clear) {
  hipLaunchKernelGGL(k_finalize, dim3(nblk(n, 256)), dim3(256), 0, st, acc, kind,
                     unit, bias, check_sign, overflow, out, n, mul, clear);
  return hipGetLastError();
}

}  // namespace petstrip
