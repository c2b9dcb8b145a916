// 3D hybrid reconstruction (include/petstrip_hybrid.h): the strip kernel
// walked per LOR of a barrel scanner over a voxel grid.
//
// Restates PET3D::Hybrid::Reconstruction (grey277/j-pet-mlem
// src/3d/hybrid/reconstruction.h) for an AMD Instinct GPU:
//   - host: LOR segments (2d/geometry/line_segment.h:21-31), LOR-pixel
//     weights, sensitivity and their normalisation (reconstruction.h:212-265)
//     with the reference's float expressions, so they equal its values;
//   - k_hybrid_frames: translate_to_frame and the grid filter
//     (reconstruction.h:97-161), one thread per response, then an
//     order-keeping compaction (wave ballots, block offsets scanned on the
//     host);
//   - k_hybrid_mlem: the iteration (reconstruction.h:267-445), one wave per
//     event: its lanes first put the per-pixel terms of the event's pixel
//     infos into an LDS table, then walk the event's (pixel info, plane)
//     pairs, pixel-major as the reference's loops, with one exponential and
//     one arctangent per voxel, keep each voxel's kernel x rho in LDS (the
//     reference's voxel cache) while they sum the denominator in a fixed
//     shuffle tree, and add weight / denominator to the next image with one
//     float atomic per voxel (as the reference sums float thread images);
//     k_hybrid_finalize makes it the image.  Image and accumulator are
//     plane-fastest on the device.
// The reference's backprojection reads each voxel's weight from a
// voxel-indexed cache, so when an event's range holds one pixel twice (a
// LOR's pixel infos from a matrix with TOF positions repeat pixels, and the
// extra pixel info one past the LOR's end, reconstruction.h:133-138, can
// repeat a pixel of the range) every entry adds the last entry's weight;
// the walk here does the same through a per-pixel-info link to the next
// entry of the same pixel in its LOR.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/petstrip_hybrid.h"

namespace petstrip {
int set_error(int code, const char* msg);
}

namespace {

using petstrip::set_error;

constexpr int kBlock = 256;              // 4 waves
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int64_t kMaxInfos = (int64_t)1 << 30;

// A LOR's segment quantities in the reference's floats
// (line_segment.h:21-31): mid point, unit direction, normal, length, the
// normal's offset; R = length / 2 is the strip frame's radius.
struct SegD {
  float mx, my, dx, dy, nx, ny, len, dfo, R;
};

// A frame event on the device.  end is the reference's (at most one past the
// LOR's last pixel info); extra is that one-past index, or -1.
struct HEvent {
  int32_t lor;
  int32_t begin, end, extra;
  float up, right, tan, sec;
  int16_t first, second, plane_begin, plane_end;
};

struct HGrid {
  int32_t w, h, planes;
  float ps, llcx, llcy, z_left;
};

struct HKernel {
  float isz, isdl, c, length;
};

SegD make_seg(const float* s) {
  SegD g;
  g.mx = (s[0] + s[2]) / 2.0f;
  g.my = (s[1] + s[3]) / 2.0f;
  const float vx = s[2] - s[0], vy = s[3] - s[1];
  const float l = std::sqrt(vx * vx + vy * vy);
  g.dx = vx / l;
  g.dy = vy / l;
  g.nx = -g.dy;
  g.ny = g.dx;
  g.len = l;
  g.dfo = s[0] * g.nx + s[1] * g.ny;
  g.R = l / 2;
  return g;
}

// GaussianKernel constants (2d/strip/gaussian_kernel.h:27-33)
HKernel make_kernel(double sigma_z, double sigma_dl, double length) {
  HKernel k;
  const float sz = (float)sigma_z, sdl = (float)sigma_dl;
  k.isz = 1 / (sz * sz);
  k.isdl = 1 / (sdl * sdl);
  k.c = (float)(1 / (2 * M_PI * M_PI) * (double)std::sqrt(k.isz * k.isz * k.isdl));
  k.length = (float)length;
  return k;
}

// ---- device ------------------------------------------------------------

// Strip kernel at (dz, dy) from the event (gaussian_kernel.h:51-79) over the
// frame sensitivity at (z, y) (gaussian_kernel.h:93-105,140-149), with the
// host's std::min / std::max.
__device__ __forceinline__ float kernel2d(const HKernel& k, float up, float t, float sec,
                                          float R, float z, float y, float right) {
  const float dz = z - right, dy = y - up;
  const float sec_sq = sec * sec;
  const float yp = dy + up;
  const float ap = -(yp - R) * sec_sq, aq = -(yp + R) * sec_sq;
  const float ar = -2 * yp * sec * t;
  const float pq = dz - dy * t;
  const float br = -2 * dy * sec;
  const float a_a = ap * k.isz * ap + aq * k.isz * aq + ar * k.isdl * ar;
  const float b_b = pq * k.isz * pq + pq * k.isz * pq + br * k.isdl * br;
  const float kv = (k.c / sqrtf(a_a)) * expf(-0.5f * b_b);
  const float L2 = k.length / 2;
  const float a = (L2 - z) / (R - y), b = (L2 + z) / (R + y);
  const float c = -(L2 + z) / (R - y), d = (-L2 + z) / (R + y);
  const float sens = (atanf(b < a ? b : a) - atanf(c < d ? d : c)) / (float)M_PI;
  return kv / sens;
}

__global__ __launch_bounds__(kBlock) void k_hybrid_frames(
    const SegD* __restrict__ seg, const int64_t* __restrict__ lor_begin,
    const int64_t* __restrict__ lor_end, const float* __restrict__ positions,
    const int32_t* __restrict__ pixels, const float* __restrict__ weights, int32_t n_infos,
    HGrid grid, HKernel k, const int16_t* __restrict__ lors, const float* __restrict__ zzd,
    int64_t n, HEvent* __restrict__ out, int32_t* __restrict__ keep) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= n) return;
  HEvent ev;
  ev.first = lors[2 * e];
  ev.second = lors[2 * e + 1];
  ev.lor = (int32_t)((int64_t)ev.first * (ev.first + 1) / 2 + ev.second);
  const SegD s = seg[ev.lor];
  const float R = s.R;
  const float zu = zzd[3 * e], zd = zzd[3 * e + 1], dl = zzd[3 * e + 2];
  // Response::calculate_tan_y_z (2d/strip/response.h:22-37)
  ev.tan = (zu - zd) / (2 * R);
  ev.up = -0.5f * (dl / sqrtf(1 + ev.tan * ev.tan));
  ev.right = 0.5f * (zu + zd + (2 * ev.up * ev.tan));
  // GaussianKernel::ellipse_bb (gaussian_kernel.h:108-128,159-160)
  const float tsq = ev.tan * ev.tan;
  ev.sec = sqrtf(1 + ev.tan * ev.tan);
  const float A = 4 * ev.sec * ev.sec * k.isdl + 2 * tsq * k.isz;
  const float B = -4 * ev.tan * k.isz;
  const float C = 2 * k.isz;
  const float B_2 = (B / 2) * (B / 2);
  const float half_up = 3 / sqrtf(A - (B_2 / C));
  const float half_right = 3 / sqrtf(C - (B_2 / A));
  const float zl = ev.right - half_right, zr = ev.right + half_right;
  // plane(z): truncation toward zero to a short (reconstruction.h:149)
  const int16_t pl = (int16_t)(int)((zl - grid.z_left) / grid.ps);
  const int16_t pr = (int16_t)(int)((zr - grid.z_left) / grid.ps);
  ev.plane_begin = pl > 0 ? pl : (int16_t)0;
  const int16_t pe = (int16_t)(pr + 1);
  ev.plane_end = pe < grid.planes ? pe : (int16_t)grid.planes;
  const float t_up = (ev.up + half_up + R) / (2 * R);
  const float t_dn = (ev.up - half_up + R) / (2 * R);
  const int32_t b0 = (int32_t)lor_begin[ev.lor], b1 = (int32_t)lor_end[ev.lor];
  // upper bound of t_up, lower bound of t_dn in [b0, b1)
  int32_t lo = b0, hi = b1;
  while (lo < hi) {
    const int32_t m = lo + (hi - lo) / 2;
    if (!(t_up < positions[m])) lo = m + 1;
    else hi = m;
  }
  ev.end = lo + 1;
  lo = b0;
  hi = b1;
  while (lo < hi) {
    const int32_t m = lo + (hi - lo) / 2;
    if (positions[m] < t_dn) lo = m + 1;
    else hi = m;
  }
  ev.begin = lo;
  ev.extra = ev.end > b1 ? ev.end - 1 : -1;
  int ok = 0;
  if (!(ev.plane_end <= 0 || ev.plane_begin >= grid.planes)) {
    const int32_t stop = ev.end < n_infos ? ev.end : n_infos;
    for (int32_t i = ev.begin; i < stop; ++i) {
      const int32_t p = pixels[i];
      const int x = (int16_t)(p & 0xffff), y = (int16_t)(p >> 16);
      if (x >= 0 && x < grid.w && y >= 0 && y < grid.h && weights[i] > 0) {
        ok = 1;
        break;
      }
    }
  }
  out[e] = ev;
  keep[e] = ok;
}

// Kept events per block, then their scatter in input order.
__global__ __launch_bounds__(kBlock) void k_count_kept(const int32_t* __restrict__ keep,
                                                       int64_t n, int32_t* __restrict__ counts) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int c = __syncthreads_count(e < n && keep[e]);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

__global__ __launch_bounds__(kBlock) void k_scatter_kept(const HEvent* __restrict__ in,
                                                         const int32_t* __restrict__ keep,
                                                         int64_t n,
                                                         const int64_t* __restrict__ offsets,
                                                         HEvent* __restrict__ out) {
  __shared__ int wave_count[kWavesPerBlock];
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool k = e < n && keep[e];
  const unsigned long long mask = __ballot(k);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) wave_count[wave] = __popcll(mask);
  __syncthreads();
  int64_t pos = offsets[blockIdx.x];
  for (int w = 0; w < wave; ++w) pos += wave_count[w];
  pos += __popcll(mask & ((1ull << lane) - 1ull));
  if (k) out[pos] = in[e];
}

__device__ __forceinline__ float wave_sum_f(float v) {
  // fixed tree: the same order on every run
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// The per-pixel part of an event's voxel terms, one entry per pixel info of
// its range, kept in LDS for the event's (pixel, plane) walk:
//   y     the pixel centre's position along the LOR from its middle
//   dyt   (y - up) tan
//   kpix  c / sqrt(a S a) exp(-isdl br^2 / 2): the kernel's factors that do
//         not depend on z (gaussian_kernel.h:51-79 splits into these and
//         exp(-isz pq^2), pq = dz - dy tan)
//   irm, irp  1 / (R - y), 1 / (R + y) of the frame sensitivity
//   ws    weight x sensitivity (the denominator's factor)
//   w2    the backprojection's weight: that of the range's last pixel info
//         of the same pixel (the extra pixel info's for a pixel it repeats)
//   vox0  the pixel's first voxel in the plane-fastest layout
// With per-plane weights the two weights are read per voxel and the entry
// holds their pixel infos instead (info, winfo).
struct PixTerm {
  float y, dyt, kpix, irm, irp;
  union {
    struct {
      float ws, w2;
    };
    struct {
      int32_t info, winfo;
    };
  };
  int32_t vox0;
};
constexpr int kMaxPix = 192;  // table entries per wave; longer ranges walk directly
constexpr int kCacheVox = 1024;  // voxels per wave whose kernel x rho pass 2 reuses
// next-image copies the events add to in turn (summed by k_hybrid_finalize):
// fewer waves on the same addresses at once (9.37 -> 8.97 ms with four at
// 160 x 160 x 128, 1e6 events)
constexpr int kAccCopies = 4;

// z-dependent part of a voxel's weight for pixel term T at plane centre zc:
// exp(-isz pq^2) kpix over the frame sensitivity (gaussian_kernel.h:140-149)
// as products with the pixel's reciprocals.  kernel2d <= 0 (the reference's
// skip) is returned as is.
__device__ __forceinline__ float voxel_kernel(const HKernel& k, const PixTerm& T, float zc,
                                              float right) {
  const float pq = (zc - right) - T.dyt;
  const float kz = T.kpix * __expf(-k.isz * pq * pq);
  const float L2 = k.length / 2;
  const float a = (L2 - zc) * T.irm, b = (L2 + zc) * T.irp;
  const float c = -(L2 + zc) * T.irm, d = (zc - L2) * T.irp;
  // atan u - atan v = atan((u - v) / (1 + u v)), plus pi (or minus pi)
  // when 1 + u v < 0 and u > 0 (u < 0): one arctangent instead of two
  const float u = b < a ? b : a, v = c < d ? d : c;
  const float q = 1 + u * v;
  float span = atanf((u - v) * __builtin_amdgcn_rcpf(q));
  if (q < 0) span += u > 0 ? (float)M_PI : -(float)M_PI;
  return kz * (float)M_PI * __builtin_amdgcn_rcpf(span);
}

// One wave per event.  The image and the accumulator are plane-fastest
// (voxel (x, y, z) at z + planes (x + W y)), so the lanes of a wave, which
// walk consecutive planes of a pixel, read and add contiguous words.
// S3: with a 3D sensitivity volume (sens3, sens_depth planes).  MP: with
// per-plane weights as well (wmp: sens_depth weights per pixel info,
// plane-fastest), the reference's multi-plane geometry
// (reconstruction.h:345-356).
template <bool S3, bool MP = false>
__global__ __launch_bounds__(kBlock) void k_hybrid_mlem(
    const SegD* __restrict__ seg, const int32_t* __restrict__ pixels,
    const float* __restrict__ weights, const float* __restrict__ sens, int32_t n_infos,
    HGrid grid, HKernel k, const HEvent* __restrict__ events, int64_t n,
    const float* __restrict__ rho, float* __restrict__ acc, const float* __restrict__ sens3,
    int sens_depth, const float* __restrict__ wmp, const int32_t* __restrict__ next_same) {
  static_assert(S3 || !MP, "per-plane weights come with their sensitivity volume");
  __shared__ PixTerm table_all[kWavesPerBlock][kMaxPix];
  __shared__ float cache_all[kWavesPerBlock][kCacheVox];
  const int lane = threadIdx.x & 63;
  PixTerm* table = table_all[threadIdx.x >> 6];
  float* cache = cache_all[threadIdx.x >> 6];
  const int64_t wave0 = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int64_t n_waves = (int64_t)gridDim.x * kWavesPerBlock;
  const int planes = grid.planes;
  const int64_t np = (int64_t)grid.w * grid.h;
  for (int64_t e = wave0; e < n; e += n_waves) {
    const HEvent ev = events[e];
    const SegD s = seg[ev.lor];
    const int32_t end = ev.end < n_infos ? ev.end : n_infos;
    const int n_pix = end - ev.begin;
    const int n_pl = ev.plane_end - ev.plane_begin;
    if (n_pix <= 0 || n_pl <= 0) continue;
    const int32_t extra_pixel = ev.extra >= 0 && ev.extra < n_infos ? pixels[ev.extra] : -1;
    const float extra_w = extra_pixel != -1 ? weights[ev.extra] : 0.0f;
    const float sec_sq = ev.sec * ev.sec;
    // per-pixel terms (the pixel's position along the LOR as the reference
    // forms it, line_segment.h:80-84)
    auto pixel_term = [&](int i) {
      PixTerm T;
      const int32_t info = ev.begin + i;
      const int32_t p = pixels[info];
      const int px = (int16_t)(p & 0xffff), py = (int16_t)(p >> 16);
      const float cx = grid.llcx + (float)px * grid.ps, cy = grid.llcy + (float)py * grid.ps;
      T.y = (cx - s.mx) * s.dx + (cy - s.my) * s.dy;
      const float dy = T.y - ev.up;
      const float yp = dy + ev.up;
      const float ap = -(yp - s.R) * sec_sq, aq = -(yp + s.R) * sec_sq;
      const float ar = -2 * yp * ev.sec * ev.tan;
      const float br = -2 * dy * ev.sec;
      const float a_a = ap * k.isz * ap + aq * k.isz * aq + ar * k.isdl * ar;
      T.kpix = k.c * __builtin_amdgcn_rsqf(a_a) * __expf(-0.5f * (br * k.isdl * br));
      T.dyt = dy * ev.tan;
      T.irm = __builtin_amdgcn_rcpf(s.R - T.y);
      T.irp = __builtin_amdgcn_rcpf(s.R + T.y);
      const int32_t pix = px + grid.w * py;
      // the range's last pixel info of this pixel (the links run forward
      // within the LOR; the extra one past its end comes last of all)
      int32_t last = info;
      for (int32_t nx = next_same[info]; nx >= 0 && nx < end; nx = next_same[nx]) last = nx;
      if constexpr (MP) {
        T.info = info;
        T.winfo = p == extra_pixel ? ev.extra : last;
      } else {
        const float w = weights[info];
        T.ws = w * sens[pix];
        T.w2 = p == extra_pixel ? extra_w : (last == info ? w : weights[last]);
      }
      T.vox0 = pix * planes;
      return T;
    };
    const bool tabled = n_pix <= kMaxPix;
    if (tabled) {
      for (int i = lane; i < n_pix; i += 64) table[i] = pixel_term(i);
      // the wave's own writes, read back by its other lanes
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    const int64_t total = (int64_t)n_pix * n_pl;
    const int step_i = 64 / n_pl, step_z = 64 - step_i * n_pl;
    const int i0 = lane / n_pl, z0 = lane - i0 * n_pl;
    float den = 0.0f;
    {
      int i = i0, z = z0;
      for (int64_t j = lane; j < total; j += 64) {
        const PixTerm T = tabled ? table[i] : pixel_term(i);
        const int iz = ev.plane_begin + z;
        const float zc = ((float)iz + 0.5f) * grid.ps + grid.z_left;
        const float k2 = voxel_kernel(k, T, zc, ev.right);
        // kernel x rho for pass 2 (0 for the reference's skipped voxels)
        float kr = 0.0f;
        if (!(k2 <= 0)) {
          kr = k2 * rho[T.vox0 + iz];
          float ws = MP ? 1.0f : T.ws;
          if constexpr (S3) {
            // a volume of D planes, read at plane |iz - D|
            // (reconstruction.h:345-362); |iz - D| = D is one plane past the
            // volume there and reads plane D - 1 here
            int p = iz < sens_depth ? sens_depth - iz : iz - sens_depth;
            p = p < sens_depth - 1 ? p : sens_depth - 1;
            if constexpr (MP) ws = wmp[(int64_t)T.info * sens_depth + p];
            ws *= sens3[T.vox0 + p];
          }
          den += kr * ws;
        }
        if (j < kCacheVox) cache[j] = kr;
        i += step_i;
        z += step_z;
        if (z >= n_pl) {
          z -= n_pl;
          ++i;
        }
      }
    }
    den = wave_sum_f(den);
    if (den > 0) {
      const float inv = 1 / den;
      float* acc_e = acc + (int64_t)(e & (kAccCopies - 1)) * np * planes;
      int i = i0, z = z0;
      for (int64_t j = lane; j < total; j += 64) {
        const PixTerm T = tabled ? table[i] : pixel_term(i);
        const int iz = ev.plane_begin + z;
        const int32_t vox = T.vox0 + iz;
        float kr;
        if (j < kCacheVox) {
          kr = cache[j];
        } else {
          const float zc = ((float)iz + 0.5f) * grid.ps + grid.z_left;
          const float k2 = voxel_kernel(k, T, zc, ev.right);
          kr = !(k2 <= 0) ? k2 * rho[vox] : 0.0f;
        }
        float w2;
        if constexpr (MP) {
          int p = iz < sens_depth ? sens_depth - iz : iz - sens_depth;
          p = p < sens_depth - 1 ? p : sens_depth - 1;
          w2 = wmp[(int64_t)T.winfo * sens_depth + p];
        } else {
          w2 = T.w2;
        }
        if (kr != 0.0f) unsafeAtomicAdd(acc_e + vox, kr * w2 * inv);
        i += step_i;
        z += step_z;
        if (z >= n_pl) {
          z -= n_pl;
          ++i;
        }
      }
    }
    // the table is rewritten for the next event only after every lane's
    // reads of this one
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ __launch_bounds__(kBlock) void k_hybrid_finalize(float* __restrict__ acc,
                                                            float* __restrict__ rho, int64_t n) {
  const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (v >= n) return;
  float t = 0.0f;
  for (int c = 0; c < kAccCopies; ++c) {
    t += acc[c * n + v];
    acc[c * n + v] = 0.0f;
  }
  rho[v] = t;
}

__global__ __launch_bounds__(kBlock) void k_count_voxels(const HEvent* __restrict__ events,
                                                         int64_t n, int32_t n_infos,
                                                         unsigned long long* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long v = 0;
  if (e < n) {
    const HEvent ev = events[e];
    const int32_t end = ev.end < n_infos ? ev.end : n_infos;
    const int np = end - ev.begin, nz = ev.plane_end - ev.plane_begin;
    if (np > 0 && nz > 0) v = (unsigned long long)np * nz;
  }
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, v);
}

template <typename T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    return set_error(PETSTRIP_ENOMEM, "device allocation failed");
  }
  return PETSTRIP_OK;
}
template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

int hip_fail(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return set_error(e == hipErrorOutOfMemory ? PETSTRIP_ENOMEM : PETSTRIP_EDEVICE, m.c_str());
}

#define HTRY(expr)                                 \
  do {                                             \
    hipError_t _e = (expr);                        \
    if (_e != hipSuccess) return hip_fail(_e, #expr); \
  } while (0)

}  // namespace

struct petstrip_hybrid {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t it0 = nullptr, it1 = nullptr, k0 = nullptr, k1 = nullptr;
  HGrid grid{};
  HKernel kern{};
  int32_t n_detectors = 0;
  int64_t n_lors = 0, n_infos = 0, n_voxels_grid = 0;
  std::vector<int64_t> info_begin, info_end;
  std::vector<float> weights, sens;
  std::vector<float> raw_weights;  // before normalisation
  std::vector<int64_t> pix_index;  // pixel (x + W y) of every pixel info
  SegD* d_seg = nullptr;
  int64_t* d_lbeg = nullptr;
  int64_t* d_lend = nullptr;
  float* d_pos = nullptr;
  int32_t* d_pix = nullptr;
  int32_t* d_next = nullptr;  // next pixel info of the same pixel in the LOR, or -1
  float* d_w = nullptr;
  float* d_sens = nullptr;
  // external sensitivity volume (sens_depth planes, 0: none), plane-fastest
  // with the image's stride: plane p of pixel pix at pix * planes + p
  float* d_sens3 = nullptr;
  int sens_depth = 0;
  // per-plane weights (0: one set), sens_depth of them per pixel info,
  // plane-fastest
  int mp = 0;
  float* d_wmp = nullptr;
  // host copies for read-back: the normalised weight planes (mp of
  // n_infos, plane-major; empty for one set) and the sensitivity volume
  // (sens_depth planes of rows x columns; empty for the 2D map)
  std::vector<float> plane_weights, sens_volume;
  float* d_rho = nullptr;
  float* d_acc = nullptr;
  HEvent* d_events = nullptr;
  int64_t n_events = 0, cap_events = 0;
  int64_t n_voxels = 0;
  double last_iteration_ms = 0, last_kernel_ms = 0;
};

namespace {

int upload_weights(petstrip_hybrid* h) {
  HTRY(hipMemcpyAsync(h->d_w, h->weights.data(), h->n_infos * sizeof(float),
                      hipMemcpyHostToDevice, h->stream));
  HTRY(hipMemcpyAsync(h->d_sens, h->sens.data(), h->sens.size() * sizeof(float),
                      hipMemcpyHostToDevice, h->stream));
  HTRY(hipStreamSynchronize(h->stream));
  return PETSTRIP_OK;
}

int count_voxels(petstrip_hybrid* h) {
  unsigned long long* d = nullptr;
  int rc;
  if ((rc = dalloc(&d, 1))) return rc;
  unsigned long long v = 0;
  hipError_t e = hipMemsetAsync(d, 0, sizeof(v), h->stream);
  if (e == hipSuccess && h->n_events > 0) {
    hipLaunchKernelGGL(k_count_voxels, dim3((unsigned)((h->n_events + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, h->stream, h->d_events, h->n_events,
                       (int32_t)h->n_infos, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&v, d, sizeof(v), hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  dfree(d);
  if (e != hipSuccess) return hip_fail(e, "voxel count");
  h->n_voxels = (int64_t)v;
  return PETSTRIP_OK;
}

int reserve_events(petstrip_hybrid* h, int64_t n) {
  if (n > INT32_MAX) return set_error(PETSTRIP_EINVAL, "more than 2^31 - 1 events");
  if (n <= h->cap_events) return PETSTRIP_OK;
  dfree(h->d_events);
  h->cap_events = 0;
  int rc;
  if ((rc = dalloc(&h->d_events, (size_t)n))) return rc;
  h->cap_events = n;
  return PETSTRIP_OK;
}

int one_iteration(petstrip_hybrid* h) {
  HTRY(hipEventRecord(h->it0, h->stream));
  HTRY(hipEventRecord(h->k0, h->stream));
  if (h->n_events > 0) {
    // one wave per event; enough waves to fill the device several times over
    const int64_t blocks = std::min<int64_t>((h->n_events + kWavesPerBlock - 1) / kWavesPerBlock,
                                             256 * 64);
    auto kern = h->mp > 0           ? k_hybrid_mlem<true, true>
                : h->sens_depth > 0 ? k_hybrid_mlem<true>
                                    : k_hybrid_mlem<false>;
    hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(kBlock), 0, h->stream, h->d_seg,
                       h->d_pix, h->d_w, h->d_sens, (int32_t)h->n_infos, h->grid, h->kern,
                       h->d_events, h->n_events, h->d_rho, h->d_acc, h->d_sens3,
                       h->sens_depth, h->d_wmp, h->d_next);
    HTRY(hipGetLastError());
  }
  HTRY(hipEventRecord(h->k1, h->stream));
  hipLaunchKernelGGL(k_hybrid_finalize,
                     dim3((unsigned)((h->n_voxels_grid + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, h->stream, h->d_acc, h->d_rho, h->n_voxels_grid);
  HTRY(hipGetLastError());
  HTRY(hipEventRecord(h->it1, h->stream));
  HTRY(hipEventSynchronize(h->it1));
  float ms = 0;
  HTRY(hipEventElapsedTime(&ms, h->it0, h->it1));
  h->last_iteration_ms = ms;
  HTRY(hipEventElapsedTime(&ms, h->k0, h->k1));
  h->last_kernel_ms = ms;
  return PETSTRIP_OK;
}

}  // namespace

extern "C" {

int petstrip_hybrid_create(const petstrip_hybrid_geometry* g, double sigma_z, double sigma_dl,
                           double length, int32_t flags, int32_t device,
                           petstrip_hybrid** out) {
  if (!g || !out) return set_error(PETSTRIP_EINVAL, "bad argument");
  *out = nullptr;
  if (g->n_detectors < 1 || g->n_detectors > 32767 || g->n_columns < 1 || g->n_rows < 1 ||
      g->n_planes < 1 || g->n_columns > 32767 || g->n_rows > 32767 || g->n_planes > 32767 ||
      !(g->pixel_size > 0) || g->n_pixel_infos < 1 || g->n_pixel_infos >= kMaxInfos)
    return set_error(PETSTRIP_EINVAL, "bad hybrid geometry sizes");
  if (!g->segments || !g->widths || !g->info_begin || !g->info_end || !g->pixels ||
      !g->positions || ((flags & PETSTRIP_HYBRID_GEOMETRY_WEIGHTS) && !g->weights))
    return set_error(PETSTRIP_EINVAL, "hybrid geometry array missing");
  if (!(sigma_z > 0) || !(sigma_dl > 0) || !(length > 0))
    return set_error(PETSTRIP_EINVAL, "sigma_z, sigma_dl and length must be positive");
  const int64_t n_lors = (int64_t)(g->n_detectors + 1) * g->n_detectors / 2;
  for (int64_t l = 0; l < n_lors; ++l)
    if (g->info_begin[l] < 0 || g->info_end[l] < g->info_begin[l] ||
        g->info_end[l] > g->n_pixel_infos)
      return set_error(PETSTRIP_EINVAL, "hybrid geometry: bad pixel-info range");
  for (int64_t i = 0; i < g->n_pixel_infos; ++i) {
    const int x = g->pixels[2 * i], y = g->pixels[2 * i + 1];
    if (x < 0 || x >= g->n_columns || y < 0 || y >= g->n_rows)
      return set_error(PETSTRIP_EINVAL, "hybrid geometry: pixel outside the grid");
  }
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || n_dev < 1 || device < 0 || device >= n_dev)
    return set_error(PETSTRIP_EDEVICE, "no such HIP device");
  auto* h = new petstrip_hybrid();
  h->device = device;
  h->n_detectors = g->n_detectors;
  h->n_lors = n_lors;
  h->n_infos = g->n_pixel_infos;
  h->grid.w = g->n_columns;
  h->grid.h = g->n_rows;
  h->grid.planes = g->n_planes;
  h->grid.ps = (float)g->pixel_size;
  // PixelGrid::lower_left_center (2d/geometry/pixel_grid.h:22)
  h->grid.llcx = (float)g->lower_left_x + h->grid.ps / 2;
  h->grid.llcy = (float)g->lower_left_y + h->grid.ps / 2;
  h->grid.z_left = (float)g->z_left;
  h->kern = make_kernel(sigma_z, sigma_dl, length);
  h->n_voxels_grid = (int64_t)g->n_columns * g->n_rows * g->n_planes;
  h->info_begin.assign(g->info_begin, g->info_begin + n_lors);
  h->info_end.assign(g->info_end, g->info_end + n_lors);
  std::vector<SegD> segs(n_lors);
  for (int64_t l = 0; l < n_lors; ++l) segs[l] = make_seg(g->segments + 4 * l);
  // weights: the Gaussian in the distance to the LOR, sigma 0.3 width, a
  // double normalisation times a float exponential (reconstruction.h:212-234)
  h->weights.resize(h->n_infos);
  if (flags & PETSTRIP_HYBRID_GEOMETRY_WEIGHTS) {
    std::memcpy(h->weights.data(), g->weights, h->n_infos * sizeof(float));
  } else {
    for (int64_t l = 0; l < n_lors; ++l) {
      const SegD& s = segs[l];
      const float sw = 0.3f * g->widths[l];
      const double norm = 1 / (sw * std::sqrt(2 * M_PI));
      const float inv2 = 1 / (2 * sw * sw);
      for (int64_t i = g->info_begin[l]; i < g->info_end[l]; ++i) {
        const float cx = h->grid.llcx + (float)g->pixels[2 * i] * h->grid.ps;
        const float cy = h->grid.llcy + (float)g->pixels[2 * i + 1] * h->grid.ps;
        const float d = (cx * s.nx + cy * s.ny) - s.dfo;
        h->weights[i] = (float)(norm * std::exp(-d * d * inv2));
      }
    }
  }
  // sensitivity and the normalised weights (reconstruction.h:236-265)
  const size_t np = (size_t)g->n_columns * g->n_rows;
  h->raw_weights = h->weights;
  h->pix_index.resize(h->n_infos);
  for (int64_t i = 0; i < h->n_infos; ++i)
    h->pix_index[i] = g->pixels[2 * i] + (int64_t)g->n_columns * g->pixels[2 * i + 1];
  h->sens.assign(np, (flags & PETSTRIP_HYBRID_SENSITIVITY_ONE) ? 1.0f : 0.0f);
  if (!(flags & PETSTRIP_HYBRID_SENSITIVITY_ONE))
    for (int64_t i = 0; i < h->n_infos; ++i)
      h->sens[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1]] += h->weights[i];
  for (int64_t i = 0; i < h->n_infos; ++i)
    h->weights[i] /= h->sens[g->pixels[2 * i] + (size_t)g->n_columns * g->pixels[2 * i + 1]];
  std::vector<int32_t> pix(h->n_infos);
  for (int64_t i = 0; i < h->n_infos; ++i)
    pix[i] = (int32_t)(uint16_t)g->pixels[2 * i] | ((int32_t)g->pixels[2 * i + 1] << 16);
  // links between a LOR's pixel infos of one pixel, walked backwards with
  // the last index seen per pixel
  std::vector<int32_t> next(h->n_infos, -1);
  {
    std::vector<int32_t> seen(np, -1);
    for (int64_t l = 0; l < n_lors; ++l) {
      for (int64_t i = g->info_end[l] - 1; i >= g->info_begin[l]; --i) {
        int32_t& s = seen[h->pix_index[i]];
        next[i] = s;
        s = (int32_t)i;
      }
      for (int64_t i = g->info_begin[l]; i < g->info_end[l]; ++i) seen[h->pix_index[i]] = -1;
    }
  }

  auto fail_free = [&](int rc) {
    petstrip_hybrid_destroy(h);
    return rc;
  };
  if (hipSetDevice(device) != hipSuccess) return fail_free(set_error(PETSTRIP_EDEVICE, "hipSetDevice"));
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&h->it0) != hipSuccess || hipEventCreate(&h->it1) != hipSuccess ||
      hipEventCreate(&h->k0) != hipSuccess || hipEventCreate(&h->k1) != hipSuccess)
    return fail_free(set_error(PETSTRIP_EDEVICE, "stream or event creation failed"));
  int rc;
  if ((rc = dalloc(&h->d_seg, n_lors)) || (rc = dalloc(&h->d_lbeg, n_lors)) ||
      (rc = dalloc(&h->d_lend, n_lors)) || (rc = dalloc(&h->d_pos, h->n_infos)) ||
      (rc = dalloc(&h->d_pix, h->n_infos)) || (rc = dalloc(&h->d_next, h->n_infos)) ||
      (rc = dalloc(&h->d_w, h->n_infos)) ||
      (rc = dalloc(&h->d_sens, np)) || (rc = dalloc(&h->d_rho, h->n_voxels_grid)) ||
      (rc = dalloc(&h->d_acc, kAccCopies * h->n_voxels_grid)))
    return fail_free(rc);
  std::vector<float> ones(h->n_voxels_grid, 1.0f);
  hipError_t e = hipMemcpy(h->d_seg, segs.data(), n_lors * sizeof(SegD), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_lbeg, g->info_begin, n_lors * sizeof(int64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_lend, g->info_end, n_lors * sizeof(int64_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_pos, g->positions, h->n_infos * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_pix, pix.data(), h->n_infos * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_next, next.data(), h->n_infos * sizeof(int32_t), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->d_rho, ones.data(), h->n_voxels_grid * sizeof(float),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemset(h->d_acc, 0, kAccCopies * h->n_voxels_grid * sizeof(float));
  if (e != hipSuccess) return fail_free(hip_fail(e, "hybrid geometry upload"));
  if ((rc = upload_weights(h))) return fail_free(rc);
  *out = h;
  return PETSTRIP_OK;
}

void petstrip_hybrid_destroy(petstrip_hybrid* h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  dfree(h->d_seg);
  dfree(h->d_lbeg);
  dfree(h->d_lend);
  dfree(h->d_pos);
  dfree(h->d_pix);
  dfree(h->d_next);
  dfree(h->d_w);
  dfree(h->d_sens);
  dfree(h->d_sens3);
  dfree(h->d_wmp);
  dfree(h->d_rho);
  dfree(h->d_acc);
  dfree(h->d_events);
  for (hipEvent_t ev : {h->it0, h->it1, h->k0, h->k1})
    if (ev) (void)hipEventDestroy(ev);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int petstrip_hybrid_set_responses(petstrip_hybrid* h, const int16_t* lors, const float* zzd,
                                  int64_t n) {
  if (!h || n < 0 || (n > 0 && (!lors || !zzd)))
    return set_error(PETSTRIP_EINVAL, "bad argument");
  for (int64_t e = 0; e < n; ++e) {
    const int f = lors[2 * e], s = lors[2 * e + 1];
    if (s < 0 || f < s || f >= h->n_detectors)
      return set_error(PETSTRIP_EINVAL, "response LOR out of range (first >= second >= 0)");
  }
  HTRY(hipSetDevice(h->device));
  int rc;
  if ((rc = reserve_events(h, n))) return rc;
  h->n_events = 0;
  if (n > 0) {
    int16_t* d_l = nullptr;
    float* d_z = nullptr;
    HEvent* d_tmp = nullptr;
    int32_t* d_keep = nullptr;
    int32_t* d_counts = nullptr;
    int64_t* d_off = nullptr;
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    auto cleanup = [&]() {
      dfree(d_l);
      dfree(d_z);
      dfree(d_tmp);
      dfree(d_keep);
      dfree(d_counts);
      dfree(d_off);
    };
    if ((rc = dalloc(&d_l, 2 * n)) || (rc = dalloc(&d_z, 3 * n)) ||
        (rc = dalloc(&d_tmp, n)) || (rc = dalloc(&d_keep, n)) ||
        (rc = dalloc(&d_counts, blocks)) || (rc = dalloc(&d_off, blocks))) {
      cleanup();
      return rc;
    }
    std::vector<int32_t> counts(blocks);
    std::vector<int64_t> off(blocks);
    hipError_t e = hipMemcpyAsync(d_l, lors, 2 * n * sizeof(int16_t), hipMemcpyHostToDevice,
                                  h->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_z, zzd, 3 * n * sizeof(float), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_hybrid_frames, dim3((unsigned)blocks), dim3(kBlock), 0, h->stream,
                         h->d_seg, h->d_lbeg, h->d_lend, h->d_pos, h->d_pix, h->d_w,
                         (int32_t)h->n_infos, h->grid, h->kern, d_l, d_z, n, d_tmp, d_keep);
      e = hipGetLastError();
    }
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_count_kept, dim3((unsigned)blocks), dim3(kBlock), 0, h->stream,
                         d_keep, n, d_counts);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(counts.data(), d_counts, blocks * sizeof(int32_t),
                         hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    int64_t total = 0;
    for (int64_t b = 0; b < blocks; ++b) {
      off[b] = total;
      total += counts[b];
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_off, off.data(), blocks * sizeof(int64_t), hipMemcpyHostToDevice,
                         h->stream);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_scatter_kept, dim3((unsigned)blocks), dim3(kBlock), 0, h->stream,
                         d_tmp, d_keep, n, d_off, h->d_events);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    cleanup();
    if (e != hipSuccess) return hip_fail(e, "hybrid frame events");
    h->n_events = total;
  }
  return count_voxels(h);
}

int petstrip_hybrid_set_events(petstrip_hybrid* h, const petstrip_hybrid_event* events,
                               int64_t n) {
  if (!h || n < 0 || (n > 0 && !events)) return set_error(PETSTRIP_EINVAL, "bad argument");
  std::vector<HEvent> dev(n);
  for (int64_t e = 0; e < n; ++e) {
    const auto& s = events[e];
    if (s.second < 0 || s.first < s.second || s.first >= h->n_detectors || s.begin < 0 ||
        s.end < s.begin || s.end > h->n_infos + 1 || s.plane_begin < 0 ||
        s.plane_end > h->grid.planes)
      return set_error(PETSTRIP_EINVAL, "frame event out of range");
    HEvent& d = dev[e];
    d.first = s.first;
    d.second = s.second;
    d.lor = (int32_t)((int64_t)s.first * (s.first + 1) / 2 + s.second);
    d.up = s.up;
    d.right = s.right;
    d.tan = s.tan;
    d.sec = s.sec;
    d.begin = (int32_t)s.begin;
    d.end = (int32_t)s.end;
    d.extra = s.end > h->info_end[d.lor] ? (int32_t)s.end - 1 : -1;
    d.plane_begin = s.plane_begin;
    d.plane_end = s.plane_end;
  }
  HTRY(hipSetDevice(h->device));
  int rc;
  if ((rc = reserve_events(h, n))) return rc;
  if (n > 0)
    HTRY(hipMemcpy(h->d_events, dev.data(), n * sizeof(HEvent), hipMemcpyHostToDevice));
  h->n_events = n;
  return count_voxels(h);
}

int petstrip_hybrid_get_events(petstrip_hybrid* h, petstrip_hybrid_event* events) {
  if (!h || (h->n_events > 0 && !events)) return set_error(PETSTRIP_EINVAL, "bad argument");
  if (h->n_events == 0) return PETSTRIP_OK;
  std::vector<HEvent> dev(h->n_events);
  HTRY(hipSetDevice(h->device));
  HTRY(hipMemcpy(dev.data(), h->d_events, h->n_events * sizeof(HEvent), hipMemcpyDeviceToHost));
  for (int64_t e = 0; e < h->n_events; ++e) {
    const HEvent& d = dev[e];
    petstrip_hybrid_event& s = events[e];
    s.first = d.first;
    s.second = d.second;
    s.up = d.up;
    s.right = d.right;
    s.tan = d.tan;
    s.sec = d.sec;
    s.begin = d.begin;
    s.end = d.end;
    s.plane_begin = d.plane_begin;
    s.plane_end = d.plane_end;
  }
  return PETSTRIP_OK;
}

int petstrip_hybrid_set_weights(petstrip_hybrid* h, const float* weights,
                                const float* sensitivity) {
  if (!h) return set_error(PETSTRIP_EINVAL, "bad argument");
  if (weights) std::memcpy(h->weights.data(), weights, h->n_infos * sizeof(float));
  if (sensitivity) {
    std::memcpy(h->sens.data(), sensitivity, h->sens.size() * sizeof(float));
    h->sens_depth = 0;  // the map replaces an external volume
    h->mp = 0;
    h->plane_weights.clear();
    h->sens_volume.clear();
  }
  HTRY(hipSetDevice(h->device));
  return upload_weights(h);
}

// reconstruction_cmd.cpp:499-519 with --sensitivity: the geometry weights
// normalised by the volume's plane 0 (reconstruction.h:255-271) and the
// volume in the denominator (reconstruction.h:357-362).
int petstrip_hybrid_set_sensitivity_3d(petstrip_hybrid* h, const float* sensitivity,
                                       int32_t depth) {
  if (!h || !sensitivity) return set_error(PETSTRIP_EINVAL, "bad argument");
  const int planes = h->grid.planes;
  if (depth < 1 || depth != planes / 2)
    return set_error(PETSTRIP_EINVAL, "the sensitivity volume must have n_planes / 2 planes");
  const size_t np = (size_t)h->grid.w * h->grid.h;
  std::vector<float> t((size_t)h->n_voxels_grid, 0.0f);
  for (int p = 0; p < depth; ++p)
    for (size_t pix = 0; pix < np; ++pix) t[pix * planes + p] = sensitivity[pix + np * p];
  for (int64_t i = 0; i < h->n_infos; ++i)
    h->weights[i] = h->raw_weights[i] / sensitivity[h->pix_index[i]];
  h->sens.assign(np, 1.0f);
  HTRY(hipSetDevice(h->device));
  int rc;
  if (!h->d_sens3 && (rc = dalloc(&h->d_sens3, (size_t)h->n_voxels_grid))) return rc;
  HTRY(hipMemcpyAsync(h->d_sens3, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice,
                      h->stream));
  h->sens_depth = depth;
  h->mp = 0;
  h->plane_weights.clear();
  h->sens_volume.assign(sensitivity, sensitivity + (size_t)depth * np);
  return upload_weights(h);
}

int petstrip_hybrid_set_plane_weights(petstrip_hybrid* h, const float* raw_weights,
                                      int32_t n_planes_half, int32_t sensitivity_mode,
                                      const float* sensitivity) {
  if (!h || !raw_weights || sensitivity_mode < 0 || sensitivity_mode > 2 ||
      (sensitivity_mode == 2 && !sensitivity))
    return set_error(PETSTRIP_EINVAL, "bad argument");
  const int planes = h->grid.planes;
  const int64_t ni = h->n_infos;
  const size_t np = (size_t)h->grid.w * h->grid.h;
  if (n_planes_half == 1) {
    // one set: the geometry weights with the 2D map or the external volume
    h->raw_weights.assign(raw_weights, raw_weights + ni);
    if (sensitivity_mode == 2) return petstrip_hybrid_set_sensitivity_3d(h, sensitivity, planes / 2);
    h->sens.assign(np, sensitivity_mode == 1 ? 1.0f : 0.0f);
    if (sensitivity_mode == 0)
      for (int64_t i = 0; i < ni; ++i) h->sens[h->pix_index[i]] += raw_weights[i];
    for (int64_t i = 0; i < ni; ++i) h->weights[i] = raw_weights[i] / h->sens[h->pix_index[i]];
    h->sens_depth = 0;
    h->mp = 0;
    h->plane_weights.clear();
    h->sens_volume.clear();
    HTRY(hipSetDevice(h->device));
    return upload_weights(h);
  }
  if (n_planes_half < 1 || n_planes_half != planes / 2)
    return set_error(PETSTRIP_EINVAL,
                     "you must specify either one or half no. planes matrices");
  const int H = n_planes_half;
  // the sensitivity volume (calculate_sensitivity: per pixel and plane, the
  // plane's weights in pixel-info order) and the normalised weights
  std::vector<float> vol((size_t)H * np, sensitivity_mode == 1 ? 1.0f : 0.0f);
  if (sensitivity_mode == 2) {
    std::memcpy(vol.data(), sensitivity, vol.size() * sizeof(float));
  } else if (sensitivity_mode == 0) {
    for (int64_t i = 0; i < ni; ++i)
      for (int p = 0; p < H; ++p) vol[h->pix_index[i] + np * p] += raw_weights[p * ni + i];
  }
  std::vector<float> w((size_t)ni * H);  // plane-fastest per pixel info
  for (int64_t i = 0; i < ni; ++i)
    for (int p = 0; p < H; ++p)
      w[(size_t)i * H + p] = raw_weights[p * ni + i] / vol[h->pix_index[i] + np * p];
  std::vector<float> t((size_t)h->n_voxels_grid, 0.0f);
  for (int p = 0; p < H; ++p)
    for (size_t pix = 0; pix < np; ++pix) t[pix * planes + p] = vol[pix + np * p];
  // plane 0 is what the frame events' grid filter reads (reconstruction.h:152-161)
  h->raw_weights.assign(raw_weights, raw_weights + ni);
  for (int64_t i = 0; i < ni; ++i) h->weights[i] = w[(size_t)i * H];
  h->sens.assign(np, 1.0f);
  HTRY(hipSetDevice(h->device));
  int rc;
  if (!h->d_sens3 && (rc = dalloc(&h->d_sens3, (size_t)h->n_voxels_grid))) return rc;
  dfree(h->d_wmp);
  if ((rc = dalloc(&h->d_wmp, w.size()))) return rc;
  HTRY(hipMemcpyAsync(h->d_sens3, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice,
                      h->stream));
  HTRY(hipMemcpyAsync(h->d_wmp, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice,
                      h->stream));
  h->sens_depth = H;
  h->mp = H;
  h->plane_weights.resize(w.size());
  for (int p = 0; p < H; ++p)
    for (int64_t i = 0; i < ni; ++i) h->plane_weights[p * ni + i] = w[(size_t)i * H + p];
  h->sens_volume = std::move(vol);
  return upload_weights(h);
}

int petstrip_hybrid_get_plane_weights(petstrip_hybrid* h, float* weights, float* sensitivity,
                                      int32_t* n_weight_planes,
                                      int32_t* n_sensitivity_planes) {
  if (!h) return set_error(PETSTRIP_EINVAL, "bad argument");
  const size_t np = (size_t)h->grid.w * h->grid.h;
  const int32_t wp = h->mp > 0 ? h->mp : 1;
  const int32_t sp = h->sens_depth > 0 ? h->sens_depth : 1;
  if (n_weight_planes) *n_weight_planes = wp;
  if (n_sensitivity_planes) *n_sensitivity_planes = sp;
  if (weights) {
    const float* src = h->mp > 0 ? h->plane_weights.data() : h->weights.data();
    std::memcpy(weights, src, (size_t)wp * h->n_infos * sizeof(float));
  }
  if (sensitivity) {
    const float* src = h->sens_depth > 0 ? h->sens_volume.data() : h->sens.data();
    std::memcpy(sensitivity, src, (size_t)sp * np * sizeof(float));
  }
  return PETSTRIP_OK;
}

int petstrip_hybrid_get_weights(petstrip_hybrid* h, float* weights, float* sensitivity) {
  if (!h) return set_error(PETSTRIP_EINVAL, "bad argument");
  if (weights) std::memcpy(weights, h->weights.data(), h->n_infos * sizeof(float));
  if (sensitivity) std::memcpy(sensitivity, h->sens.data(), h->sens.size() * sizeof(float));
  return PETSTRIP_OK;
}

// The image crosses the boundary as planes x rows x columns (the reference's
// VoxelMap, voxel.h:27-29) and lives on the device plane-fastest.
int petstrip_hybrid_set_rho(petstrip_hybrid* h, const float* rho) {
  if (!h || !rho) return set_error(PETSTRIP_EINVAL, "bad argument");
  const int64_t np = (int64_t)h->grid.w * h->grid.h, P = h->grid.planes;
  std::vector<float> t(h->n_voxels_grid);
  for (int64_t z = 0; z < P; ++z)
    for (int64_t p = 0; p < np; ++p) t[p * P + z] = rho[z * np + p];
  HTRY(hipSetDevice(h->device));
  HTRY(hipMemcpy(h->d_rho, t.data(), h->n_voxels_grid * sizeof(float), hipMemcpyHostToDevice));
  return PETSTRIP_OK;
}

int petstrip_hybrid_get_rho(petstrip_hybrid* h, float* rho) {
  if (!h || !rho) return set_error(PETSTRIP_EINVAL, "bad argument");
  const int64_t np = (int64_t)h->grid.w * h->grid.h, P = h->grid.planes;
  std::vector<float> t(h->n_voxels_grid);
  HTRY(hipSetDevice(h->device));
  HTRY(hipMemcpy(t.data(), h->d_rho, h->n_voxels_grid * sizeof(float), hipMemcpyDeviceToHost));
  for (int64_t p = 0; p < np; ++p)
    for (int64_t z = 0; z < P; ++z) rho[z * np + p] = t[p * P + z];
  return PETSTRIP_OK;
}

int petstrip_hybrid_iterate(petstrip_hybrid* h, int32_t n) {
  if (!h || n < 0) return set_error(PETSTRIP_EINVAL, "bad argument");
  HTRY(hipSetDevice(h->device));
  for (int32_t i = 0; i < n; ++i) {
    int rc = one_iteration(h);
    if (rc) return rc;
  }
  return PETSTRIP_OK;
}

int petstrip_hybrid_run(petstrip_hybrid* h, float* rho_inout, int32_t n_blocks,
                        int32_t n_iterations_in_block, int32_t start_iteration,
                        petstrip_hybrid_output_cb output, petstrip_progress_cb progress,
                        void* user) {
  if (!h || !rho_inout || n_blocks < 0 || n_iterations_in_block < 0)
    return set_error(PETSTRIP_EINVAL, "bad argument");
  int rc;
  if ((rc = petstrip_hybrid_set_rho(h, rho_inout))) return rc;
  auto emit = [&](int iteration) -> int {
    if ((rc = petstrip_hybrid_get_rho(h, rho_inout))) return rc;
    if (output && output(user, iteration, rho_inout, h->grid.w, h->grid.h, h->grid.planes))
      return set_error(PETSTRIP_ECALLBACK, "output callback stopped the run");
    return PETSTRIP_OK;
  };
  for (int b = 0; b < n_blocks; ++b) {
    for (int i = 0; i < n_iterations_in_block; ++i) {
      const int it = b * n_iterations_in_block + i;
      if (it < start_iteration) continue;
      if (progress) progress(user, it, 0);
      if ((rc = petstrip_hybrid_iterate(h, 1))) return rc;
      if (progress) progress(user, it, 1);
      // the first block also outputs iterations 1-5, 10, 15, 20, 30, 50, 100
      if (!b && i < n_iterations_in_block - 1 &&
          (i < 5 || i == 9 || i == 14 || i == 19 || i == 29 || i == 49 || i == 99))
        if ((rc = emit(i + 1))) return rc;
    }
    const int it = (b + 1) * n_iterations_in_block;
    if (it <= start_iteration) continue;
    if ((rc = emit(it))) return rc;
  }
  if (progress) progress(user, n_blocks * n_iterations_in_block, 0);
  return PETSTRIP_OK;
}

int petstrip_hybrid_stats_get(petstrip_hybrid* h, petstrip_hybrid_stats* out) {
  if (!h || !out) return set_error(PETSTRIP_EINVAL, "bad argument");
  out->n_events = h->n_events;
  out->n_voxels = h->n_voxels;
  out->last_iteration_ms = h->last_iteration_ms;
  out->last_kernel_ms = h->last_kernel_ms;
  return PETSTRIP_OK;
}

}  // extern "C"
