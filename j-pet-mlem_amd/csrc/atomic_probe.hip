// Atomic-throughput probes (libpetstrip_probe.so, a measurement aid beside
// the reconstruction library).  They time, on an otherwise idle device, the
// two kinds of atomic update the binned kernel makes:
//   - global atomics in the pattern of its window flush: each wave adds to a
//     tile of 16 columns x 4 rows of a W x H image (64-byte row segments),
//     one atomic per lane, no return value, waves on distinct tiles;
//   - LDS ds_add_u64 (and, for comparison, ds_add_u32 and ds_add_f32) in
//     the pattern of its walks: the 64 lanes of a wave on
//     64 consecutive 64-bit words (no bank conflicts), 16 waves per
//     workgroup.
// The rates give the ceilings that the kernel's atomic counts are compared
// with.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "petstrip_probe.h"

namespace {

constexpr int kGlobalThreads = 256;
constexpr int kLdsThreads = 1024;
constexpr int kLdsWords = kLdsThreads / 64 * 256;  // four 64-word slots per wave

__device__ __forceinline__ void atomic_one(float* p) { unsafeAtomicAdd(p, 1.0f); }
__device__ __forceinline__ void atomic_one(double* p) { unsafeAtomicAdd(p, 1.0); }
__device__ __forceinline__ void atomic_one(unsigned long long* p) { atomicAdd(p, 1ull); }

// Wave w adds to tiles t = w, w + step, w + 2 step, ... (mod ntiles); with at
// most ntiles waves in the grid, no two waves are on one tile at once.
template <typename T>
__global__ __launch_bounds__(kGlobalThreads) void k_global_atomics(T* img, int width,
                                                                   int ntiles_x, int ntiles,
                                                                   int step, int reps) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (kGlobalThreads / 64) + (threadIdx.x >> 6);
  const int c = lane & 15, r = lane >> 4;
  int t = wave;  // < ntiles by the launch
  for (int i = 0; i < reps; ++i) {
    const int ty = t / ntiles_x, tx = t - ty * ntiles_x;
    atomic_one(img + ((size_t)(ty * 4 + r) * width + tx * 16 + c));
    t += step;
    if (t >= ntiles) t -= ntiles;
  }
}

// T = unsigned long long (ds_add_u64), unsigned (ds_add_u32) or float
// (ds_add_f32).
template <typename T>
__global__ __launch_bounds__(kLdsThreads) void k_lds_atomics(unsigned long long* sink,
                                                             int reps) {
  __shared__ T buf[kLdsWords];
  for (int i = threadIdx.x; i < kLdsWords; i += kLdsThreads) buf[i] = T(0);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T* p = buf + wave * 256 + lane;
  for (int i = 0; i < reps; ++i)
    __hip_atomic_fetch_add(p + (i & 3) * 64, T(1), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
  __syncthreads();
  // one word per workgroup keeps the updates observable
  if (threadIdx.x == 0) sink[blockIdx.x] = (unsigned long long)(buf[0] + buf[kLdsWords - 1]);
}

template <typename T>
hipError_t time_global(int width, int height, int reps, double* atomics, float* ms) {
  const int ntiles_x = width / 16, ntiles = ntiles_x * (height / 4);
  const int nblocks = ntiles / (kGlobalThreads / 64);
  if (nblocks < 1) return hipErrorInvalidValue;
  const int nwaves = nblocks * (kGlobalThreads / 64);
  const int step = 97 % ntiles;
  T* img = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t rc = hipMalloc(&img, (size_t)width * height * sizeof(T));
  if (rc == hipSuccess) rc = hipMemset(img, 0, (size_t)width * height * sizeof(T));
  if (rc == hipSuccess) rc = hipEventCreate(&e0);
  if (rc == hipSuccess) rc = hipEventCreate(&e1);
  if (rc == hipSuccess) {
    // warm-up, then the timed launch
    hipLaunchKernelGGL(k_global_atomics<T>, dim3(nblocks), dim3(kGlobalThreads), 0, 0, img,
                       width, ntiles_x, ntiles, step, 16);
    rc = hipGetLastError();
  }
  if (rc == hipSuccess) rc = hipEventRecord(e0, 0);
  if (rc == hipSuccess) {
    hipLaunchKernelGGL(k_global_atomics<T>, dim3(nblocks), dim3(kGlobalThreads), 0, 0, img,
                       width, ntiles_x, ntiles, step, reps);
    rc = hipGetLastError();
  }
  if (rc == hipSuccess) rc = hipEventRecord(e1, 0);
  if (rc == hipSuccess) rc = hipEventSynchronize(e1);
  if (rc == hipSuccess) rc = hipEventElapsedTime(ms, e0, e1);
  *atomics = (double)nwaves * 64.0 * reps;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (img) (void)hipFree(img);
  return rc;
}

template <typename T>
hipError_t time_lds(int reps, double* atomics, float* ms) {
  int dev = 0, ncu = 0;
  hipError_t rc = hipGetDevice(&dev);
  if (rc == hipSuccess) rc = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  if (rc != hipSuccess) return rc;
  const int nblocks = 4 * ncu;
  unsigned long long* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  rc = hipMalloc(&sink, (size_t)nblocks * sizeof(unsigned long long));
  if (rc == hipSuccess) rc = hipEventCreate(&e0);
  if (rc == hipSuccess) rc = hipEventCreate(&e1);
  if (rc == hipSuccess) {
    hipLaunchKernelGGL(k_lds_atomics<T>, dim3(nblocks), dim3(kLdsThreads), 0, 0, sink, 64);
    rc = hipGetLastError();
  }
  if (rc == hipSuccess) rc = hipEventRecord(e0, 0);
  if (rc == hipSuccess) {
    hipLaunchKernelGGL(k_lds_atomics<T>, dim3(nblocks), dim3(kLdsThreads), 0, 0, sink, reps);
    rc = hipGetLastError();
  }
  if (rc == hipSuccess) rc = hipEventRecord(e1, 0);
  if (rc == hipSuccess) rc = hipEventSynchronize(e1);
  if (rc == hipSuccess) rc = hipEventElapsedTime(ms, e0, e1);
  *atomics = (double)nblocks * kLdsThreads * reps;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (sink) (void)hipFree(sink);
  return rc;
}

}  // namespace

extern "C" int petstrip_probe_atomics(int device, int kind, int32_t width, int32_t height,
                                      int32_t reps, double* atomics_per_second,
                                      double* milliseconds) {
  if (!atomics_per_second || reps < 1 || kind < PETSTRIP_PROBE_GLOBAL_F32 ||
      kind > PETSTRIP_PROBE_LDS_F32)
    return (int)hipErrorInvalidValue;
  const bool lds = kind >= PETSTRIP_PROBE_LDS_U64;
  if (!lds && (width < 16 || height < 4 || width > (1 << 15) || height > (1 << 15)))
    return (int)hipErrorInvalidValue;
  hipError_t rc = hipSetDevice(device);
  if (rc != hipSuccess) return (int)rc;
  double n = 0.0;
  float ms = 0.0f;
  switch (kind) {
    case PETSTRIP_PROBE_GLOBAL_F32: rc = time_global<float>(width, height, reps, &n, &ms); break;
    case PETSTRIP_PROBE_GLOBAL_F64: rc = time_global<double>(width, height, reps, &n, &ms); break;
    case PETSTRIP_PROBE_GLOBAL_U64:
      rc = time_global<unsigned long long>(width, height, reps, &n, &ms);
      break;
    case PETSTRIP_PROBE_LDS_U32: rc = time_lds<unsigned>(reps, &n, &ms); break;
    case PETSTRIP_PROBE_LDS_F32: rc = time_lds<float>(reps, &n, &ms); break;
    default: rc = time_lds<unsigned long long>(reps, &n, &ms); break;
  }
  if (rc != hipSuccess) return (int)rc;
  *atomics_per_second = ms > 0.0f ? n / (ms * 1e-3) : 0.0;
  if (milliseconds) *milliseconds = ms;
  return 0;
}
