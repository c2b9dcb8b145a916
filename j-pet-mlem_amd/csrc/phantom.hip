// Seeded list-mode events for the strip scanner from an elliptical phantom,
// on the host (threads) or on the device (one lane per event slot); both run
// phantom_slot (phantom.h), so the same seed gives the same events.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/petstrip.h"
#include "phantom.h"

namespace petstrip {
int set_error(int code, const char* msg);

// Region table and scanner constants of a phantom; PETSTRIP_OK or an error.
int phantom_spec(const petstrip_scanner* sc, const double* regions, int32_t n_regions,
                 uint64_t seed, PhantomSpec& p) {
  if (!sc || n_regions <= 0 || !regions)
    return set_error(PETSTRIP_EINVAL, "bad phantom arguments");
  if (n_regions > kPhantomMaxRegions)
    return set_error(PETSTRIP_EINVAL, "phantom has more than 64 regions");
  double total = 0;
  p.n_regions = n_regions;
  for (int i = 0; i < n_regions; ++i) {
    const double* r = regions + 6 * i;
    PhantomRegion& g = p.reg[i];
    g.x = r[0];
    g.y = r[1];
    g.a = r[2];
    g.b = r[3];
    if (!(g.a > 0) || !(g.b > 0) || !(r[5] >= 0))
      return set_error(PETSTRIP_EINVAL,
                       "phantom region needs positive axes and intensity >= 0");
    const double ang = r[4] * M_PI / 180;
    g.s = sin(ang);
    g.c = cos(ang);
    g.A = g.c * g.c / (g.a * g.a) + g.s * g.s / (g.b * g.b);
    g.B = g.s * g.s / (g.a * g.a) + g.c * g.c / (g.b * g.b);
    g.C = g.s * g.c * (1 / (g.a * g.a) - 1 / (g.b * g.b));
    total += r[5] * M_PI * g.a * g.b;  // weight = intensity * area
    g.cdf = total;
  }
  if (!(total > 0)) return set_error(PETSTRIP_EINVAL, "phantom has zero total intensity");
  for (int i = 0; i < n_regions; ++i) p.reg[i].cdf /= total;
  p.R = sc->radius;
  p.half_L = sc->scintillator_length / 2;
  p.sz = sc->sigma_z;
  p.sdl = sc->sigma_dl;
  p.k0 = (uint32_t)seed;
  p.k1 = (uint32_t)(seed >> 32);
  return PETSTRIP_OK;
}

namespace {

__global__ __launch_bounds__(256) void k_phantom(const PhantomSpec* __restrict__ spec,
                                                 int64_t n, float* __restrict__ out,
                                                 int* __restrict__ fail) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float ev[3];
  if (!phantom_slot(*spec, i, ev)) {
    atomicOr(fail, 1);
    ev[0] = ev[1] = ev[2] = 0.0f;
  }
  out[3 * i] = ev[0];
  out[3 * i + 1] = ev[1];
  out[3 * i + 2] = ev[2];
}

}  // namespace

// Device generation into out (3 floats per event) on stream st; *failed
// set when a slot found no detected emission.
int phantom_device(const PhantomSpec& p, int64_t n, float* out, hipStream_t st,
                   bool* failed) {
  *failed = false;
  if (n <= 0) return PETSTRIP_OK;
  PhantomSpec* d_spec = nullptr;
  int* d_fail = nullptr;
  auto done = [&](int rc) {
    if (d_spec) (void)hipFree(d_spec);
    if (d_fail) (void)hipFree(d_fail);
    return rc;
  };
  auto herr = [&](hipError_t e, const char* what) {
    return done(set_error(e == hipErrorOutOfMemory ? PETSTRIP_ENOMEM : PETSTRIP_EDEVICE,
                          (std::string(what) + ": " + hipGetErrorString(e)).c_str()));
  };
  hipError_t e;
  if ((e = hipMalloc((void**)&d_spec, sizeof(PhantomSpec))) != hipSuccess)
    return herr(e, "hipMalloc");
  if ((e = hipMalloc((void**)&d_fail, sizeof(int))) != hipSuccess)
    return herr(e, "hipMalloc");
  if ((e = hipMemcpyAsync(d_spec, &p, sizeof(PhantomSpec), hipMemcpyHostToDevice, st)) !=
      hipSuccess)
    return herr(e, "hipMemcpyAsync");
  if ((e = hipMemsetAsync(d_fail, 0, sizeof(int), st)) != hipSuccess)
    return herr(e, "hipMemsetAsync");
  hipLaunchKernelGGL(k_phantom, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     d_spec, n, out, d_fail);
  if ((e = hipGetLastError()) != hipSuccess) return herr(e, "k_phantom");
  int f = 0;
  if ((e = hipMemcpyAsync(&f, d_fail, sizeof(int), hipMemcpyDeviceToHost, st)) !=
      hipSuccess)
    return herr(e, "hipMemcpyAsync");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return herr(e, "hipStreamSynchronize");
  *failed = f != 0;
  return done(PETSTRIP_OK);
}

}  // namespace petstrip

extern "C" int petstrip_phantom_events(const petstrip_scanner* sc,
                                       const double* regions, int32_t n_regions,
                                       uint64_t seed, int64_t n, float* out,
                                       int n_threads) {
  using namespace petstrip;
  if (n < 0 || (n > 0 && !out)) return set_error(PETSTRIP_EINVAL, "bad phantom arguments");
  PhantomSpec p;
  int rc;
  if ((rc = phantom_spec(sc, regions, n_regions, seed, p))) return rc;
  if (n_threads < 1) n_threads = 1;
  n_threads = (int)std::min<int64_t>(n_threads, std::max<int64_t>(n / 4096, 1));
  std::vector<int> status((size_t)n_threads, 0);
  auto work = [&](int tid) {
    const int64_t lo = n * tid / n_threads, hi = n * (tid + 1) / n_threads;
    for (int64_t i = lo; i < hi; ++i) {
      if (!phantom_slot(p, i, out + 3 * i)) {
        status[(size_t)tid] = 1;
        return;
      }
    }
  };
  std::vector<std::thread> th;
  for (int t = 1; t < n_threads; ++t) th.emplace_back(work, t);
  work(0);
  for (auto& t : th) t.join();
  for (int s : status)
    if (s)
      return set_error(PETSTRIP_ESTATE,
                       "no detected emission after 2^16 draws in a slot");
  return PETSTRIP_OK;
}

extern "C" int petstrip_phantom_events_device(const petstrip_scanner* sc,
                                              const double* regions, int32_t n_regions,
                                              uint64_t seed, int64_t n, float* d_out,
                                              int device) {
  using namespace petstrip;
  if (n < 0 || (n > 0 && !d_out)) return set_error(PETSTRIP_EINVAL, "bad phantom arguments");
  PhantomSpec p;
  int rc;
  if ((rc = phantom_spec(sc, regions, n_regions, seed, p))) return rc;
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess)
    return set_error(PETSTRIP_EDEVICE, (std::string("hipSetDevice: ") +
                                        hipGetErrorString(e)).c_str());
  bool failed = false;
  if ((rc = phantom_device(p, n, d_out, nullptr, &failed))) return rc;
  if (failed)
    return set_error(PETSTRIP_ESTATE, "no detected emission after 2^16 draws in a slot");
  return PETSTRIP_OK;
}
