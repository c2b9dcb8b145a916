// Seeded list-mode events for the strip scanner from a phantom of ellipses,
// rectangles and points, on the host (threads) or on the device (one lane per
// event slot); both run the slot functions of phantom.h, so the same seed
// gives the same events.  Also the per-emission outputs of the phantom
// command (exact responses, emission events, the emitted / detected / naive
// images, the tan-bins map) and its analytic-kernel map.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "../../include/petstrip.h"
#include "phantom.h"
#include "strip_geometry.h"
#include "strip_kernels.h"

namespace petstrip {
int set_error(int code, const char* msg);
StripConst scanner_const(const petstrip_scanner& sc, double sigma_z, double sigma_dl);

namespace {

// Regions of seven doubles (kind, x, y, a, b, angle_deg, intensity) into the
// spec; weights as the reference's (phantom.h:44-67): intensity x area for
// ellipses (pi a b) and rectangles (4 a b), the intensity for points.
int spec_from(const petstrip_scanner* sc, const double* regions, int32_t n_regions,
              int32_t flags, uint64_t seed, PhantomSpec& p) {
  if (!sc || n_regions <= 0 || !regions)
    return set_error(PETSTRIP_EINVAL, "bad phantom arguments");
  if (n_regions > kPhantomMaxRegions)
    return set_error(PETSTRIP_EINVAL, "phantom has more than 64 regions");
  double total = 0;
  p.n_regions = n_regions;
  p.additive = (flags & PETSTRIP_PHANTOM_ADDITIVE) ? 1 : 0;
  for (int i = 0; i < n_regions; ++i) {
    const double* r = regions + 7 * i;
    PhantomRegion& g = p.reg[i];
    g.kind = (int)r[0];
    if (g.kind != kRegionEllipse && g.kind != kRegionRectangle && g.kind != kRegionPoint)
      return set_error(PETSTRIP_EINVAL, "unknown phantom region kind");
    g.x = r[1];
    g.y = r[2];
    g.a = r[3];
    g.b = r[4];
    if (!(r[6] >= 0)) return set_error(PETSTRIP_EINVAL, "phantom intensity must be >= 0");
    if (g.kind != kRegionPoint && (!(g.a > 0) || !(g.b > 0)))
      return set_error(PETSTRIP_EINVAL, "phantom region needs positive axes");
    const double ang = g.kind == kRegionEllipse ? r[5] * M_PI / 180 : 0.0;
    g.s = sin(ang);
    g.c = cos(ang);
    if (g.kind == kRegionEllipse) {
      g.A = g.c * g.c / (g.a * g.a) + g.s * g.s / (g.b * g.b);
      g.B = g.s * g.s / (g.a * g.a) + g.c * g.c / (g.b * g.b);
      g.C = g.s * g.c * (1 / (g.a * g.a) - 1 / (g.b * g.b));
      total += r[6] * M_PI * g.a * g.b;
    } else {
      g.A = g.B = g.C = 0;
      total += g.kind == kRegionRectangle ? r[6] * 4 * g.a * g.b : r[6];
    }
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

// Ellipse sextuples (x, y, a, b, angle_deg, intensity) as region septuples.
std::vector<double> ellipses7(const double* regions, int32_t n_regions) {
  std::vector<double> r7((size_t)std::max(n_regions, 0) * 7);
  for (int i = 0; i < n_regions; ++i) {
    r7[7 * i] = kRegionEllipse;
    for (int k = 0; k < 6; ++k) r7[7 * i + 1 + k] = regions[6 * i + k];
  }
  return r7;
}

// One slot of the phantom command: returns 1 with an event (responses with
// and without error, the emission's point and direction (sin, cos phi)), 0
// for an undetected emission (emission mode), -1 when the draws ran out.
// Histogram increments go through hist (atomics on the device, a thread's
// own images on the host).
template <class Hist>
PH_HD int phantom_slot_full(const PhantomSpec& p, int64_t i, bool detected_mode,
                            const PhantomGrid& g, Hist& hist, float resp[3],
                            float exact_out[3], float event_out[4],
                            uint32_t* n_emitted) {
  uint32_t ctr[4] = {(uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0, 0};
  double px, py, phi, exact[3];
  uint32_t ne = 0;
  for (uint32_t draw = 0; ph_next_emission(p, ctr, &draw, &px, &py, &phi); ++draw) {
    ++ne;
    const int pix = ph_pixel(g, px, py);
    hist.emitted(pix);
    if (ph_detect(p, px, py, phi, exact)) {
      ph_noise(p, ctr, draw, exact, resp);
      hist.detected(pix);
      // centre from the response with error (response.h:22-37, float)
      const float zu = resp[0], zd = resp[1], dl = resp[2];
      const float t = (zu - zd) / (2.0f * (float)p.R);
      const float y = -0.5f * dl / sqrtf(1.0f + t * t);
      const float z = 0.5f * (zu + zd + 2.0f * y * t);
      const int npix = ph_pixel(g, z, y);
      hist.naive(npix);
      if (g.tan_bins > 0 && npix >= 0) {
        const double plane = floor((t + g.max_tan) / (2 * g.max_tan / g.tan_bins));
        if (plane >= 0 && plane < g.tan_bins)
          hist.tan(npix + (int64_t)plane * g.n_columns * g.n_rows);
      }
      exact_out[0] = (float)exact[0];
      exact_out[1] = (float)exact[1];
      exact_out[2] = (float)exact[2];
      event_out[0] = (float)px;
      event_out[1] = (float)py;
      event_out[2] = (float)sin(phi);
      event_out[3] = (float)cos(phi);
      *n_emitted = ne;
      return 1;
    }
    if (!detected_mode) {
      *n_emitted = ne;
      return 0;
    }
  }
  *n_emitted = ne;
  return -1;
}

struct DevHist {
  int32_t *em, *det, *nai, *tanb;
  __device__ void emitted(int i) { if (em && i >= 0) atomicAdd(em + i, 1); }
  __device__ void detected(int i) { if (det && i >= 0) atomicAdd(det + i, 1); }
  __device__ void naive(int i) { if (nai && i >= 0) atomicAdd(nai + i, 1); }
  __device__ void tan(int64_t i) { if (tanb) atomicAdd(tanb + i, 1); }
};

struct HostHist {
  int32_t *em, *det, *nai, *tanb;
  void emitted(int i) { if (em && i >= 0) ++em[i]; }
  void detected(int i) { if (det && i >= 0) ++det[i]; }
  void naive(int i) { if (nai && i >= 0) ++nai[i]; }
  void tan(int64_t i) { if (tanb) ++tanb[i]; }
};

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

// Slots [s0, s0 + m) of the phantom command into chunk buffers.
__global__ __launch_bounds__(256) void k_phantom_full(
    const PhantomSpec* __restrict__ spec, const PhantomGrid* __restrict__ grid,
    int64_t s0, int64_t m, int detected_mode, float* __restrict__ resp,
    float* __restrict__ exact, float* __restrict__ events, int8_t* __restrict__ flag,
    DevHist hist, unsigned long long* __restrict__ n_emitted) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  float r[3] = {0, 0, 0}, x[3] = {0, 0, 0}, e[4] = {0, 0, 0, 0};
  uint32_t ne = 0;
  const int f = phantom_slot_full(*spec, s0 + j, detected_mode != 0, *grid, hist, r, x, e, &ne);
  for (int k = 0; k < 3; ++k) {
    resp[3 * j + k] = r[k];
    exact[3 * j + k] = x[k];
  }
  for (int k = 0; k < 4; ++k) events[4 * j + k] = e[k];
  flag[j] = (int8_t)f;
  atomicAdd(n_emitted, (unsigned long long)ne);
}

int herr(hipError_t e, const char* what) {
  return set_error(e == hipErrorOutOfMemory ? PETSTRIP_ENOMEM : PETSTRIP_EDEVICE,
                   (std::string(what) + ": " + hipGetErrorString(e)).c_str());
}

}  // namespace

// Region table and scanner constants of an ellipse phantom; PETSTRIP_OK or
// an error.
int phantom_spec(const petstrip_scanner* sc, const double* regions, int32_t n_regions,
                 uint64_t seed, PhantomSpec& p) {
  if (!regions || n_regions <= 0) return set_error(PETSTRIP_EINVAL, "bad phantom arguments");
  if (n_regions > kPhantomMaxRegions)
    return set_error(PETSTRIP_EINVAL, "phantom has more than 64 regions");
  const std::vector<double> r7 = ellipses7(regions, n_regions);
  return spec_from(sc, r7.data(), n_regions, 0, seed, p);
}

// Device generation into out (3 floats per event) on stream st; *failed
// set when a slot found no detected emission.
int phantom_device(const PhantomSpec& p, int64_t n, float* out, hipStream_t st,
                   bool* failed) {
  *failed = false;
  if (n <= 0) return PETSTRIP_OK;
  if (n > (int64_t)INT32_MAX) return set_error(PETSTRIP_EINVAL, "more than 2^31-1 events");
  PhantomSpec* d_spec = nullptr;
  int* d_fail = nullptr;
  auto done = [&](int rc) {
    if (d_spec) (void)hipFree(d_spec);
    if (d_fail) (void)hipFree(d_fail);
    return rc;
  };
  hipError_t e;
  if ((e = hipMalloc((void**)&d_spec, sizeof(PhantomSpec))) != hipSuccess)
    return done(herr(e, "hipMalloc"));
  if ((e = hipMalloc((void**)&d_fail, sizeof(int))) != hipSuccess)
    return done(herr(e, "hipMalloc"));
  if ((e = hipMemcpyAsync(d_spec, &p, sizeof(PhantomSpec), hipMemcpyHostToDevice, st)) !=
      hipSuccess)
    return done(herr(e, "hipMemcpyAsync"));
  if ((e = hipMemsetAsync(d_fail, 0, sizeof(int), st)) != hipSuccess)
    return done(herr(e, "hipMemsetAsync"));
  hipLaunchKernelGGL(k_phantom, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     d_spec, n, out, d_fail);
  if ((e = hipGetLastError()) != hipSuccess) return done(herr(e, "k_phantom"));
  int f = 0;
  if ((e = hipMemcpyAsync(&f, d_fail, sizeof(int), hipMemcpyDeviceToHost, st)) != hipSuccess)
    return done(herr(e, "hipMemcpyAsync"));
  if ((e = hipStreamSynchronize(st)) != hipSuccess)
    return done(herr(e, "hipStreamSynchronize"));
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
  if ((e = hipSetDevice(device)) != hipSuccess) return herr(e, "hipSetDevice");
  bool failed = false;
  if ((rc = phantom_device(p, n, d_out, nullptr, &failed))) return rc;
  if (failed)
    return set_error(PETSTRIP_ESTATE, "no detected emission after 2^16 draws in a slot");
  return PETSTRIP_OK;
}

extern "C" int petstrip_phantom_generate(const petstrip_scanner* sc, const double* regions,
                                         int32_t n_regions, int32_t flags, uint64_t seed,
                                         int64_t n, const petstrip_phantom_grid* grid,
                                         petstrip_phantom_output* out, int device) {
  using namespace petstrip;
  if (!out || n < 0 || (n > 0 && !out->responses))
    return set_error(PETSTRIP_EINVAL, "bad phantom arguments");
  if (flags & ~(PETSTRIP_PHANTOM_ADDITIVE | PETSTRIP_PHANTOM_DETECTED))
    return set_error(PETSTRIP_EINVAL, "unknown phantom flags");
  PhantomSpec p;
  int rc;
  if ((rc = spec_from(sc, regions, n_regions, flags, seed, p))) return rc;
  PhantomGrid g{1, 1, 1.0, 0.0, 0.0, 0, 1.0};
  const bool images = out->emitted || out->detected || out->naive || out->tan_bins_map;
  if (images) {
    if (!grid || grid->n_columns < 1 || grid->n_rows < 1 || !(grid->pixel_size > 0) ||
        grid->tan_bins < 0 || (out->tan_bins_map && (grid->tan_bins < 1 || !(grid->max_tan > 0))))
      return set_error(PETSTRIP_EINVAL, "bad phantom image grid");
    g = PhantomGrid{grid->n_columns, grid->n_rows, grid->pixel_size, grid->lower_left_x,
                    grid->lower_left_y, out->tan_bins_map ? grid->tan_bins : 0, grid->max_tan};
  }
  const size_t npix = (size_t)g.n_columns * g.n_rows;
  const size_t ntan = npix * (size_t)std::max(g.tan_bins, 1);
  const bool detected_mode = (flags & PETSTRIP_PHANTOM_DETECTED) != 0;
  for (int32_t* im : {out->emitted, out->detected, out->naive})
    if (im) std::fill(im, im + npix, 0);
  if (out->tan_bins_map) std::fill(out->tan_bins_map, out->tan_bins_map + ntan, 0);
  out->n_responses = 0;
  out->n_emitted = 0;
  const int64_t kChunk = (int64_t)1 << 22;
  // compacts a chunk's slots into the caller's arrays
  auto gather = [&](int64_t m, const int8_t* flag, const float* r, const float* x,
                    const float* e) -> int {
    for (int64_t j = 0; j < m; ++j) {
      if (flag[j] < 0)
        return set_error(PETSTRIP_ESTATE, "no detected emission after 2^16 draws in a slot");
      if (flag[j] == 0) continue;
      const int64_t o = out->n_responses++;
      std::copy(r + 3 * j, r + 3 * j + 3, out->responses + 3 * o);
      if (out->exact) std::copy(x + 3 * j, x + 3 * j + 3, out->exact + 3 * o);
      if (out->events) std::copy(e + 4 * j, e + 4 * j + 4, out->events + 4 * o);
    }
    return PETSTRIP_OK;
  };
  if (device < 0) {
    // host threads, each with its own images, summed at the end
    const int nt = (int)std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency()));
    std::vector<float> r(3 * kChunk), x(3 * kChunk), e(4 * kChunk);
    std::vector<int8_t> flag(kChunk);
    std::vector<std::vector<int32_t>> im((size_t)nt * 4);
    std::vector<uint64_t> ne((size_t)nt, 0);
    for (int t = 0; t < nt; ++t) {
      if (out->emitted) im[4 * t].assign(npix, 0);
      if (out->detected) im[4 * t + 1].assign(npix, 0);
      if (out->naive) im[4 * t + 2].assign(npix, 0);
      if (out->tan_bins_map) im[4 * t + 3].assign(ntan, 0);
    }
    for (int64_t s0 = 0; s0 < n; s0 += kChunk) {
      const int64_t m = std::min(kChunk, n - s0);
      auto work = [&](int t) {
        HostHist h{out->emitted ? im[4 * t].data() : nullptr,
                   out->detected ? im[4 * t + 1].data() : nullptr,
                   out->naive ? im[4 * t + 2].data() : nullptr,
                   out->tan_bins_map ? im[4 * t + 3].data() : nullptr};
        for (int64_t j = m * t / nt; j < m * (t + 1) / nt; ++j) {
          uint32_t k = 0;
          flag[j] = (int8_t)phantom_slot_full(p, s0 + j, detected_mode, g, h, &r[3 * j],
                                              &x[3 * j], &e[4 * j], &k);
          ne[t] += k;
        }
      };
      std::vector<std::thread> th;
      for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
      work(0);
      for (auto& t : th) t.join();
      if ((rc = gather(m, flag.data(), r.data(), x.data(), e.data()))) return rc;
    }
    int32_t* dst[4] = {out->emitted, out->detected, out->naive, out->tan_bins_map};
    for (int t = 0; t < nt; ++t) {
      out->n_emitted += (int64_t)ne[t];
      for (int k = 0; k < 4; ++k)
        if (dst[k])
          for (size_t i = 0; i < im[4 * t + k].size(); ++i) dst[k][i] += im[4 * t + k][i];
    }
    return PETSTRIP_OK;
  }
  // device: chunks of slots, compacted on the host
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return herr(e, "hipSetDevice");
  std::vector<void*> dev;
  auto dalloc = [&](size_t bytes, void** ptr) -> hipError_t {
    hipError_t err = hipMalloc(ptr, bytes);
    if (err == hipSuccess) dev.push_back(*ptr);
    return err;
  };
  auto done = [&](int code) {
    for (void* ptr : dev) (void)hipFree(ptr);
    return code;
  };
  PhantomSpec* d_spec;
  PhantomGrid* d_grid;
  float *d_r, *d_x, *d_e;
  int8_t* d_flag;
  unsigned long long* d_ne;
  DevHist h{nullptr, nullptr, nullptr, nullptr};
  if ((e = dalloc(sizeof(PhantomSpec), (void**)&d_spec)) ||
      (e = dalloc(sizeof(PhantomGrid), (void**)&d_grid)) ||
      (e = dalloc(12 * kChunk, (void**)&d_r)) || (e = dalloc(12 * kChunk, (void**)&d_x)) ||
      (e = dalloc(16 * kChunk, (void**)&d_e)) || (e = dalloc(kChunk, (void**)&d_flag)) ||
      (e = dalloc(sizeof(unsigned long long), (void**)&d_ne)) ||
      (out->emitted && (e = dalloc(4 * npix, (void**)&h.em))) ||
      (out->detected && (e = dalloc(4 * npix, (void**)&h.det))) ||
      (out->naive && (e = dalloc(4 * npix, (void**)&h.nai))) ||
      (out->tan_bins_map && (e = dalloc(4 * ntan, (void**)&h.tanb))))
    return done(herr(e, "hipMalloc"));
  if ((e = hipMemcpy(d_spec, &p, sizeof p, hipMemcpyHostToDevice)) ||
      (e = hipMemcpy(d_grid, &g, sizeof g, hipMemcpyHostToDevice)) ||
      (e = hipMemset(d_ne, 0, sizeof(unsigned long long))) ||
      (h.em && (e = hipMemset(h.em, 0, 4 * npix))) ||
      (h.det && (e = hipMemset(h.det, 0, 4 * npix))) ||
      (h.nai && (e = hipMemset(h.nai, 0, 4 * npix))) ||
      (h.tanb && (e = hipMemset(h.tanb, 0, 4 * ntan))))
    return done(herr(e, "hipMemcpy"));
  std::vector<float> r(3 * kChunk), x(3 * kChunk), ev(4 * kChunk);
  std::vector<int8_t> flag(kChunk);
  for (int64_t s0 = 0; s0 < n; s0 += kChunk) {
    const int64_t m = std::min(kChunk, n - s0);
    hipLaunchKernelGGL(k_phantom_full, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, 0,
                       d_spec, d_grid, s0, m, detected_mode ? 1 : 0, d_r, d_x, d_e, d_flag,
                       h, d_ne);
    if ((e = hipGetLastError()) != hipSuccess) return done(herr(e, "k_phantom_full"));
    if ((e = hipMemcpy(flag.data(), d_flag, m, hipMemcpyDeviceToHost)) ||
        (e = hipMemcpy(r.data(), d_r, 12 * m, hipMemcpyDeviceToHost)) ||
        (out->exact && (e = hipMemcpy(x.data(), d_x, 12 * m, hipMemcpyDeviceToHost))) ||
        (out->events && (e = hipMemcpy(ev.data(), d_e, 16 * m, hipMemcpyDeviceToHost))))
      return done(herr(e, "hipMemcpy"));
    if ((rc = gather(m, flag.data(), r.data(), x.data(), ev.data()))) return done(rc);
  }
  unsigned long long total = 0;
  if ((e = hipMemcpy(&total, d_ne, sizeof total, hipMemcpyDeviceToHost)) ||
      (h.em && (e = hipMemcpy(out->emitted, h.em, 4 * npix, hipMemcpyDeviceToHost))) ||
      (h.det && (e = hipMemcpy(out->detected, h.det, 4 * npix, hipMemcpyDeviceToHost))) ||
      (h.nai && (e = hipMemcpy(out->naive, h.nai, 4 * npix, hipMemcpyDeviceToHost))) ||
      (h.tanb && (e = hipMemcpy(out->tan_bins_map, h.tanb, 4 * ntan, hipMemcpyDeviceToHost))))
    return done(herr(e, "hipMemcpy"));
  out->n_emitted = (int64_t)total;
  return done(PETSTRIP_OK);
}

extern "C" int petstrip_phantom_tan_kernel_map(const petstrip_scanner* sc,
                                               const petstrip_phantom_grid* grid,
                                               double kernel_point_x, double kernel_point_y,
                                               double scale, float* out, int device) {
  using namespace petstrip;
  if (!sc || !grid || !out || grid->n_columns < 1 || grid->n_rows < 1 ||
      grid->tan_bins < 1 || !(grid->pixel_size > 0) || !(grid->max_tan > 0))
    return set_error(PETSTRIP_EINVAL, "bad tan kernel map arguments");
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return herr(e, "hipSetDevice");
  const StripConst k = scanner_const(*sc, sc->sigma_z, sc->sigma_dl);
  const size_t n = (size_t)grid->n_columns * grid->n_rows * grid->tan_bins;
  float* d = nullptr;
  if ((e = hipMalloc((void**)&d, n * sizeof(float))) != hipSuccess) return herr(e, "hipMalloc");
  e = launch_tan_kernel_map(nullptr, k, grid->n_columns, grid->n_rows, grid->tan_bins,
                            (float)grid->pixel_size, (float)grid->lower_left_x,
                            (float)grid->lower_left_y, (float)grid->max_tan,
                            (float)kernel_point_x, (float)kernel_point_y, (float)scale, d);
  if (e == hipSuccess) e = hipMemcpy(out, d, n * sizeof(float), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return herr(e, "tan kernel map");
  return PETSTRIP_OK;
}
