// One MLEM iteration that allocates nothing (petstrip_iterate_noalloc).
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc): the
// iteration of petstrip_iterate on the per-event kernels
// (PETSTRIP_PATH_PER_EVENT: event setup, one wave per event for the
// denominator and for the backprojection into an FP64 accumulator, then the
// conversion to the new image), with every buffer beyond the caller's
// inputs and outputs in the file-scope device array `scratch` below.
//
// The array lives in this file's code object, which holds no kernel, so a
// device pays its memory only once this entry point has run there; it then
// keeps it until the process ends.  Its device address is looked up once per
// device with hipGetSymbolAddress (the array's name in host code is not a
// device address) and only that address is handed to kernels.  Every call
// on one device shares the array: calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "petstrip.h"
#include "strip_fused.h"
#include "strip_geometry.h"
#include "strip_kernels.h"

namespace petstrip {
int set_error(int code, const char* msg);
StripConst scanner_const(const petstrip_scanner& sc, double sigma_z, double sigma_dl);
int scanner_check(const petstrip_scanner* sc);
}  // namespace petstrip

namespace {

constexpr size_t kScratchBytes = PETSTRIP_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up once).
int scratch_address(int device, unsigned char** out) {
  using petstrip::set_error;
  if (device < 0 || device >= kMaxDevices)
    return set_error(PETSTRIP_EDEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(PETSTRIP_EDEVICE,
                       (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e))
                           .c_str());
    g_scratch_ptr[device] = ptr;
  }
  *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
  return PETSTRIP_OK;
}

constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" int petstrip_noalloc_max_events(int32_t n_y_pixels, int32_t n_z_pixels,
                                           int64_t* max_events) {
  if (!max_events || n_y_pixels < 1 || n_z_pixels < 1)
    return petstrip::set_error(PETSTRIP_EINVAL, "bad argument");
  const size_t acc = align256((size_t)n_y_pixels * (size_t)n_z_pixels * sizeof(double));
  if (acc > kScratchBytes) {
    *max_events = 0;
    return PETSTRIP_OK;
  }
  // geometry, denominator and weight per event, each array 256-byte aligned
  const size_t per_event = sizeof(petstrip::EventGeom) + 2 * sizeof(float);
  const size_t room = kScratchBytes - acc - 3 * 256;
  *max_events = (int64_t)(room / per_event);
  return PETSTRIP_OK;
}

extern "C" int petstrip_iterate_noalloc(const petstrip_scanner* scanner,
                                        const float* d_responses, int64_t n,
                                        const float* d_rho, float* d_rho_out, void* stream) {
  using petstrip::set_error;
  int rc;
  if ((rc = petstrip::scanner_check(scanner))) return rc;
  if (n < 0 || (n > 0 && !d_responses) || !d_rho || !d_rho_out)
    return set_error(PETSTRIP_EINVAL, "bad argument");
  int64_t max_n = 0;
  if ((rc = petstrip_noalloc_max_events(scanner->n_y_pixels, scanner->n_z_pixels, &max_n)))
    return rc;
  if (n > max_n)
    return set_error(PETSTRIP_EINVAL,
                     ("petstrip_iterate_noalloc: " + std::to_string(n) + " events exceed the " +
                      std::to_string(max_n) + " its " +
                      std::to_string(kScratchBytes >> 20) + " MiB scratch holds at this image size")
                         .c_str());
  int device = 0;
  hipError_t e = hipGetDevice(&device);
  if (e != hipSuccess) return set_error(PETSTRIP_EDEVICE, hipGetErrorString(e));
  unsigned char* base = nullptr;
  if ((rc = scratch_address(device, &base))) return rc;
  const size_t n_pix = (size_t)scanner->n_y_pixels * (size_t)scanner->n_z_pixels;
  // scratch layout: FP64 accumulator, event geometry, denominators, weights
  double* acc = reinterpret_cast<double*>(base);
  size_t off = align256(n_pix * sizeof(double));
  petstrip::EventGeom* geom = reinterpret_cast<petstrip::EventGeom*>(base + off);
  off += align256((size_t)n * sizeof(petstrip::EventGeom));
  float* lambda = reinterpret_cast<float*>(base + off);
  off += align256((size_t)n * sizeof(float));
  float* inv = reinterpret_cast<float*>(base + off);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const petstrip::StripConst sg =
      petstrip::scanner_const(*scanner, scanner->sigma_z, scanner->sigma_dl);
  // the accumulator keeps the last call's contents: clear it (a fill kernel,
  // which graph capture records as a kernel node)
  if ((e = petstrip::launch_fill(st, reinterpret_cast<float*>(acc), 0.0f,
                                 (int64_t)(2 * n_pix))) ==
          hipSuccess &&
      (e = petstrip::launch_event_setup(st, sg, d_responses, n, geom)) == hipSuccess &&
      (e = petstrip::launch_denominator(st, sg, sg, geom, n, d_rho, nullptr, lambda, inv)) ==
          hipSuccess &&
      (e = petstrip::launch_backproject(st, sg, sg, geom, n, d_rho, inv, acc)) == hipSuccess)
    e = petstrip::launch_finalize(st, acc, petstrip::GLOBAL_F64, 1.0, 0.0f, 0, nullptr,
                                  d_rho_out, (int64_t)n_pix);
  if (e != hipSuccess)
    return set_error(PETSTRIP_EDEVICE,
                     (std::string("kernel launch: ") + hipGetErrorString(e)).c_str());
  return PETSTRIP_OK;
}
