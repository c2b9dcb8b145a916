/* libpetstrip_probe.so: atomic-throughput probes for the measurement report.
 *
 * No counterpart in the reference (grey277/j-pet-mlem): the reference's GPU
 * path scatters one global atomic per ellipse pixel
 * (src/2d/strip/cuda/warp_granularity.cuh:145-150); this build's binned
 * kernel makes those updates LDS atomics and flushes one global atomic per
 * window pixel.  The probes measure the device's rate for each kind so that
 * the kernel's atomic counts (petstrip_stats.n_ellipse_pixels and
 * n_window_pixels) can be set against them.  Measurement aid only: nothing in
 * libpetstrip_hip.so depends on it. */
#ifndef PETSTRIP_PROBE_H
#define PETSTRIP_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  PETSTRIP_PROBE_GLOBAL_F32 = 0, /* global atomic add, float (the FP32 flush)  */
  PETSTRIP_PROBE_GLOBAL_F64 = 1, /* global atomic add, double (the FP64 flush) */
  PETSTRIP_PROBE_GLOBAL_U64 = 2, /* global atomic add, uint64 (fixed point)    */
  PETSTRIP_PROBE_LDS_U64 = 3,    /* LDS ds_add_u64 (the walks' updates)        */
  PETSTRIP_PROBE_LDS_U32 = 4,    /* LDS ds_add_u32                             */
  PETSTRIP_PROBE_LDS_F32 = 5     /* LDS ds_add_f32                             */
};

/* Times one launch of `reps` atomics per lane on `device` and writes the
 * device-wide rate (atomics per second) and the launch time.  Global kinds
 * add into a width x height image (width >= 16, height >= 4, each at most
 * 32768) in tiles of 16 columns x 4 rows per wave; the LDS kind ignores the
 * image size.  Returns 0, or a nonzero HIP error code. */
int petstrip_probe_atomics(int device, int kind, int32_t width, int32_t height,
                           int32_t reps, double* atomics_per_second,
                           double* milliseconds);

#ifdef __cplusplus
}
#endif

#endif /* PETSTRIP_PROBE_H */
