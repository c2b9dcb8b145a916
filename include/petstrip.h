/*
 * petstrip: MI355X (gfx950) list-mode MLEM for the 2D strip PET scanner.
 *
 * C ABI of libpetstrip_hip.so.  It replaces the reference's --gpu entry
 * point, the C++ template
 *   PET2D::Strip::GPU::Reconstruction::run<F>(scanner, responses, n,
 *       input_output, n_iteration_blocks, n_iterations_in_block, output,
 *       progress, device, n_blocks, n_threads_per_block, device_name)
 * declared in src/2d/strip/cuda/reconstruction.h:31-44 and defined in
 * src/2d/strip/cuda/reconstruction.cu:31-113 (grey277/j-pet-mlem).
 *
 * Conventions
 *   - Every call returns 0 on success or a negative PETSTRIP_E* code; the
 *     message of the last failure on the calling thread is returned by
 *     petstrip_last_error().  Nothing aborts the process (the reference
 *     abort()s on any CUDA error, src/util/cuda/debug.h:17-27).
 *   - Host buffers belong to the caller; device buffers belong to the handle.
 *   - Images are float32, W = n_z_pixels columns by H = n_y_pixels rows,
 *     row-major, row 0 on top (y = +H*pixel_height/2), column 0 on the left
 *     (z = -W*pixel_width/2), as PixelMap and Scanner use them
 *     (src/2d/geometry/pixel_map.h:15-40, src/2d/strip/scanner.h:61-66).
 *   - Responses are packed float32 triples (z_u, z_d, dl)
 *     (src/2d/strip/response.h:13-45).
 *   - One handle drives one device on one HIP stream; calls on one handle
 *     must not overlap.  Callbacks run on the calling thread.
 */
#ifndef PETSTRIP_H
#define PETSTRIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PETSTRIP_ABI_VERSION 1

enum {
  PETSTRIP_OK = 0,
  PETSTRIP_EINVAL = -1,   /* bad argument (null pointer, bad size, ...) */
  PETSTRIP_EDEVICE = -2,  /* no such device / HIP runtime failure       */
  PETSTRIP_ENOMEM = -3,   /* device or host allocation failed           */
  PETSTRIP_ESTATE = -4,   /* call out of order (e.g. no events loaded)  */
  PETSTRIP_ECALLBACK = -5 /* a callback asked to stop                   */
};

/* Scanner geometry, the fields of the Scanner constructor in its order
 * (src/2d/strip/scanner.h:39-49).  Values are converted to float exactly as
 * the reference's Scanner<float, short> stores them. */
typedef struct petstrip_scanner {
  double radius;              /* R: the strips lie at y = +-R [m]       */
  double scintillator_length; /* L [m]                                  */
  int32_t n_y_pixels;         /* H                                      */
  int32_t n_z_pixels;         /* W                                      */
  double pixel_height;        /* [m]                                    */
  double pixel_width;         /* [m]                                    */
  double sigma_z;             /* TOF sigma along z [m]                  */
  double sigma_dl;            /* TOF sigma of delta-l [m]               */
  double center_y;            /* scanner centre, usually 0              */
  double center_z;
} petstrip_scanner;

/* How the backprojection is accumulated. */
enum {
  PETSTRIP_ACCUM_FP32 = 0,        /* plain float sums                       */
  PETSTRIP_ACCUM_COMPENSATED = 1, /* compensated (two-sum) float sums       */
  PETSTRIP_ACCUM_FIXED64 = 2      /* int64 fixed point, 2^-40 units:
                                     bit-reproducible run to run          */
};

/* Which per-event update is computed. */
enum {
  /* CPU semantics (the parity target): Gaussian kernel, no sensitivity,
   * rho'[p] = sum_e K_e(p) rho[p] / sum_q K_e(q) rho[q]
   * (src/2d/strip/reconstruction.h:220-298). */
  PETSTRIP_UPDATE_CPU = 0,
  /* The reference --gpu update: denominator weighted by the 5-point pixel
   * sensitivity (src/2d/strip/cuda/reconstruction/warp_granularity.cuh:
   * 101-108,145-150), guarded inverse; with PETSTRIP_KERNEL_ANALYTIC it is
   * the reference --gpu iteration (cuda/reconstruction.cu:45-52). */
  PETSTRIP_UPDATE_SENSITIVITY_DENOMINATOR = 1
};

/* Which kernel value is computed inside the (Gaussian) 3-sigma support. */
enum {
  PETSTRIP_KERNEL_GAUSSIAN = 0,  /* GaussianKernel (gaussian_kernel.h:51-79)  */
  /* AnalyticKernel (src/2d/strip/analytic_kernel.h:28-62); runs on the
   * per-event kernels; no sigma derivative */
  PETSTRIP_KERNEL_ANALYTIC = 1
};

/* Which kernels run an iteration. */
enum {
  PETSTRIP_PATH_AUTO = 0,      /* fused binned kernel when it applies       */
  PETSTRIP_PATH_PER_EVENT = 1  /* one wave per event, global atomics        */
};

typedef struct petstrip_options {
  int32_t accumulation; /* PETSTRIP_ACCUM_*                                */
  int32_t update;       /* PETSTRIP_UPDATE_*                               */
  int32_t path;         /* PETSTRIP_PATH_*                                 */
  int32_t row_scan;     /* verification switch: 1 = find every row's ellipse
                           interval by scanning instead of from the roots of
                           its quadratic; results are identical, slower     */
  int32_t force_bands;  /* verification switch: 1 = run the fused passes in
                           64-row bands (the form used when a window does
                           not fit LDS) even when windows fit; 2 = in bands
                           of 64 rows x 128 columns (the form used when a
                           row band does not fit either)                   */
  int32_t kernel;       /* PETSTRIP_KERNEL_*                               */
  int32_t row_table;    /* 0 = the binned walks keep each event's box rows
                           (ellipse columns, kernel row factor) in a device
                           table built once per event set when it fits in
                           a quarter of free device memory; 1 = no table,
                           every iteration forms the rows again (identical
                           results)                                      */
  int32_t pass1_walk;   /* 0 = the binned iterations sum each row's
                           denominator terms as polynomials in the kernel's
                           geometric ratio by Horner's rule; 1 = by
                           advancing a chain of kernel values (results agree
                           to FP32 rounding; the compensated accumulation
                           and the sigma derivative always use the chain) */
} petstrip_options;

typedef struct petstrip_handle petstrip_handle;

/* Per-run counters of the last iteration. */
typedef struct petstrip_stats {
  int64_t n_events;          /* events on the device                     */
  int64_t n_box_pixels;      /* sum over events of clamped-box pixels    */
  int64_t n_ellipse_pixels;  /* sum over events of 3-sigma-ellipse pixels*/
  int64_t n_zero_denominator;/* events whose denominator was not > 0     */
  double flop_alg;           /* sum_e 15 N_box + 41 N_ell + 50           */
  double last_iteration_ms;  /* device time of the last iteration        */
  double last_kernel_ms;     /* device time of its event kernel(s)       */
  int32_t fused;             /* 1: iterations use the binned kernel, 2: in
                                64-row bands, 3: in bands of 64 rows x 128
                                columns, 0: the per-event kernels       */
  int32_t n_work_items;      /* workgroups of the binned kernel          */
  int64_t lds_bytes;         /* LDS per workgroup of the binned kernel   */
  int64_t n_window_pixels;   /* binned kernel: sum over work items of their
                                window pixels, the most global atomics one
                                pass-2 launch makes (one per nonzero window
                                pixel); every ellipse pixel's update is an
                                LDS atomic.  0 for the per-event kernels,
                                whose updates are all global atomics     */
  int64_t row_table_bytes;   /* device bytes of the row table (0: none)   */
  int64_t horner_walks;      /* 1: the binned iterations' denominators by
                                Horner's rule (options pass1_walk), 0: by
                                the chain                                 */
  int64_t n_items_two_per_cu; /* unbanded binned kernel: work items whose
                                 windows (32-bit sums, the FP32 iteration)
                                 fit the LDS of two workgroups per CU; the
                                 events bin by box width as well when the
                                 widest window does not                  */
} petstrip_stats;

/* Called after an output is ready.  iteration >= 1 for images, -1 for the
 * sensitivity map (as the reference's output delegate,
 * cuda/reconstruction.cu:63-70,98-110).  image is W*H floats, valid only
 * during the call.  Return nonzero to stop (run returns
 * PETSTRIP_ECALLBACK). */
typedef int (*petstrip_output_cb)(void* user, int iteration,
                                  const float* image, int32_t width,
                                  int32_t height);
/* Progress, as the reference's progress delegate (completed, finished). */
typedef void (*petstrip_progress_cb)(void* user, int completed, int finished);

/* --- lifetime ---------------------------------------------------------- */

/* Returns PETSTRIP_ABI_VERSION. */
int petstrip_abi_version(void);
/* Message of the last failure on this thread ("" if none). */
const char* petstrip_last_error(void);
/* Number of visible HIP devices (0 without a GPU; no error). */
int petstrip_device_count(int* count);

/* Creates a handle on `device` (cudaSetDevice + properties,
 * cuda/reconstruction.cu:54-57).  opts may be NULL for defaults. */
int petstrip_create(const petstrip_scanner* scanner, int device,
                    const petstrip_options* opts, petstrip_handle** out);
int petstrip_destroy(petstrip_handle* h);
/* Device marketing name (reference: device_name delegate). */
int petstrip_device_name(petstrip_handle* h, char* buf, size_t len);
/* The HIP stream (hipStream_t) the handle launches on. */
int petstrip_stream(petstrip_handle* h, void** stream);
int petstrip_set_options(petstrip_handle* h, const petstrip_options* opts);

/* --- data ---------------------------------------------------------------- */

/* Uploads n responses (z_u, z_d, dl) from host memory (replaces the
 * ResponsesSOA upload, src/2d/strip/cuda/gpu_events_soa.h:15-44) and
 * prepares per-event geometry.  n may be 0. */
int petstrip_set_events(petstrip_handle* h, const float* responses,
                        int64_t n);
/* Same from device memory already holding n packed triples. */
int petstrip_set_events_device(petstrip_handle* h,
                               const float* d_responses, int64_t n);
/* Sets / reads the current image rho (W*H floats, host memory). */
int petstrip_set_rho(petstrip_handle* h, const float* rho);
int petstrip_get_rho(petstrip_handle* h, float* rho);
/* Sets rho from device memory. */
int petstrip_set_rho_device(petstrip_handle* h, const float* d_rho);
/* Device pointer of the current image (valid until the next call). */
int petstrip_rho_device(petstrip_handle* h, const float** d_rho);

/* --- MLEM ---------------------------------------------------------------- */

/* Runs n_iterations MLEM iterations on the device-resident image and
 * events, then waits for the device. */
int petstrip_iterate(petstrip_handle* h, int n_iterations);

/* The reference run(): rho_inout is the start image and receives every
 * output; images are reported through `output` at iterations 1-5, 10, 15,
 * 20, 30, 50, 100 of the first block and at the end of every block, and the
 * 5-point sensitivity map first with iteration -1
 * (cuda/reconstruction.cu:63-70,79-110).  Callbacks may be NULL. */
int petstrip_run(petstrip_handle* h, float* rho_inout, int n_iteration_blocks,
                 int n_iterations_in_block, petstrip_output_cb output,
                 petstrip_progress_cb progress, void* user);

/* Sensitivity map: mode 0 = Scanner::sensitivity at pixel centres
 * (src/2d/strip/reconstruction.h:67-73), mode 1 = 5-point
 * Scanner::pixel_sensitivity (src/2d/strip/scanner.h:174-185). */
int petstrip_sensitivity(petstrip_handle* h, int mode, float* out);

int petstrip_stats_get(petstrip_handle* h, petstrip_stats* out);
/* Per-event footprints {box pixels, ellipse pixels} (host arrays of n). */
int petstrip_footprints(petstrip_handle* h, int32_t* n_box, int32_t* n_ell);

/* --- differentiation ----------------------------------------------------- */

/* Forward projector: lambda[e] = sum_{p in E_e} K_e(p) rho[p] for the
 * current image (the MLEM denominator, reconstruction.h:264-290).  lambda
 * is a host array of n floats (may be NULL: then only kept on device). */
int petstrip_forward(petstrip_handle* h, float* lambda);

/* Adjoint of the forward projector and the sigma derivative.  With upstream
 * weights g_e (host array of n, or NULL for g_e = 1/lambda_e, the Poisson
 * list-mode log-likelihood l = sum_e log lambda_e - sum_p s_p rho_p with
 * s_p = 1):
 *   grad_rho[p]   = sum_e g_e K_e(p)  - (g == NULL ? 1 : 0)
 *   grad_sigma[0] = sum_e g_e d lambda_e / d sigma_z
 *   grad_sigma[1] = sum_e g_e d lambda_e / d sigma_dl
 * with the support (ellipse and box) held at the current sigma.  Kernels
 * are recomputed, not stored.  loglik (nullable) receives l.  Uses the
 * lambda of the last petstrip_forward (computes it if absent).  On the
 * fused path grad_sigma and loglik are summed per workgroup and then in a
 * fixed order, so they are bit-reproducible run to run. */
int petstrip_backward(petstrip_handle* h, const float* g, float* grad_rho,
                      double grad_sigma[2], double* loglik);

/* Changes (sigma_z, sigma_dl) of the kernel.  freeze_support = 1 keeps the
 * ellipse/box support of the previous sigma (for finite differences). */
int petstrip_set_sigma(petstrip_handle* h, double sigma_z, double sigma_dl,
                       int freeze_support);

/* --- event generation (input side) --------------------------------------- */

/* Seeded list-mode events from an elliptical phantom, independent of the
 * thread count: slot i draws from Philox4x32-10 keyed by (seed, i) until an
 * emission is detected.  Regions are (x, y, a, b, angle_deg, intensity)
 * sextuples; non-additive: earlier regions lie on top (the reference
 * phantom, src/2d/geometry/phantom.h:183-215, detection and noise
 * src/2d/strip/scanner.h:107-113,195-267).  out receives n triples. */
int petstrip_phantom_events(const petstrip_scanner* scanner,
                            const double* regions, int32_t n_regions,
                            uint64_t seed, int64_t n, float* out,
                            int n_threads);

/* The same events generated on device `device` into d_out (device memory,
 * n triples), one lane per slot.  Host and device runs of one seed agree
 * except where the two libm implementations differ in the last bit of a
 * double.  No counterpart in the reference (its phantom is host-only,
 * src/2d/strip/phantom_cmd.cpp:214-394, with a shared random engine). */
int petstrip_phantom_events_device(const petstrip_scanner* scanner,
                                   const double* regions, int32_t n_regions,
                                   uint64_t seed, int64_t n, float* d_out,
                                   int device);

/* The phantom command's generator (src/2d/strip/phantom_cmd.cpp:136-444,
 * src/common/phantom_monte_carlo.h:33-109).  Regions are septuples
 * (kind, x, y, a, b, angle_deg, intensity) with kind 0 = ellipse (semi-axes
 * a, b, rotated by angle), 1 = rectangle (centre x, y, half-sides a, b),
 * 2 = point (x, y); weights intensity x area, intensity for points
 * (src/2d/geometry/phantom.h:44-67).  Flags: PETSTRIP_PHANTOM_ADDITIVE
 * (regions overlap additively, phantom.h:193-203) and
 * PETSTRIP_PHANTOM_DETECTED (n counts detected events, as --detected;
 * otherwise n counts emissions and only the detected ones are returned).
 * The same Philox stream as petstrip_phantom_events: in the detected mode
 * with ellipses only, the responses equal that function's.  device < 0
 * runs on host threads (the same events). */
#define PETSTRIP_PHANTOM_ADDITIVE 1
#define PETSTRIP_PHANTOM_DETECTED 2

/* Image grid of the histograms: PixelGrid (src/2d/geometry/pixel_grid.h),
 * n_columns along z, row 0 at the bottom (lower_left_y); tan planes of the
 * tan-bins map as VariableVoxelSizeVoxelGrid (src/3d/geometry/voxel_grid.h
 * :78-121) from -max_tan in steps of 2 max_tan / tan_bins. */
typedef struct petstrip_phantom_grid {
  int32_t n_columns, n_rows;
  double pixel_size;
  double lower_left_x, lower_left_y;
  int32_t tan_bins;
  double max_tan;
} petstrip_phantom_grid;

/* Host outputs; nullable ones are skipped.  The images are overwritten. */
typedef struct petstrip_phantom_output {
  float* responses;      /* capacity n x 3: (z_u, z_d, dl) with detector error */
  float* exact;          /* n x 3: the responses without error               */
  float* events;         /* n x 4: emission point and (sin phi, cos phi)     */
  int32_t* emitted;      /* n_columns x n_rows: origins of every emission    */
  int32_t* detected;     /* origins of the detected emissions                */
  int32_t* naive;        /* centres of the responses with error              */
  int32_t* tan_bins_map; /* n_columns x n_rows x tan_bins, those centres by tan */
  int64_t n_responses;   /* out: responses written                           */
  int64_t n_emitted;     /* out: emissions made                              */
} petstrip_phantom_output;

int petstrip_phantom_generate(const petstrip_scanner* scanner, const double* regions,
                              int32_t n_regions, int32_t flags, uint64_t seed,
                              int64_t n, const petstrip_phantom_grid* grid,
                              petstrip_phantom_output* out, int device);

/* The phantom command's --tan-bins kernel map (phantom_cmd.cpp:396-440):
 * per voxel of the grid, the analytic kernel (analytic_kernel.h:28-62) at
 * the kernel point seen from the voxel centre with the plane's tan, times
 * 4 R pixel^2 dtan sqrt(1 + tan^2) and scale (the reference's
 * 1/sensitivity(kernel point) with --detected, else 1), FP32 on device. */
int petstrip_phantom_tan_kernel_map(const petstrip_scanner* scanner,
                                    const petstrip_phantom_grid* grid,
                                    double kernel_point_x, double kernel_point_y,
                                    double scale, float* out, int device);

/* --- one iteration that allocates nothing ----------------------------------- */

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): one MLEM iteration of the CPU update (the one
 * petstrip_iterate computes) on the per-event kernels
 * (PETSTRIP_PATH_PER_EVENT), without a handle, calling no hipMalloc, hipFree
 * or other allocating routine.  Its event geometry, per-event denominators
 * and weights and its FP64 accumulator live in a static device array of
 * PETSTRIP_NOALLOC_SCRATCH_BYTES, which each device that runs this call holds
 * from its first call until the process ends; so 56 n + 8 W H bytes (plus
 * alignment) must fit it: petstrip_noalloc_max_events gives the largest n
 * for an image size, and larger n fails with PETSTRIP_EINVAL before any
 * device work.  d_responses (n packed triples), d_rho and d_rho_out (W*H
 * floats; may be the same buffer) are device memory on the current device;
 * the work is enqueued on `stream` (a hipStream_t, NULL for the default
 * stream) and the call returns without waiting.  Every call on one device
 * shares the scratch array: calls that overlap on one device race. */
#define PETSTRIP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int petstrip_noalloc_max_events(int32_t n_y_pixels, int32_t n_z_pixels,
                                int64_t* max_events);
int petstrip_iterate_noalloc(const petstrip_scanner* scanner, const float* d_responses,
                             int64_t n, const float* d_rho, float* d_rho_out,
                             void* stream);

/* --- response files (input side; host only, no device needed) ------------- */

/* Reads a text response file as the reference commands read `.txt`
 * responses, with its fast text reader: Strip::Reconstruction::
 * fast_load_txt_events (src/2d/strip/reconstruction.h:169-196) and the 3D
 * hybrid Reconstruction's (src/3d/hybrid/reconstruction.h:165-188), both
 * over util::text_parser (src/util/text_parser.h:29-175).  Its floats are
 * formed in FP32 as float(whole) + float(fraction) x 0.1f^digits (the power
 * by repeated float multiplication, the sum in one fused multiply-add),
 * then scaled by the exponent (e > 0 multiplies by 10 x 0.1f^(e-1), as
 * there), then signed; they are not the correctly rounded values of the
 * decimal strings.  Lines end at '\n' (a last line without one is not
 * read); fields are separated by spaces only.
 *   columns = 3: "z_u z_d dl" per line;
 *   columns = 5: "i j z_u z_d dl" (--3d-response, the 3D hybrid command);
 *                i, j go to lors (2 per line, int16) when it is not NULL.
 * responses == NULL: *n_out = the number of lines (the reader's first
 * pass, reconstruction.h:171-173), nothing is parsed.  Otherwise up to
 * `capacity` responses (3 floats each) are written and *n_out is their
 * number.  A line that does not parse fails with "error line: <line>"
 * (the reference prints the same and throws). */
int petstrip_read_text_responses(const char* path, int32_t columns, float* responses,
                                 int16_t* lors, int64_t capacity, int64_t* n_out);
/* The same float reader on n space-separated fields of one string
 * (util::text_parser::operator>> <float>, text_parser.h:140-186); consumed
 * (nullable) receives the characters read. */
int petstrip_parse_text_floats(const char* text, int32_t n, float* out, int32_t* consumed);

/* Copies the handle's n_events responses (z_u, z_d, dl triples) to the host. */
int petstrip_get_events(petstrip_handle* h, float* responses);

/* Generates n phantom events on the handle's device straight into its event
 * buffer (the scanner's sigmas set the noise) and prepares them as
 * petstrip_set_events does. */
int petstrip_set_phantom_events(petstrip_handle* h, const double* regions,
                                int32_t n_regions, uint64_t seed, int64_t n);

#ifdef __cplusplus
}
#endif

#endif
