// Host driver and C ABI of libpetstrip_hip.so (declared in include/petstrip.h).
//
// Replaces PET2D::Strip::GPU::Reconstruction::run
// (src/2d/strip/cuda/reconstruction.cu:31-113): owns the device buffers and
// one HIP stream, uploads the responses once, runs iterations on
// device-resident data and reports images through callbacks.  Errors are
// returned as codes with a message (the reference aborts).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/petstrip.h"
#include "phantom.h"
#include "strip_fused.h"
#include "strip_geometry.h"
#include "strip_kernels.h"

using petstrip::EventGeom;
using petstrip::StripConst;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                        \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess)                                                    \
      return fail(_e == hipErrorOutOfMemory ? PETSTRIP_ENOMEM               \
                                            : PETSTRIP_EDEVICE,             \
                  std::string(#expr) + ": " + hipGetErrorString(_e));        \
  } while (0)

int hiperr(hipError_t e) {
  if (e == hipSuccess) return PETSTRIP_OK;
  return fail(e == hipErrorOutOfMemory ? PETSTRIP_ENOMEM : PETSTRIP_EDEVICE,
              std::string("kernel launch: ") + hipGetErrorString(e));
}

// Scanner<float, short> and GaussianKernel<float> constants, computed with the
// reference's float operation order (scanner.h:50-67, gaussian_kernel.h:27-33).
StripConst make_const(const petstrip_scanner& sc, double sigma_z,
                      double sigma_dl) {
  StripConst s;
  s.R = (float)sc.radius;
  s.L = (float)sc.scintillator_length;
  s.ph = (float)sc.pixel_height;
  s.pw = (float)sc.pixel_width;
  float cy = (float)sc.center_y, cz = (float)sc.center_z;
  float size_y = (float)(short)sc.n_y_pixels * s.ph;
  float size_z = (float)(short)sc.n_z_pixels * s.pw;
  s.tl_y = cy + size_y / 2;
  s.tl_z = cz - size_z / 2;
  s.tl_y_half_h = s.tl_y - s.ph / 2;
  s.tl_z_half_w = s.tl_z + s.pw / 2;
  s.half_L = s.L / 2;
  float sz = (float)sigma_z, sdl = (float)sigma_dl;
  s.sigma_z = sz;
  s.sigma_dl = sdl;
  s.isz = 1 / (sz * sz);
  s.isdl = 1 / (sdl * sdl);
  s.c = (float)(1 / (2 * M_PI * M_PI) * (double)sqrtf(s.isz * s.isz * s.isdl));
  s.W = sc.n_z_pixels;
  s.H = sc.n_y_pixels;
  return s;
}

int check_scanner(const petstrip_scanner* sc) {
  if (!sc) return fail(PETSTRIP_EINVAL, "scanner is NULL");
  if (sc->n_y_pixels < 2 || sc->n_z_pixels < 2 || sc->n_y_pixels > 32767 ||
      sc->n_z_pixels > 32767)
    return fail(PETSTRIP_EINVAL,
                "n_y_pixels and n_z_pixels must be in [2, 32767]");
  if (!(sc->pixel_height > 0) || !(sc->pixel_width > 0))
    return fail(PETSTRIP_EINVAL, "pixel sizes must be positive");
  if (!(sc->sigma_z > 0) || !(sc->sigma_dl > 0))
    return fail(PETSTRIP_EINVAL, "sigma_z and sigma_dl must be positive");
  if (!(sc->radius > 0) || !(sc->scintillator_length > 0))
    return fail(PETSTRIP_EINVAL, "radius and length must be positive");
  return PETSTRIP_OK;
}

template <typename T>
int dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? PETSTRIP_ENOMEM : PETSTRIP_EDEVICE,
                std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                    " bytes): " + hipGetErrorString(e));
  return PETSTRIP_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace

namespace petstrip {
int set_error(int code, const char* msg) { return fail(code, msg); }
StripConst scanner_const(const petstrip_scanner& sc, double sigma_z, double sigma_dl) {
  return make_const(sc, sigma_z, sigma_dl);
}
int scanner_check(const petstrip_scanner* sc) { return check_scanner(sc); }
int phantom_spec(const petstrip_scanner* sc, const double* regions, int32_t n_regions,
                 uint64_t seed, PhantomSpec& p);
int phantom_device(const PhantomSpec& p, int64_t n, float* out, hipStream_t st,
                   bool* failed);
}  // namespace petstrip

struct petstrip_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evk0 = nullptr, evk1 = nullptr;
  petstrip_scanner scanner{};
  petstrip_options opts{};
  StripConst sg{};  // constants the support (ellipse, box) was built with
  StripConst sk{};  // constants of the kernel value
  std::string name;
  size_t n_pix = 0;
  int64_t n_events = 0;
  size_t cap_events = 0;
  float* d_resp = nullptr;
  EventGeom* d_geom = nullptr;
  float* d_lambda = nullptr;
  float* d_inv = nullptr;
  int32_t* d_nbox = nullptr;
  int32_t* d_nell = nullptr;
  float* d_rho = nullptr;
  float* d_out = nullptr;
  float* d_sens = nullptr;  // 5-point sensitivity for the --gpu variant
  double* d_scalars = nullptr;
  double* d_scal_part = nullptr;  // per-workgroup partials of the scalars
  size_t cap_scal_part = 0;
  bool lambda_valid = false;
  petstrip_stats stats{};
  // spatially binned events for the fused iteration
  size_t lds_limit = 65536;
  int32_t* d_bin = nullptr;
  int32_t* d_count = nullptr;
  EventGeom* d_sgeom = nullptr;
  int32_t* d_sidx = nullptr;
  petstrip::WorkItem* d_items = nullptr;
  size_t cap_items = 0;
  int n_items = 0;
  // the unbanded launches' item ranges (height class x window width class)
  petstrip::ItemRanges ranges;
  int64_t window_px = 0;  // sum over work items of window pixels
  int max_rows = 1;
  size_t lds_bytes = 0;
  int col_stride = 0;  // LDS column stride of the fused windows (0: none fits)
  int max_ww = 1;
  // banded form (windows taller than LDS allows): per-band partial sums
  bool banded = false;
  int nbands = 1;
  int ncb = 1;      // column-band slots per event (1: row bands only)
  int cshift = 30;  // log2 of the column-band width (30: no column bands)
  int64_t n_sorted = 0;    // events with a non-empty box (the binned ones)
  // groups of whole chunks of the banded form: item range and sorted-event
  // range, pipelined over two streams (pass 2 of a group beside pass 1 of the
  // next)
  struct BandGroup {
    int item_begin, item_end;
    int64_t event_begin, event_end;
  };
  std::vector<BandGroup> groups;
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_group[8] = {};
  hipEvent_t ev_join = nullptr;
  float* d_part = nullptr;
  float* d_lam_sorted = nullptr;
  size_t cap_part = 0;
  // two words on the device: the events whose last denominator was not > 0,
  // then the overflow flag of the fixed-point image (an int); read back
  // together into h_status (pinned) at the end of an iterate call
  unsigned long long* d_nzero = nullptr;
  int* d_flag = nullptr;  // = (int*)(d_nzero + 1)
  unsigned long long* h_status = nullptr;
  // the denominators of the events outside the bins (empty boxes) are 0:
  // the unbinned iteration writes those of the binned events only
  bool lambda_clean = false;
  // sum of rho, exact after set_rho and an estimate after iterations
  // (sum rho' = events with a positive denominator); it sets the
  // fixed-point unit of the log-likelihood gradient
  double rho_sum = 0.0;
  // G(k) = 2^(-(k du)^2) of the fused walks, for the current sk and pw
  float* d_gtab = nullptr;
  float* d_mtab = nullptr;  // (k G(k), k^2 G(k)) pairs
  size_t gtab_n = 0;
  float s_pix = 0.0f, du = -1.0f;
  bool fused_ok = false;    // forward, adjoint and gradient on the fused walks
  bool fused_iter = false;  // the iteration on the fused walks
  void* d_acc = nullptr;  // W*H accumulator, 8 bytes per pixel
  // the row table of the unbanded walks (per sorted event rowtab_stride
  // entries): built when events or the support change, its kernel values
  // rebuilt when the kernel constants change
  petstrip::RowEntry* d_rowtab = nullptr;
  size_t cap_rowtab = 0;  // entries
  int64_t rowtab_stride = 0;
  bool rowtab_on = false;
  StripConst rowtab_sk{};  // kernel constants of its u0, lsc and b1
  float rowtab_spix = -1.0f, rowtab_du = -1.0f;
};

namespace {

int ensure_device(petstrip_handle* h) {
  HIP_TRY(hipSetDevice(h->device));
  return PETSTRIP_OK;
}

// Event arrays sized for n events.
int reserve_events(petstrip_handle* h, int64_t n) {
  // sorted event indices are int32
  if (n > (int64_t)INT32_MAX)
    return fail(PETSTRIP_EINVAL, "more than 2^31-1 events in one handle");
  if ((size_t)n <= h->cap_events && h->d_geom) return PETSTRIP_OK;
  dfree(h->d_resp);
  dfree(h->d_geom);
  dfree(h->d_lambda);
  dfree(h->d_inv);
  dfree(h->d_nbox);
  dfree(h->d_nell);
  dfree(h->d_bin);
  dfree(h->d_sgeom);
  dfree(h->d_sidx);
  h->cap_events = 0;
  h->rowtab_on = false;
  size_t c = (size_t)std::max<int64_t>(n, 1);
  int rc;
  if ((rc = dalloc(&h->d_resp, 3 * c))) return rc;
  if ((rc = dalloc(&h->d_geom, c))) return rc;
  if ((rc = dalloc(&h->d_lambda, c))) return rc;
  if ((rc = dalloc(&h->d_inv, c))) return rc;
  if ((rc = dalloc(&h->d_nbox, c))) return rc;
  if ((rc = dalloc(&h->d_nell, c))) return rc;
  if ((rc = dalloc(&h->d_bin, c))) return rc;
  if ((rc = dalloc(&h->d_sgeom, c))) return rc;
  if ((rc = dalloc(&h->d_sidx, c))) return rc;
  h->cap_events = c;
  return PETSTRIP_OK;
}

// Spatial bins of the box centres: 16 columns by 8 rows keeps the windows
// of the tall (sigma_dl > sigma_z) ellipses within a 128-row column stride.
constexpr int kBinW = 16;
constexpr int kBinH = 8;
constexpr int kChunk = petstrip::kItemEvents;  // events per work item
constexpr int kMinGroupItems = 2048;  // banded items per pipelined group, at least
// MLEM backprojection values K rho / lambda <= 1 are summed in 2^-31 units
// (one conversion per value), or in 2^-40 units in the deterministic
// accumulation mode (two conversions)
constexpr int kMlemFixedBits = 31;
constexpr int kFixed64Bits = 40;
// log-likelihood gradient (K / lambda, unbounded): values up to 2^30.5 units
// take one conversion, up to 2^52 the split one, beyond that a direct FP64
// atomic.  The unit puts a typical value, 1 / (mean rho x mean ellipse
// pixels), at 2^kGradTypicalBits units: each value is then rounded to about
// 1e-6 of itself, and values up to 2^10 times the typical one (cold pixels,
// small denominators) still take the single conversion.
constexpr int kGradTypicalBits = 20;
// The FP32 MLEM iteration sums the same values K / lambda (its image is rho
// times the sums) in 32-bit windows: each work item of the binned kernel
// picks its own unit from its window's rho and event count (strip_fused.h,
// kLog2Sum32), values above its cap add to the image directly.  The launch's
// unit, a typical value at 2^13 units, serves the items whose window sums to
// zero.
constexpr int kMlemTypicalBits = 13;

// Along a box row the reference's float ellipse test has no holes when the
// quadratic's column-to-column curvature dominates the rounding of q.  With
// q convex in the column, q(c) <= (q(c-1) + q(c+1))/2 - C pw^2 exactly; if
// the float q is within E of the exact one, a column rejected between two
// accepted ones needs 2E > C pw^2.  E is bounded by |dq/dz| times the
// rounding of dz (|dq/dz| <= 2 sqrt(9 C) + 1 on the ellipse, dz rounded to
// about two units in the last place of the largest |z|) plus the rounding
// of q near 9; rows are one interval when C pw^2 > 4 E, twice the margin
// the argument needs.  The fused kernel and the footprint counts rely on it.
bool rows_are_intervals(const StripConst& s) {
  double C = 2.0 * s.isz;
  double pw = s.pw;
  double zmax = std::max(std::fabs((double)s.tl_z), std::fabs((double)s.tl_z) + s.W * pw) + 1.0;
  double ulp_z = zmax * 1.2e-7;
  double dq = 2.0 * std::sqrt(9.0 * C) + 1.0;
  double err = dq * ulp_z + 9.0 * 2.4e-7;
  return C * pw * pw > 4.0 * err;
}

// The per-column tables of the fused walks (see strip_fused.hip): G(k) =
// 2^(-(k du)^2), du = pw sqrt(isz log2 e) with the kernel's sigma_z, and the
// (k G(k), k^2 G(k)) pairs of the sigma moments, each rounded once from
// double; rebuilt when sigma_z or the width change.
int update_kernel_table(petstrip_handle* h) {
  const float s_pix = sqrtf(h->sk.isz * 1.44269504088896341f);
  const float du = h->sg.pw * s_pix;
  const size_t n = (size_t)h->sg.W + 8;
  // the tables depend on du alone, but u = pq s_pix takes the new sigma's
  // s_pix even when du rounds to the same float
  h->s_pix = s_pix;
  if (h->d_gtab && du == h->du && n == h->gtab_n) return PETSTRIP_OK;
  std::vector<float> t(n), m(2 * n);
  for (size_t k = 0; k < n; ++k) {
    const double x = (double)k * (double)du;
    const double gk = std::exp2(-x * x);
    t[k] = (float)gk;
    m[2 * k] = (float)((double)k * gk);
    m[2 * k + 1] = (float)((double)k * (double)k * gk);
  }
  if (n != h->gtab_n) {
    dfree(h->d_gtab);
    dfree(h->d_mtab);
    h->gtab_n = 0;
    int rc;
    if ((rc = dalloc(&h->d_gtab, n))) return rc;
    if ((rc = dalloc(&h->d_mtab, 2 * n))) return rc;
    h->gtab_n = n;
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  HIP_TRY(hipMemcpy(h->d_gtab, t.data(), n * sizeof(float), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(h->d_mtab, m.data(), 2 * n * sizeof(float), hipMemcpyHostToDevice));
  h->s_pix = s_pix;
  h->du = du;
  return PETSTRIP_OK;
}

bool rowtab_kernel_ok(const petstrip_handle* h) {
  return h->rowtab_on && h->rowtab_spix == h->s_pix && h->rowtab_du == h->du &&
         memcmp(&h->rowtab_sk, &h->sk, sizeof(StripConst)) == 0;
}

// (Re)computes the row table for the current support, kernel constants and
// binned events.  The unbanded walks then read each box row's ellipse
// columns, u0, row factor and chain ratio instead of forming them every
// iteration: 16 bytes per (event, row) against ~100 VALU instructions per
// row group.  Kept off in the banded form (its rows are clipped to bands),
// with options.row_table = 1, for images wider than the entries' 16-bit
// columns, and when the table would take more than a quarter of the free
// device memory.
int build_row_table(petstrip_handle* h) {
  h->rowtab_on = false;
  if (h->banded || !h->fused_iter || h->opts.row_table != 0 || h->n_sorted == 0 ||
      h->sg.W > 65535)
    return PETSTRIP_OK;
  // slots of whole row groups: a wave's loads of a group stay in its event
  const int64_t stride = ((int64_t)h->max_rows + 63) & ~(int64_t)63;
  // three spare events past the last (the idle segments of a packed wave
  // step read there)
  const size_t need = (size_t)(h->n_sorted + 3) * (size_t)stride;
  if (need > h->cap_rowtab) {
    dfree(h->d_rowtab);
    h->cap_rowtab = 0;
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    if (need * sizeof(petstrip::RowEntry) > free_b / 4) return PETSTRIP_OK;
    int rc;
    if ((rc = dalloc(&h->d_rowtab, need))) return rc;
    h->cap_rowtab = need;
  }
  h->rowtab_stride = stride;
  HIP_TRY(hipMemsetAsync(h->d_rowtab + (size_t)h->n_sorted * stride, 0,
                         3 * (size_t)stride * sizeof(petstrip::RowEntry), h->stream));
  HIP_TRY(petstrip::launch_row_table(h->stream, h->sg, h->sk, h->s_pix, h->du, h->d_sgeom,
                                     h->n_sorted, h->opts.row_scan, stride, h->d_rowtab));
  h->rowtab_on = true;
  h->rowtab_sk = h->sk;
  h->rowtab_spix = h->s_pix;
  h->rowtab_du = h->du;
  return PETSTRIP_OK;
}

// Sorts events by box height class (taller than 32 rows, 17-32, at most 16:
// the walks take one, two or four events per wave) and then into kBinW x
// kBinH bins of their box centre, cuts each bin into chunks of kChunk events
// and sizes each chunk's LDS window.
int prepare_fused(petstrip_handle* h) {
  h->fused_ok = false;
  h->n_items = 0;
  h->lambda_clean = false;
  h->ranges = petstrip::ItemRanges{};
  int rc0;
  if ((rc0 = update_kernel_table(h))) return rc0;
  const int64_t n = h->n_events;
  const int W = h->sg.W, H = h->sg.H;
  std::vector<petstrip::WorkItem> items;
  int64_t acc = 0;
  int max_ww = 1, max_wh = 1, max_rows = 1;
  int rc;
  // Bins of kBinW x bin_h pixels.  A window is about bin_h - 1 rows taller
  // than its tallest box: when that takes it past the column stride the
  // boxes alone need (64-row multiples), the events are binned again with
  // bins low enough to keep it, which halves (or better) the window's LDS.
  // wide_cols: boxes wider than this bin apart from the others (their
  // windows need more LDS than two workgroups per CU leave)
  constexpr int kClasses = petstrip::kHeightClasses * petstrip::kWidthClasses;
  int class_first[kClasses + 1] = {};
  auto bin_events = [&](int bin_h, int wide_cols) -> int {
    const int nbx = (W + kBinW - 1) / kBinW, nby = (H + bin_h - 1) / bin_h;
    const int n_spatial = nbx * nby;
    const int nb = n_spatial * kClasses;
    dfree(h->d_count);
    int rc1;
    if ((rc1 = dalloc(&h->d_count, (size_t)nb * 2))) return rc1;
    HIP_TRY(hipMemsetAsync(h->d_count, 0, (size_t)nb * 2 * sizeof(int32_t), h->stream));
    HIP_TRY(petstrip::launch_bin_count(h->stream, h->d_geom, n, kBinW, bin_h, nbx, n_spatial,
                                       wide_cols, h->d_bin, h->d_count));
    std::vector<int32_t> cnt((size_t)nb), off((size_t)nb);
    HIP_TRY(hipMemcpyAsync(cnt.data(), h->d_count, nb * sizeof(int32_t),
                           hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    acc = 0;
    items.clear();
    for (int b = 0; b < nb; ++b) {
      // the first item of each (height, width) class
      if (b % n_spatial == 0) class_first[b / n_spatial] = (int)items.size();
      off[(size_t)b] = (int32_t)acc;
      for (int64_t s0 = acc; s0 < acc + cnt[(size_t)b]; s0 += kChunk) {
        petstrip::WorkItem it{};
        it.begin = (int32_t)s0;
        it.end = (int32_t)std::min<int64_t>(s0 + kChunk, acc + cnt[(size_t)b]);
        items.push_back(it);
      }
      acc += cnt[(size_t)b];
    }
    class_first[kClasses] = (int)items.size();
    int32_t* d_cursor = h->d_count + nb;
    HIP_TRY(hipMemcpyAsync(d_cursor, off.data(), nb * sizeof(int32_t),
                           hipMemcpyHostToDevice, h->stream));
    HIP_TRY(petstrip::launch_bin_scatter(h->stream, h->d_geom, n, h->d_bin, d_cursor,
                                         h->d_sgeom, h->d_sidx));
    h->n_items = (int)items.size();
    if (items.size() > h->cap_items) {
      dfree(h->d_items);
      h->cap_items = 0;
      if ((rc1 = dalloc(&h->d_items, items.size()))) return rc1;
      h->cap_items = items.size();
    }
    if (!items.empty()) {
      HIP_TRY(hipMemcpyAsync(h->d_items, items.data(),
                             items.size() * sizeof(petstrip::WorkItem),
                             hipMemcpyHostToDevice, h->stream));
      HIP_TRY(petstrip::launch_item_windows(h->stream, h->d_sgeom, h->d_items,
                                            h->n_items));
      HIP_TRY(hipMemcpyAsync(items.data(), h->d_items,
                             items.size() * sizeof(petstrip::WorkItem),
                             hipMemcpyDeviceToHost, h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    max_ww = max_wh = max_rows = 1;
    for (auto& it : items) {
      max_ww = std::max(max_ww, it.ww);
      max_wh = std::max(max_wh, it.wh);
      max_rows = std::max(max_rows, it.max_rows);
    }
    return PETSTRIP_OK;
  };
  constexpr int kNoSplit = 1 << 30;
  if ((rc = bin_events(kBinH, kNoSplit))) return rc;
  {
    // a bin's centres span bin_h rows, so its window spans at most
    // max_rows + bin_h - 1 rows
    const int want = petstrip::fused_col_stride(max_rows);
    int bh = kBinH;
    while (bh > 1 && max_rows + bh - 1 > want) bh >>= 1;
    const bool lower = want != 0 && petstrip::fused_col_stride(max_wh) != want &&
                       bh < kBinH && max_rows + bh - 1 <= want;
    if (!lower) bh = kBinH;
    // Windows too wide for two workgroups per CU (the 32-bit windows of the
    // iteration): the narrower boxes bin apart, so that their items' windows
    // fit (a window spans its boxes plus the bin's kBinW centre columns and
    // the rounding of the centres).
    const int hp = lower ? want : petstrip::fused_col_stride(max_wh);
    int wide_cols = kNoSplit;
    if (hp != 0 &&
        petstrip::fused_lds_bytes(max_ww, hp, petstrip::MODE_MLEM_K) > petstrip::kTwoPerCuLds) {
      int t4 = 0;
      while (petstrip::fused_lds_bytes(t4 + 1, hp, petstrip::MODE_MLEM_K) <=
             petstrip::kTwoPerCuLds)
        ++t4;
      if (t4 - (kBinW + 3) >= 16) wide_cols = t4 - (kBinW + 3);
    }
    if ((lower || wide_cols != kNoSplit) && (rc = bin_events(bh, wide_cols))) return rc;
  }
  h->max_rows = max_rows;
  h->max_ww = max_ww;
  h->n_sorted = acc;
  h->banded = false;
  h->col_stride = petstrip::fused_col_stride(max_wh);
  // Every mode accumulates its LDS window in exact fixed point (ds_add_u64
  // runs at several times the rate of ds_add_f32 on gfx950); the
  // accumulation mode only picks the global format.  A workgroup of 16
  // waves holds one CU's LDS for the larger windows.
  h->lds_bytes = h->col_stride
                     ? petstrip::fused_lds_bytes(max_ww, h->col_stride, petstrip::MODE_MLEM)
                     : (size_t)-1;
  h->ncb = 1;
  h->cshift = 30;
  if ((h->lds_bytes > h->lds_limit || h->opts.force_bands) && !items.empty()) {
    // banded form: each chunk's window cut into bands of kBandRows image
    // rows, and into bands of 2^kColBandShift columns when a row band of the
    // widest window still exceeds LDS
    const int br = petstrip::kBandRows;
    const bool cols =
        petstrip::fused_lds_bytes(max_ww, br, petstrip::MODE_MLEM) > h->lds_limit ||
        h->opts.force_bands == 2;
    const int cshift = cols ? petstrip::kColBandShift : 30;
    const int cb = 1 << cshift;
    std::vector<petstrip::WorkItem> bands;
    std::vector<int> chunk_first;  // first band item of each chunk
    chunk_first.reserve(items.size() + 1);
    for (const auto& it : items) {
      chunk_first.push_back((int)bands.size());
      if (it.wh <= 0 || it.ww <= 0) continue;
      for (int b = it.wy0 / br; b <= (it.wy0 + it.wh - 1) / br; ++b) {
        for (int c = it.wx0 >> cshift; c <= (it.wx0 + it.ww - 1) >> cshift; ++c) {
          petstrip::WorkItem bi = it;
          bi.wy0 = std::max(it.wy0, b * br);
          bi.wh = std::min(it.wy0 + it.wh, (b + 1) * br) - bi.wy0;
          bi.wx0 = (int32_t)std::max<int64_t>(it.wx0, (int64_t)c * cb);
          bi.ww = (int32_t)(std::min<int64_t>(it.wx0 + it.ww, (int64_t)(c + 1) * cb) - bi.wx0);
          bi.pad = b;
          bands.push_back(bi);
        }
      }
    }
    const int band_ww = cols ? std::min(max_ww, cb) : max_ww;
    const size_t lds = petstrip::fused_lds_bytes(band_ww, br, petstrip::MODE_MLEM);
    if (lds <= h->lds_limit) {
      h->banded = true;
      h->col_stride = br;
      h->lds_bytes = lds;
      h->max_ww = band_ww;
      // slots per event: every row band and column band its box can reach
      h->ncb = cols ? (max_ww + cb - 1) / cb + 1 : 1;
      h->cshift = cshift;
      h->nbands = ((max_rows + br - 1) / br + 1) * h->ncb;
      const size_t need = (size_t)std::max<int64_t>(acc, 1) * h->nbands * 4;
      if (need > h->cap_part) {
        dfree(h->d_part);
        h->cap_part = 0;
        if ((rc = dalloc(&h->d_part, need))) return rc;
        h->cap_part = need;
      }
      dfree(h->d_lam_sorted);
      if ((rc = dalloc(&h->d_lam_sorted, (size_t)std::max<int64_t>(acc, 1)))) return rc;
      // up to 8 groups of whole chunks, each enough items to fill the device
      const int n_chunks = (int)items.size();
      const int n_groups = (int)std::max<size_t>(
          1, std::min<size_t>(8, bands.size() / (size_t)kMinGroupItems));
      h->groups.clear();
      int c0 = 0;
      for (int g = 0; g < n_groups; ++g) {
        const size_t target = bands.size() * (size_t)(g + 1) / n_groups;
        int c1 = c0;
        while (c1 < n_chunks && (g == n_groups - 1 || (size_t)chunk_first[c1] < target)) ++c1;
        if (c1 == c0) continue;
        petstrip_handle::BandGroup bg;
        bg.item_begin = chunk_first[c0];
        bg.item_end = c1 < n_chunks ? chunk_first[c1] : (int)bands.size();
        bg.event_begin = items[c0].begin;
        bg.event_end = c1 < n_chunks ? items[c1].begin : acc;
        h->groups.push_back(bg);
        c0 = c1;
      }
      items.swap(bands);
      h->n_items = (int)items.size();
      // banded items walk one event per wave (launch_strip_banded)
      if (items.size() > h->cap_items) {
        dfree(h->d_items);
        if ((rc = dalloc(&h->d_items, items.size()))) return rc;
        h->cap_items = items.size();
      }
      HIP_TRY(hipMemcpyAsync(h->d_items, items.data(),
                             items.size() * sizeof(petstrip::WorkItem),
                             hipMemcpyHostToDevice, h->stream));
      HIP_TRY(hipStreamSynchronize(h->stream));
    }
  }
  h->window_px = 0;
  for (const auto& it : items) h->window_px += (int64_t)it.ww * it.wh;
  // the unbanded launches: one range per non-empty (height, width) class
  h->ranges = petstrip::ItemRanges{};
  h->ranges.n_items = h->n_items;
  if (!h->banded) {
    for (int k = 0; k < kClasses; ++k) {
      const int b0 = class_first[k], b1 = class_first[k + 1];
      if (b1 <= b0) continue;
      int mw = 1;
      for (int i = b0; i < b1; ++i) mw = std::max(mw, items[(size_t)i].ww);
      const int cls = k / petstrip::kWidthClasses;
      const int r = h->ranges.n++;
      h->ranges.begin[r] = b0;
      h->ranges.end[r] = b1;
      h->ranges.pack[r] = cls == 0 ? 1 : cls == 1 ? 2 : 4;
      h->ranges.max_ww[r] = mw;
    }
  }
  {
    const size_t need = (size_t)std::max<int64_t>(h->n_items, petstrip::kCombineBlocks) *
                        petstrip::kScalStride;
    if (need > h->cap_scal_part) {
      dfree(h->d_scal_part);
      h->cap_scal_part = 0;
      if ((rc = dalloc(&h->d_scal_part, need))) return rc;
      h->cap_scal_part = need;
    }
  }
  const bool walks = h->lds_bytes <= h->lds_limit && rows_are_intervals(h->sg) &&
                     h->opts.path == PETSTRIP_PATH_AUTO;
  h->fused_ok = walks && h->opts.update == PETSTRIP_UPDATE_CPU &&
                h->opts.kernel == PETSTRIP_KERNEL_GAUSSIAN;
  // the FP32 iteration also takes the reference --gpu variant: rho times the
  // sensitivity in pass 1 (MLEM_K), and the analytic kernel evaluated per
  // pixel (MLEM_KA: one window, every row group of a box cached)
  const bool fp32 = h->opts.accumulation == PETSTRIP_ACCUM_FP32;
  h->fused_iter =
      h->fused_ok ||
      (walks && fp32 &&
       (h->opts.kernel == PETSTRIP_KERNEL_GAUSSIAN ||
        (h->opts.kernel == PETSTRIP_KERNEL_ANALYTIC && !h->banded && h->max_rows <= 256)));
  return build_row_table(h);
}

// Geometry and footprint statistics for the responses in d_resp.
int prepare_events(petstrip_handle* h) {
  const int64_t n = h->n_events;
  HIP_TRY(petstrip::launch_event_setup(h->stream, h->sg, h->d_resp, n,
                                       h->d_geom));
  // footprints from the row intervals the fused walks use when rows are
  // hole-free, else by testing every box pixel
  if (rows_are_intervals(h->sg))
    HIP_TRY(petstrip::launch_row_counts(h->stream, h->sg, h->d_geom, n, h->opts.row_scan,
                                        h->d_nbox, h->d_nell));
  else
    HIP_TRY(petstrip::launch_footprint(h->stream, h->sg, h->d_geom, n, h->d_nbox,
                                       h->d_nell));
  std::vector<int32_t> nb((size_t)n), ne((size_t)n);
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(nb.data(), h->d_nbox, n * sizeof(int32_t),
                           hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(ne.data(), h->d_nell, n * sizeof(int32_t),
                           hipMemcpyDeviceToHost, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  int64_t sb = 0, se = 0;
  for (int64_t i = 0; i < n; ++i) {
    sb += nb[(size_t)i];
    se += ne[(size_t)i];
  }
  h->stats.n_events = n;
  h->stats.n_box_pixels = sb;
  h->stats.n_ellipse_pixels = se;
  h->stats.flop_alg = 15.0 * (double)sb + 41.0 * (double)se + 50.0 * (double)n;
  h->lambda_valid = false;
  return prepare_fused(h);
}

const float* denominator_weight(petstrip_handle* h) {
  return h->opts.update == PETSTRIP_UPDATE_SENSITIVITY_DENOMINATOR ? h->d_sens
                                                                   : nullptr;
}

// One MLEM iteration: rho <- rho * backprojection(inv) (kept as the
// reference form out[p] += K rho[p] inv_e).
petstrip::FusedArgs fused_args(petstrip_handle* h) {
  petstrip::FusedArgs a{};
  a.sg = h->sg;
  a.sk = h->sk;
  a.sgeom = h->d_sgeom;
  a.sidx = h->d_sidx;
  a.items = h->d_items;
  a.rho = h->d_rho;
  a.rho_weight = nullptr;
  a.lambda = h->d_lambda;
  a.g = nullptr;
  a.acc = h->d_acc;
  a.acc_kind = petstrip::GLOBAL_F32;
  a.log2_scale = (float)kMlemFixedBits;
  a.split = 0;
  a.inv_fixed_scale = std::ldexp(1.0, -kMlemFixedBits);
  a.scalars = h->d_scalars;
  a.scal_part = h->d_scal_part;
  a.row_scan = h->opts.row_scan;
  a.s_pix = h->s_pix;
  a.du = h->du;
  a.gtab = h->d_gtab;
  a.mtab = h->d_mtab;
  a.part = h->d_part;
  a.nbands = h->nbands;
  a.ncb = h->ncb;
  a.cshift = h->cshift;
  a.lam_sorted = h->d_lam_sorted;
  a.horner = h->opts.pass1_walk == 0 ? 1 : 0;
  a.log2_ell = (float)std::log2(std::max(
      1.0, (double)h->stats.n_ellipse_pixels / (double)std::max<int64_t>(h->stats.n_events, 1)));
  if (h->rowtab_on) {
    a.rowtab = h->d_rowtab;
    a.rowtab_stride = h->rowtab_stride;
    a.rowtab_kernel = rowtab_kernel_ok(h) ? 1 : 0;
  }
  return a;
}

// The forward projection of the fused forms: lambda per event and, with
// sigma, the log-likelihood terms and the weighted sigma derivative in the
// scalars (weights g, or 1/lambda when a.g is null).  The caller zeroes
// lambda and the scalars.
int fused_forward(petstrip_handle* h, const petstrip::FusedArgs& a, bool sigma) {
  if (!h->banded)
    return hiperr(petstrip::launch_strip_fused(h->stream, petstrip::MODE_FORWARD,
                                               h->col_stride, 0, a, h->ranges));
  const size_t part = (size_t)std::max<int64_t>(h->n_sorted, 1) * h->nbands * 4;
  HIP_TRY(hipMemsetAsync(h->d_part, 0, part * sizeof(float), h->stream));
  HIP_TRY(petstrip::launch_strip_banded(
      h->stream, sigma ? petstrip::MODE_FORWARD : petstrip::MODE_MLEM, 1, a, h->n_items,
      petstrip::fused_lds_bytes(h->max_ww, h->col_stride, petstrip::MODE_FORWARD)));
  HIP_TRY(petstrip::launch_band_combine(h->stream, a, h->n_sorted, sigma ? 1 : 0));
  return PETSTRIP_OK;
}

// Pass 1 (mode1, BAND = 1), the combine and pass 2 (mode2, BAND = 2) of the
// banded form, group by group: pass 1 and the combine of a group run on the
// handle's stream and its pass 2 on the second stream once they are done,
// so pass 2 of one group (bound by LDS atomics) runs beside pass 1 of the
// next (bound by VALU issue).  The handle's stream waits for both at the end.
int banded_pipeline(petstrip_handle* h, const petstrip::FusedArgs& a, int mode1,
                    bool sigma, int mode2, size_t lds2) {
  const size_t part = (size_t)std::max<int64_t>(h->n_sorted, 1) * h->nbands * 4;
  HIP_TRY(hipMemsetAsync(h->d_part, 0, part * sizeof(float), h->stream));
  const size_t lds1 =
      petstrip::fused_lds_bytes(h->max_ww, h->col_stride, petstrip::MODE_FORWARD);
  int k = 0;
  for (const auto& g : h->groups) {
    petstrip::FusedArgs b = a;
    b.items = a.items + g.item_begin;
    const int ni = g.item_end - g.item_begin;
    HIP_TRY(petstrip::launch_strip_banded(h->stream, mode1, 1, b, ni, lds1));
    petstrip::FusedArgs c = a;  // the group's sorted events
    c.part = a.part + (size_t)g.event_begin * h->nbands * 4;
    c.sidx = a.sidx + g.event_begin;
    c.lam_sorted = a.lam_sorted + g.event_begin;
    HIP_TRY(petstrip::launch_band_combine(h->stream, c, g.event_end - g.event_begin,
                                          sigma ? 1 : 0));
    hipEvent_t ev = h->ev_group[k++ % 8];
    HIP_TRY(hipEventRecord(ev, h->stream));
    HIP_TRY(hipStreamWaitEvent(h->stream2, ev, 0));
    HIP_TRY(petstrip::launch_strip_banded(h->stream2, mode2, 2, b, ni, lds2));
  }
  HIP_TRY(hipEventRecord(h->ev_join, h->stream2));
  HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));
  return PETSTRIP_OK;
}

int grad_unit_bits(const petstrip_handle* h, int typical_bits = kGradTypicalBits);

int one_iteration(petstrip_handle* h) {
  const int64_t n = h->n_events;
  if (h->fused_iter) {
    const int accm = h->opts.accumulation;
    // the table's kernel values after a change of the kernel's sigmas
    int rc0;
    if (h->rowtab_on && !rowtab_kernel_ok(h) && (rc0 = build_row_table(h))) return rc0;
    petstrip::FusedArgs a = fused_args(h);
    a.rho_weight = denominator_weight(h);  // the --gpu variant's sensitivity
    // FP32: the 32-bit windows' rows above an item's cap add their values
    // to the image directly, possibly many to one pixel, so the image sums
    // in FP64 (the flush adds one value per window pixel)
    a.acc_kind = accm == PETSTRIP_ACCUM_FIXED64 ? petstrip::GLOBAL_U64 : petstrip::GLOBAL_F64;
    // FP32: pass 2 adds K / lambda_e (as the gradient does) and the sums are
    // multiplied by rho at the end; its unit puts a typical value at
    // 2^kGradTypicalBits.  The other forms add K rho / lambda_e.
    const bool k_form = accm == PETSTRIP_ACCUM_FP32;
    const int mode2 = h->opts.kernel == PETSTRIP_KERNEL_ANALYTIC ? petstrip::MODE_MLEM_KA
                      : k_form                                    ? petstrip::MODE_MLEM_K
                                                                  : petstrip::MODE_MLEM;
    if (accm == PETSTRIP_ACCUM_FIXED64) {
      // the deterministic image sums every value in 2^-40 units
      a.log2_scale = (float)kFixed64Bits;
      a.split = 1;
      a.inv_fixed_scale = std::ldexp(1.0, -kFixed64Bits);
    } else if (k_form) {
      const int k = grad_unit_bits(h, kMlemTypicalBits);
      a.log2_scale = (float)k;
      a.inv_fixed_scale = std::ldexp(1.0, -k);
    }
    // the unbanded launch writes every binned event's denominator; the
    // others stay at the 0 of the last fill
    if (h->banded || !h->lambda_clean) {
      HIP_TRY(petstrip::launch_fill(h->stream, h->d_lambda, 0.0f, n));
      h->lambda_clean = true;
    }
    HIP_TRY(hipEventRecord(h->evk0, h->stream));
    if (h->banded) {
      int rc;
      const size_t lds2 = petstrip::fused_lds_bytes(h->max_ww, h->col_stride, mode2, 2);
      if ((rc = banded_pipeline(h, a, petstrip::MODE_MLEM, false, mode2, lds2))) return rc;
    } else {
      HIP_TRY(petstrip::launch_strip_fused(h->stream, mode2, h->col_stride,
                                           accm == PETSTRIP_ACCUM_COMPENSATED ? 1 : 0, a,
                                           h->ranges));
    }
    HIP_TRY(hipEventRecord(h->evk1, h->stream));
    // float accumulators already hold image units; the u64 one holds 2^-40
    double unit = a.acc_kind == petstrip::GLOBAL_U64 ? a.inv_fixed_scale : 1.0;
    HIP_TRY(petstrip::launch_finalize(h->stream, h->d_acc, a.acc_kind, unit, 0.0f, 1,
                                      h->d_flag, h->d_out, (int64_t)h->n_pix,
                                      k_form ? h->d_rho : nullptr, h->d_nzero));
    std::swap(h->d_rho, h->d_out);
    h->lambda_valid = false;
    // the next unit's image scale: sum rho' = events with a positive
    // denominator, at most n
    h->rho_sum = (double)n;
    return PETSTRIP_OK;
  }
  // per-event kernels: compensated denominators in the compensated mode;
  // the deterministic mode sums the image in 2^-40 units (exact, any order),
  // the others in FP64
  const int accm = h->opts.accumulation;
  const bool fixed = accm == PETSTRIP_ACCUM_FIXED64;
  HIP_TRY(hipEventRecord(h->evk0, h->stream));
  h->lambda_clean = false;
  HIP_TRY(petstrip::launch_denominator(h->stream, h->sg, h->sk, h->d_geom, n,
                                       h->d_rho, denominator_weight(h),
                                       h->d_lambda, h->d_inv, h->opts.kernel,
                                       accm == PETSTRIP_ACCUM_COMPENSATED ? 1 : 0));
  // the accumulator is zero between uses (finalize clears it)
  HIP_TRY(petstrip::launch_backproject(
      h->stream, h->sg, h->sk, h->d_geom, n, h->d_rho, h->d_inv,
      fixed ? nullptr : (double*)h->d_acc, h->opts.kernel,
      fixed ? (unsigned long long*)h->d_acc : nullptr, std::ldexp(1.0, kFixed64Bits)));
  HIP_TRY(hipEventRecord(h->evk1, h->stream));
  HIP_TRY(petstrip::launch_finalize(
      h->stream, h->d_acc, fixed ? petstrip::GLOBAL_U64 : petstrip::GLOBAL_F64,
      fixed ? std::ldexp(1.0, -kFixed64Bits) : 1.0, 0.0f, fixed ? 1 : 0,
      fixed ? h->d_flag : nullptr, h->d_out, (int64_t)h->n_pix, nullptr, h->d_nzero));
  std::swap(h->d_rho, h->d_out);
  h->lambda_valid = false;
  return PETSTRIP_OK;
}

}  // namespace

extern "C" {

int petstrip_abi_version(void) { return PETSTRIP_ABI_VERSION; }

const char* petstrip_last_error(void) { return g_last_error.c_str(); }

int petstrip_device_count(int* count) {
  if (!count) return fail(PETSTRIP_EINVAL, "count is NULL");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  return PETSTRIP_OK;
}

int petstrip_create(const petstrip_scanner* scanner, int device,
                    const petstrip_options* opts, petstrip_handle** out) {
  if (!out) return fail(PETSTRIP_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = check_scanner(scanner);
  if (rc) return rc;
  if (opts && (opts->accumulation < 0 || opts->accumulation > 2 ||
               opts->update < 0 || opts->update > 1 || opts->path < 0 ||
               opts->path > 1 || opts->row_scan < 0 || opts->row_scan > 1 ||
               opts->force_bands < 0 || opts->force_bands > 2 || opts->kernel < 0 ||
               opts->kernel > 1 || opts->row_table < 0 || opts->row_table > 1))
    return fail(PETSTRIP_EINVAL, "invalid options");
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
    return fail(PETSTRIP_EDEVICE, "no HIP device available");
  if (device < 0 || device >= count)
    return fail(PETSTRIP_EDEVICE, "device index " + std::to_string(device) +
                                      " out of range (have " +
                                      std::to_string(count) + ")");
  petstrip_handle* h = new petstrip_handle();
  h->device = device;
  h->scanner = *scanner;
  if (opts) h->opts = *opts;
  h->sg = make_const(*scanner, scanner->sigma_z, scanner->sigma_dl);
  h->sk = h->sg;
  h->n_pix = (size_t)scanner->n_y_pixels * (size_t)scanner->n_z_pixels;
  auto bail = [&](int code) {
    petstrip_destroy(h);
    return code;
  };
  if ((rc = ensure_device(h))) return bail(rc);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess)
    return bail(fail(PETSTRIP_EDEVICE, "hipGetDeviceProperties failed"));
  h->name = prop.name;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(PETSTRIP_EDEVICE, "hipStreamCreate failed"));
  for (auto& e : h->ev_group)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
      return bail(fail(PETSTRIP_EDEVICE, "hipEventCreate failed"));
  if (hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess)
    return bail(fail(PETSTRIP_EDEVICE, "hipEventCreate failed"));
  if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
      hipEventCreate(&h->evk0) != hipSuccess || hipEventCreate(&h->evk1) != hipSuccess)
    return bail(fail(PETSTRIP_EDEVICE, "hipEventCreate failed"));
  if ((rc = dalloc(&h->d_rho, h->n_pix))) return bail(rc);
  if ((rc = dalloc(&h->d_out, h->n_pix))) return bail(rc);
  if ((rc = dalloc(&h->d_sens, h->n_pix))) return bail(rc);
  if ((rc = dalloc(&h->d_scalars, 8))) return bail(rc);
  if ((rc = dalloc((double**)&h->d_acc, h->n_pix))) return bail(rc);
  if ((rc = dalloc(&h->d_nzero, 2))) return bail(rc);
  h->d_flag = reinterpret_cast<int*>(h->d_nzero + 1);
  if (hipHostMalloc((void**)&h->h_status, 2 * sizeof(unsigned long long)) != hipSuccess) {
    h->h_status = nullptr;
    return bail(fail(PETSTRIP_EDEVICE, "hipHostMalloc failed"));
  }
  if (hipMemsetAsync(h->d_nzero, 0, 2 * sizeof(unsigned long long), h->stream) != hipSuccess)
    return bail(fail(PETSTRIP_EDEVICE, "hipMemset failed"));
  if (hipMemsetAsync(h->d_acc, 0, h->n_pix * sizeof(double), h->stream) != hipSuccess)
    return bail(fail(PETSTRIP_EDEVICE, "hipMemset failed"));
  {
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin,
                              device) == hipSuccess && optin > 0)
      h->lds_limit = (size_t)optin;
    else
      h->lds_limit = prop.sharedMemPerBlock;
  }
  if ((rc = reserve_events(h, 1))) return bail(rc);
  // rho = 1 (reconstruction.h:59), 5-point sensitivity for the --gpu variant
  h->rho_sum = (double)h->n_pix;
  if (petstrip::launch_fill(h->stream, h->d_rho, 1.0f, (int64_t)h->n_pix) ||
      petstrip::launch_sensitivity(h->stream, h->sg, 1, h->d_sens) ||
      hipStreamSynchronize(h->stream) != hipSuccess)
    return bail(fail(PETSTRIP_EDEVICE, "initialisation kernels failed"));
  *out = h;
  return PETSTRIP_OK;
}

int petstrip_destroy(petstrip_handle* h) {
  if (!h) return PETSTRIP_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->stream2) (void)hipStreamSynchronize(h->stream2);
  dfree(h->d_resp);
  dfree(h->d_geom);
  dfree(h->d_lambda);
  dfree(h->d_inv);
  dfree(h->d_nbox);
  dfree(h->d_nell);
  dfree(h->d_rho);
  dfree(h->d_out);
  dfree(h->d_sens);
  dfree(h->d_scalars);
  dfree(h->d_scal_part);
  dfree(h->d_bin);
  dfree(h->d_count);
  dfree(h->d_sgeom);
  dfree(h->d_sidx);
  dfree(h->d_items);
  dfree(h->d_nzero);
  h->d_flag = nullptr;
  if (h->h_status) (void)hipHostFree(h->h_status);
  h->h_status = nullptr;
  dfree(h->d_gtab);
  dfree(h->d_mtab);
  dfree(h->d_part);
  dfree(h->d_lam_sorted);
  dfree(h->d_rowtab);
  if (h->d_acc) (void)hipFree(h->d_acc);
  h->d_acc = nullptr;
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  for (auto& e : h->ev_group)
    if (e) (void)hipEventDestroy(e);
  if (h->ev_join) (void)hipEventDestroy(h->ev_join);
  if (h->evk0) (void)hipEventDestroy(h->evk0);
  if (h->evk1) (void)hipEventDestroy(h->evk1);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  if (h->stream2) (void)hipStreamDestroy(h->stream2);
  delete h;
  return PETSTRIP_OK;
}

int petstrip_device_name(petstrip_handle* h, char* buf, size_t len) {
  if (!h || !buf || len == 0) return fail(PETSTRIP_EINVAL, "bad argument");
  snprintf(buf, len, "%s", h->name.c_str());
  return PETSTRIP_OK;
}

int petstrip_stream(petstrip_handle* h, void** stream) {
  if (!h || !stream) return fail(PETSTRIP_EINVAL, "bad argument");
  *stream = (void*)h->stream;
  return PETSTRIP_OK;
}

int petstrip_set_options(petstrip_handle* h, const petstrip_options* opts) {
  if (!h || !opts) return fail(PETSTRIP_EINVAL, "bad argument");
  if (opts->accumulation < 0 || opts->accumulation > 2 || opts->update < 0 ||
      opts->update > 1 || opts->path < 0 || opts->path > 1 || opts->row_scan < 0 ||
      opts->row_scan > 1 || opts->force_bands < 0 || opts->force_bands > 2 ||
      opts->kernel < 0 || opts->kernel > 1 || opts->row_table < 0 || opts->row_table > 1)
    return fail(PETSTRIP_EINVAL, "invalid options");
  h->opts = *opts;
  h->lambda_valid = false;
  int rc;
  if ((rc = ensure_device(h))) return rc;
  // the footprint counts follow the row-interval method
  return h->n_events > 0 ? prepare_events(h) : prepare_fused(h);
}

int petstrip_set_events(petstrip_handle* h, const float* responses, int64_t n) {
  if (!h) return fail(PETSTRIP_EINVAL, "handle is NULL");
  if (n < 0 || (n > 0 && !responses))
    return fail(PETSTRIP_EINVAL, "bad responses/n");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  if ((rc = reserve_events(h, n))) return rc;
  h->n_events = n;
  if (n > 0)
    HIP_TRY(hipMemcpyAsync(h->d_resp, responses, (size_t)n * 3 * sizeof(float),
                           hipMemcpyHostToDevice, h->stream));
  return prepare_events(h);
}

int petstrip_set_events_device(petstrip_handle* h, const float* d_responses,
                               int64_t n) {
  if (!h) return fail(PETSTRIP_EINVAL, "handle is NULL");
  if (n < 0 || (n > 0 && !d_responses))
    return fail(PETSTRIP_EINVAL, "bad responses/n");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  if ((rc = reserve_events(h, n))) return rc;
  h->n_events = n;
  if (n > 0)
    HIP_TRY(hipMemcpyAsync(h->d_resp, d_responses, (size_t)n * 3 * sizeof(float),
                           hipMemcpyDeviceToDevice, h->stream));
  return prepare_events(h);
}

int petstrip_get_events(petstrip_handle* h, float* responses) {
  if (!h || (h->n_events > 0 && !responses)) return fail(PETSTRIP_EINVAL, "bad argument");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  if (h->n_events > 0)
    HIP_TRY(hipMemcpyAsync(responses, h->d_resp, (size_t)h->n_events * 3 * sizeof(float),
                           hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return PETSTRIP_OK;
}

int petstrip_set_phantom_events(petstrip_handle* h, const double* regions,
                                int32_t n_regions, uint64_t seed, int64_t n) {
  if (!h) return fail(PETSTRIP_EINVAL, "handle is NULL");
  if (n < 0) return fail(PETSTRIP_EINVAL, "bad n");
  petstrip::PhantomSpec p;
  int rc;
  if ((rc = petstrip::phantom_spec(&h->scanner, regions, n_regions, seed, p))) return rc;
  if ((rc = ensure_device(h))) return rc;
  if ((rc = reserve_events(h, n))) return rc;
  h->n_events = n;
  bool failed = false;
  if ((rc = petstrip::phantom_device(p, n, h->d_resp, h->stream, &failed))) return rc;
  if (failed) {
    h->n_events = 0;
    return fail(PETSTRIP_ESTATE, "no detected emission after 2^16 draws in a slot");
  }
  return prepare_events(h);
}

int petstrip_set_rho(petstrip_handle* h, const float* rho) {
  if (!h || !rho) return fail(PETSTRIP_EINVAL, "bad argument");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  HIP_TRY(hipMemcpyAsync(h->d_rho, rho, h->n_pix * sizeof(float),
                         hipMemcpyHostToDevice, h->stream));
  double sum = 0.0;
  for (size_t i = 0; i < h->n_pix; ++i) sum += rho[i];
  h->rho_sum = sum;
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->lambda_valid = false;
  return PETSTRIP_OK;
}

int petstrip_set_rho_device(petstrip_handle* h, const float* d_rho) {
  if (!h || !d_rho) return fail(PETSTRIP_EINVAL, "bad argument");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  HIP_TRY(hipMemcpyAsync(h->d_rho, d_rho, h->n_pix * sizeof(float),
                         hipMemcpyDeviceToDevice, h->stream));
  std::vector<float> host(h->n_pix);
  HIP_TRY(hipMemcpyAsync(host.data(), d_rho, h->n_pix * sizeof(float),
                         hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  double sum = 0.0;
  for (float v : host) sum += v;
  h->rho_sum = sum;
  h->lambda_valid = false;
  return PETSTRIP_OK;
}

int petstrip_get_rho(petstrip_handle* h, float* rho) {
  if (!h || !rho) return fail(PETSTRIP_EINVAL, "bad argument");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  HIP_TRY(hipMemcpyAsync(rho, h->d_rho, h->n_pix * sizeof(float),
                         hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return PETSTRIP_OK;
}

int petstrip_rho_device(petstrip_handle* h, const float** d_rho) {
  if (!h || !d_rho) return fail(PETSTRIP_EINVAL, "bad argument");
  *d_rho = h->d_rho;
  return PETSTRIP_OK;
}

int petstrip_iterate(petstrip_handle* h, int n_iterations) {
  if (!h) return fail(PETSTRIP_EINVAL, "handle is NULL");
  if (n_iterations < 0) return fail(PETSTRIP_EINVAL, "negative iteration count");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  for (int i = 0; i < n_iterations; ++i) {
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    if ((rc = one_iteration(h))) return rc;
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
  }
  // the last iteration's finalize cleared the counter; counter and overflow
  // flag come back in one copy
  if (n_iterations > 0) {
    HIP_TRY(petstrip::launch_count_nonpositive(h->stream, h->d_lambda, h->n_events,
                                               h->d_nzero, false));
    HIP_TRY(hipMemcpyAsync(h->h_status, h->d_nzero, 2 * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (n_iterations > 0) {
    const unsigned long long nzero = h->h_status[0];
    h->stats.n_zero_denominator = (int64_t)nzero;
    // sum rho' = events with a positive denominator (the CPU update; within a
    // small factor for the sensitivity-weighted one)
    h->rho_sum = (double)(h->n_events - (int64_t)nzero);
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->stats.last_iteration_ms = ms;
    HIP_TRY(hipEventElapsedTime(&ms, h->evk0, h->evk1));
    h->stats.last_kernel_ms = ms;
    int flag = 0;
    memcpy(&flag, &h->h_status[1], sizeof(int));
    if (flag) {
      HIP_TRY(hipMemset(h->d_flag, 0, sizeof(int)));
      return fail(PETSTRIP_ESTATE,
                  "fixed-point image overflow: a pixel sum exceeded 2^23 in 2^-40 "
                  "units");
    }
  }
  return PETSTRIP_OK;
}

int petstrip_run(petstrip_handle* h, float* rho_inout, int n_blocks,
                 int n_iter_in_block, petstrip_output_cb output,
                 petstrip_progress_cb progress, void* user) {
  if (!h || !rho_inout) return fail(PETSTRIP_EINVAL, "bad argument");
  if (n_blocks < 0 || n_iter_in_block < 0)
    return fail(PETSTRIP_EINVAL, "negative block/iteration count");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  const int W = h->scanner.n_z_pixels, H = h->scanner.n_y_pixels;
  if (output) {
    std::vector<float> sens(h->n_pix);
    HIP_TRY(hipMemcpyAsync(sens.data(), h->d_sens, h->n_pix * sizeof(float),
                           hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (output(user, -1, sens.data(), W, H))
      return fail(PETSTRIP_ECALLBACK, "output callback stopped the run");
  }
  if ((rc = petstrip_set_rho(h, rho_inout))) return rc;
  for (int block = 0; block < n_blocks; ++block) {
    for (int i = 0; i < n_iter_in_block; ++i) {
      int iteration = block * n_iter_in_block + i;
      if (progress) progress(user, iteration, 0);
      if ((rc = petstrip_iterate(h, 1))) return rc;
      if (progress) progress(user, iteration, 1);
      // first block: iterations 1-5, 10, 15, 20, 30, 50, 100 (.cu:98-104)
      if (!block && i < n_iter_in_block - 1 &&
          (i < 5 || i == 9 || i == 14 || i == 19 || i == 29 || i == 49 ||
           i == 99)) {
        if ((rc = petstrip_get_rho(h, rho_inout))) return rc;
        if (output && output(user, i + 1, rho_inout, W, H))
          return fail(PETSTRIP_ECALLBACK, "output callback stopped the run");
      }
    }
    if ((rc = petstrip_get_rho(h, rho_inout))) return rc;
    if (output && output(user, (block + 1) * n_iter_in_block, rho_inout, W, H))
      return fail(PETSTRIP_ECALLBACK, "output callback stopped the run");
  }
  if (progress) progress(user, n_blocks * n_iter_in_block, 0);
  return PETSTRIP_OK;
}

int petstrip_sensitivity(petstrip_handle* h, int mode, float* out) {
  if (!h || !out || (mode != 0 && mode != 1))
    return fail(PETSTRIP_EINVAL, "bad argument");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  HIP_TRY(petstrip::launch_sensitivity(h->stream, h->sg, mode, h->d_out));
  HIP_TRY(hipMemcpyAsync(out, h->d_out, h->n_pix * sizeof(float),
                         hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return PETSTRIP_OK;
}

int petstrip_stats_get(petstrip_handle* h, petstrip_stats* out) {
  if (!h || !out) return fail(PETSTRIP_EINVAL, "bad argument");
  *out = h->stats;
  out->fused = h->fused_iter ? (h->banded ? (h->ncb > 1 ? 3 : 2) : 1) : 0;
  out->n_work_items = h->n_items;
  out->lds_bytes = (int64_t)h->lds_bytes;
  out->n_window_pixels = h->fused_iter ? h->window_px : 0;
  out->row_table_bytes = h->rowtab_on ? (int64_t)(h->n_sorted * h->rowtab_stride *
                                                  (int64_t)sizeof(petstrip::RowEntry))
                                      : 0;
  out->horner_walks = h->fused_iter && h->opts.pass1_walk == 0 ? 1 : 0;
  out->n_items_two_per_cu = 0;
  if (h->fused_iter && !h->banded) {
    const petstrip::ItemRanges& r = h->ranges;
    for (int i = 0; i < r.n; ++i)
      if (petstrip::fused_lds_bytes(r.max_ww[i], h->col_stride, petstrip::MODE_MLEM_K) <=
          petstrip::kTwoPerCuLds)
        out->n_items_two_per_cu += r.end[i] - r.begin[i];
  }
  return PETSTRIP_OK;
}

int petstrip_footprints(petstrip_handle* h, int32_t* n_box, int32_t* n_ell) {
  if (!h || (h->n_events > 0 && (!n_box || !n_ell)))
    return fail(PETSTRIP_EINVAL, "bad argument");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  if (h->n_events > 0) {
    HIP_TRY(hipMemcpyAsync(n_box, h->d_nbox, h->n_events * sizeof(int32_t),
                           hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(n_ell, h->d_nell, h->n_events * sizeof(int32_t),
                           hipMemcpyDeviceToHost, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return PETSTRIP_OK;
}

int petstrip_forward(petstrip_handle* h, float* lambda) {
  if (!h) return fail(PETSTRIP_EINVAL, "handle is NULL");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  const int64_t n = h->n_events;
  if (h->fused_ok) {
    petstrip::FusedArgs a = fused_args(h);
    HIP_TRY(petstrip::launch_fill(h->stream, h->d_lambda, 0.0f, n));
    HIP_TRY(hipMemsetAsync(h->d_scalars, 0, 8 * sizeof(double), h->stream));
    if ((rc = fused_forward(h, a, false))) return rc;
  } else {
    h->lambda_clean = false;
    HIP_TRY(petstrip::launch_denominator(
        h->stream, h->sg, h->sk, h->d_geom, n, h->d_rho, nullptr, h->d_lambda, h->d_inv,
        h->opts.kernel, h->opts.accumulation == PETSTRIP_ACCUM_COMPENSATED ? 1 : 0));
  }
  if (lambda && n > 0)
    HIP_TRY(hipMemcpyAsync(lambda, h->d_lambda, n * sizeof(float),
                           hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->lambda_valid = true;
  return PETSTRIP_OK;
}

namespace {

// Per-event-kernel form of backward (windows too large for LDS, or the
// per-event path forced).
int backward_per_event(petstrip_handle* h, const float* g, float* grad_rho,
                       double grad_sigma[2], double* ll_terms) {
  const int64_t n = h->n_events;
  int rc;
  if (!h->lambda_valid && (rc = petstrip_forward(h, nullptr))) return rc;
  HIP_TRY(hipMemsetAsync(h->d_scalars, 0, 8 * sizeof(double), h->stream));
  if (g) {
    if (n > 0)
      HIP_TRY(hipMemcpyAsync(h->d_inv, g, n * sizeof(float), hipMemcpyHostToDevice,
                             h->stream));
  } else {
    HIP_TRY(petstrip::launch_loglik_weights(h->stream, h->d_lambda, n, h->d_inv,
                                            h->d_scalars + 2));
  }
  if (grad_rho) {
    HIP_TRY(petstrip::launch_backproject(h->stream, h->sg, h->sk, h->d_geom, n, nullptr,
                                         h->d_inv, (double*)h->d_acc, h->opts.kernel));
    // the log-likelihood gradient subtracts the sensitivity term, 1 per pixel
    HIP_TRY(petstrip::launch_finalize(h->stream, h->d_acc, petstrip::GLOBAL_F64, 1.0,
                                      g ? 0.0f : -1.0f, 0, nullptr, h->d_out,
                                      (int64_t)h->n_pix));
    HIP_TRY(hipMemcpyAsync(grad_rho, h->d_out, h->n_pix * sizeof(float),
                           hipMemcpyDeviceToHost, h->stream));
  }
  if (grad_sigma)
    HIP_TRY(petstrip::launch_sigma_grad(h->stream, h->sg, h->sk, h->d_geom, n, h->d_rho,
                                        h->d_inv, h->d_scalars));
  double scal[8];
  HIP_TRY(hipMemcpyAsync(scal, h->d_scalars, sizeof(scal), hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  if (grad_sigma) {
    grad_sigma[0] = scal[0];
    grad_sigma[1] = scal[1];
  }
  *ll_terms = scal[2];
  return PETSTRIP_OK;
}

// log2 of the fixed-point unit of the values K / lambda_e (the
// log-likelihood gradient's, the FP32 MLEM iteration's): a typical value,
// 1 / (mean rho x mean ellipse pixels), at 2^typical_bits units.
int grad_unit_bits(const petstrip_handle* h, int typical_bits) {
  const double mean_rho = h->rho_sum / (double)std::max<size_t>(h->n_pix, 1);
  const double mean_ell = (double)h->stats.n_ellipse_pixels /
                          (double)std::max<int64_t>(h->stats.n_events, 1);
  if (!(mean_rho > 0.0) || !(mean_ell > 0.0)) return 26;
  const int k = (int)std::floor(typical_bits + std::log2(mean_rho * mean_ell));
  return std::max(-60, std::min(60, k));
}

// The log-likelihood gradient (g = 1/lambda) in one GRAD launch: lambda,
// sigma derivative and log-likelihood terms in pass 1, sum_e K / lambda_e in
// pass 2 with the row setups kept from pass 1; grad = that - 1.
int gradient_fused(petstrip_handle* h, float* grad_rho, double grad_sigma[2],
                   double* ll_terms) {
  const int64_t n = h->n_events;
  petstrip::FusedArgs a = fused_args(h);
  a.acc_kind = petstrip::GLOBAL_F64;
  const int k = grad_unit_bits(h);
  a.log2_scale = (float)k;
  a.split = 0;
  a.inv_fixed_scale = std::ldexp(1.0, -k);
  HIP_TRY(petstrip::launch_fill(h->stream, h->d_lambda, 0.0f, n));
  HIP_TRY(hipMemsetAsync(h->d_scalars, 0, 8 * sizeof(double), h->stream));
  HIP_TRY(hipEventRecord(h->evk0, h->stream));
  if (h->banded) {
    int rc;
    if ((rc = banded_pipeline(
             h, a, petstrip::MODE_FORWARD, true, petstrip::MODE_GRAD,
             petstrip::fused_lds_bytes(h->max_ww, h->col_stride, petstrip::MODE_GRAD, 2))))
      return rc;
  } else {
    HIP_TRY(petstrip::launch_strip_fused(h->stream, petstrip::MODE_GRAD, h->col_stride, 0,
                                         a, h->ranges));
  }
  HIP_TRY(hipEventRecord(h->evk1, h->stream));
  HIP_TRY(petstrip::launch_finalize(h->stream, h->d_acc, petstrip::GLOBAL_F64, 1.0, -1.0f,
                                    0, nullptr, h->d_out, (int64_t)h->n_pix));
  double scal[8];
  HIP_TRY(hipMemcpyAsync(scal, h->d_scalars, sizeof(scal), hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipMemcpyAsync(grad_rho, h->d_out, h->n_pix * sizeof(float),
                         hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->lambda_valid = true;
  if (grad_sigma) {
    grad_sigma[0] = scal[1];
    grad_sigma[1] = scal[2];
  }
  *ll_terms = scal[0];
  return PETSTRIP_OK;
}

// Fused form: one FORWARD pass (lambda, sigma derivative, log-likelihood
// terms, largest and mean |g| K) and one BACKWARD pass (K g_e into the image
// in signed fixed point).  For non-negative weights whose largest |g| K is
// within a factor 256 of the mean, the unit 2^-k puts the largest value just
// below 2^31 (one conversion per value; every sum is positive, so the
// absolute rounding is small against it).  Otherwise (weights of both
// signs, whose sums cancel, or an outlier) values may reach 2^62 / n_events
// (split conversion).  Either way no per-pixel sum can exceed 2^63.
int backward_fused(petstrip_handle* h, const float* g, float* grad_rho,
                   double grad_sigma[2], double* ll_terms) {
  const int64_t n = h->n_events;
  petstrip::FusedArgs a = fused_args(h);
  if (!g && grad_rho) return gradient_fused(h, grad_rho, grad_sigma, ll_terms);
  bool mixed_signs = false;
  if (g) {
    if (n > 0)
      HIP_TRY(hipMemcpyAsync(h->d_inv, g, n * sizeof(float), hipMemcpyHostToDevice,
                             h->stream));
    a.g = h->d_inv;
    for (int64_t i = 0; i < n && !mixed_signs; ++i) mixed_signs = g[i] < 0.0f;
  }
  HIP_TRY(petstrip::launch_fill(h->stream, h->d_lambda, 0.0f, n));
  HIP_TRY(hipMemsetAsync(h->d_scalars, 0, 8 * sizeof(double), h->stream));
  {
    int rc;
    if ((rc = fused_forward(h, a, true))) return rc;
  }
  double scal[8];
  HIP_TRY(hipMemcpyAsync(scal, h->d_scalars, sizeof(scal), hipMemcpyDeviceToHost,
                         h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  h->lambda_valid = true;
  if (grad_sigma) {
    grad_sigma[0] = scal[1];
    grad_sigma[1] = scal[2];
  }
  *ll_terms = scal[0];
  if (!grad_rho) return PETSTRIP_OK;
  const float bias = g ? 0.0f : -1.0f;
  const double mmax = scal[3], mmean = scal[4] / (double)std::max<int64_t>(n, 1);
  if (!(mmax > 0.0)) {
    HIP_TRY(petstrip::launch_fill(h->stream, h->d_out, bias, (int64_t)h->n_pix));
  } else {
    const bool split = mixed_signs || mmax > 256.0 * mmean || n >= ((int64_t)1 << 31);
    int k = split ? 62 - (int)std::ceil(std::log2((double)n * mmax))
                  : 31 - (int)std::ceil(std::log2(mmax));
    k = std::max(-100, std::min(100, k));
    a.log2_scale = (float)k;
    a.split = split ? 1 : 0;
    a.inv_fixed_scale = std::ldexp(1.0, -k);
    a.acc_kind = petstrip::GLOBAL_U64;
    HIP_TRY(hipEventRecord(h->evk0, h->stream));
    if (h->banded)
      HIP_TRY(petstrip::launch_strip_banded(
          h->stream, petstrip::MODE_BACKWARD, 2, a, h->n_items,
          petstrip::fused_lds_bytes(h->max_ww, h->col_stride, petstrip::MODE_BACKWARD)));
    else
      HIP_TRY(petstrip::launch_strip_fused(h->stream, petstrip::MODE_BACKWARD,
                                           h->col_stride, 0, a, h->ranges));
    HIP_TRY(hipEventRecord(h->evk1, h->stream));
    HIP_TRY(petstrip::launch_finalize(h->stream, h->d_acc, petstrip::GLOBAL_U64,
                                      a.inv_fixed_scale, bias, 0, nullptr, h->d_out,
                                      (int64_t)h->n_pix));
  }
  HIP_TRY(hipMemcpyAsync(grad_rho, h->d_out, h->n_pix * sizeof(float),
                         hipMemcpyDeviceToHost, h->stream));
  HIP_TRY(hipStreamSynchronize(h->stream));
  return PETSTRIP_OK;
}

}  // namespace

int petstrip_backward(petstrip_handle* h, const float* g, float* grad_rho,
                      double grad_sigma[2], double* loglik) {
  if (!h) return fail(PETSTRIP_EINVAL, "handle is NULL");
  if (loglik && g) return fail(PETSTRIP_EINVAL, "loglik needs g == NULL");
  if (grad_sigma && h->opts.kernel == PETSTRIP_KERNEL_ANALYTIC)
    return fail(PETSTRIP_EINVAL, "no sigma derivative for the analytic kernel");
  int rc;
  if ((rc = ensure_device(h))) return rc;
  double ll_terms = 0.0;
  rc = h->fused_ok ? backward_fused(h, g, grad_rho, grad_sigma, &ll_terms)
                   : backward_per_event(h, g, grad_rho, grad_sigma, &ll_terms);
  if (rc) return rc;
  if (loglik) {
    std::vector<float> rho(h->n_pix);
    if ((rc = petstrip_get_rho(h, rho.data()))) return rc;
    double srho = 0;
    for (float v : rho) srho += v;
    *loglik = ll_terms - srho;
  }
  return PETSTRIP_OK;
}

int petstrip_set_sigma(petstrip_handle* h, double sigma_z, double sigma_dl,
                       int freeze_support) {
  if (!h) return fail(PETSTRIP_EINVAL, "handle is NULL");
  if (!(sigma_z > 0) || !(sigma_dl > 0))
    return fail(PETSTRIP_EINVAL, "sigmas must be positive");
  h->sk = make_const(h->scanner, sigma_z, sigma_dl);
  h->lambda_valid = false;
  if (freeze_support) {
    int rc;
    if ((rc = ensure_device(h))) return rc;
    if (h->fused_ok && (rc = update_kernel_table(h))) return rc;
  } else {
    h->sg = h->sk;
    h->scanner.sigma_z = sigma_z;
    h->scanner.sigma_dl = sigma_dl;
    int rc;
    if ((rc = ensure_device(h))) return rc;
    return prepare_events(h);
  }
  return PETSTRIP_OK;
}

}  // extern "C"
