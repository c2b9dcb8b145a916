// Fused passes over spatially binned events (strip_fused.hip) and the
// binning that feeds them (strip_bins.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "strip_geometry.h"

namespace petstrip {

// A chunk of sorted events whose boxes share one spatial bin, and the image
// window covering all their boxes.
struct WorkItem {
  int32_t begin, end;  // range in the sorted event arrays
  int32_t wx0, wy0;    // window origin (column, row)
  int32_t ww, wh;      // window size
  int32_t max_rows;    // largest box height in the chunk
  int32_t pad;         // banded items: the row band index (rows / kBandRows)
};

// A box row's walk in the row table: iv = (ellipse columns << 16) | (first
// ellipse column - the box's first column), 0 for a row without ellipse
// pixels; u0 (u at the first ellipse column), lsc (log2 of the row factor)
// and b1 (the chain's per-pixel ratio) for the kernel constants the table
// was built with.
struct __align__(16) RowEntry {
  uint32_t iv;
  float u0, lsc, b1;
};

enum { GLOBAL_F32 = 0, GLOBAL_F64 = 1, GLOBAL_U64 = 2 };
// MODE_MLEM_K: the MLEM iteration with pass-2 values K / lambda_e (no rho),
// the image then multiplied by rho per pixel (rho is constant within an
// iteration, so sum_e K rho / lambda_e = rho sum_e K / lambda_e); MODE_MLEM
// adds K rho / lambda_e per value (the fixed 2^-40 format of the
// deterministic accumulation and the compensated form keep it).
// MODE_MLEM_KA: MODE_MLEM_K with the analytic kernel (the reference --gpu
// iteration's, analytic_kernel.h:28-62) evaluated per pixel in the walks.
enum {
  MODE_MLEM = 0,
  MODE_FORWARD = 1,
  MODE_BACKWARD = 2,
  MODE_GRAD = 3,
  MODE_MLEM_K = 4,
  MODE_MLEM_KA = 5
};

struct FusedArgs {
  StripConst sg;             // constants the ellipse/box support was built with
  StripConst sk;             // constants of the kernel value
  const EventGeom* sgeom;    // sorted event geometry
  const int32_t* sidx;       // original index of each sorted event
  const WorkItem* items;
  const float* rho;          // W*H image (MLEM, FORWARD)
  const float* rho_weight;   // W*H or null: pass 1 reads rho times it (the
                             // sensitivity-weighted denominator); MLEM_K(A) only
  float* lambda;             // per-event denominators (written by MLEM/FORWARD)
  const float* g;            // upstream weights per event, or null for 1/lambda
  void* acc;                 // W*H global accumulator (MLEM, BACKWARD)
  int acc_kind;              // GLOBAL_*
  float log2_scale;          // k: pass-2 values are added in units 2^-k (the
                             // K / lambda modes: each work item its own k,
                             // this one for windows whose rho sums to zero)
  float log2_ell;            // log2 of the mean ellipse pixels per event
  int split;                 // BACKWARD: values up to 2^62 (else below 2^31)
  double inv_fixed_scale;    // 2^-k, the unit when flushing to floats
  int row_scan;              // 1: every row interval by the scanning path
  const RowEntry* rowtab;    // unbanded walks: the row table (per sorted event
                             // rowtab_stride entries, a multiple of 64, in the
                             // walks' row order: k_row_table), or null
  int64_t rowtab_stride;
  int rowtab_kernel;         // 1: the entries' u0, lsc, b1 hold for sk and du
  int horner;                // 1: pass 1 by Horner's rule (forward_rows_h),
                             // 0: by the chain (forward_rows)
  float s_pix;               // sqrt(sk.isz log2 e): u = pq s_pix
  float du;                  // sg.pw s_pix, the u step per column
  const float* gtab;         // G(k) = 2^(-(k du)^2) for k < the image width + 8
  const float* mtab;         // (k, k^2) pairs for the same k
  float* part;               // banded: per sorted event and band, (lambda,
                             // sigma_z sum, sigma_dl sum, largest row factor)
  int nbands;                // banded: band slots per event (row bands x ncb)
  int ncb;                   // banded: column-band slots per event (1: none)
  int cshift;                // banded: log2 of the column-band width (30: none)
  float* lam_sorted;         // banded: lambda per sorted event
  double* scalars;           // FORWARD: {sum log lambda, dl/dsz, dl/dsdl,
                             //           max_e |g_e| K_e, sum_e |g_e| K_e}
                             // with K_e the event's largest kernel value
  double* scal_part;         // FORWARD, GRAD: the same five per workgroup,
                             // kScalStride doubles apart, added to scalars
                             // by k_scalars_finish in a fixed order
};

constexpr int kScalStride = 8;       // doubles per workgroup in scal_part
constexpr int kCombineBlocks = 1024; // workgroups of k_band_combine (grid-stride)

// Bin of an event: (height class, width class) x (spatial bin of its box
// centre), the height classes in the order of boxes taller than 32 rows,
// 17-32 rows, at most 16, each with the boxes at most wide_cols columns wide
// first, then the wider ones (n_spatial spatial bins per class pair); empty
// boxes get -1.
constexpr int kHeightClasses = 3;
constexpr int kWidthClasses = 2;
hipError_t launch_bin_count(hipStream_t st, const EventGeom* geom, int64_t n,
                            int bin_w, int bin_h, int nbx, int n_spatial, int wide_cols,
                            int32_t* bin, int32_t* count);
hipError_t launch_bin_scatter(hipStream_t st, const EventGeom* geom, int64_t n,
                              const int32_t* bin, int32_t* cursor,
                              EventGeom* sgeom, int32_t* sidx);
hipError_t launch_item_windows(hipStream_t st, const EventGeom* sgeom,
                               WorkItem* items, int n_items);

constexpr int kWaves = 16;           // waves per workgroup of the fused kernel
constexpr int kBandRows = 64;        // rows per band of the banded form
constexpr int kColBandShift = 7;     // columns per column band: 128 (windows
                                     // too wide for LDS even in row bands)
constexpr int kMaxColStride = 256;   // largest LDS column stride (window height)
constexpr int kItemEvents = 1024;    // most events in a work item
// The iteration's K / lambda sums (MODE_MLEM_K, MODE_MLEM_KA) accumulate in
// 32-bit LDS windows: in a work item of n events (n <= kItemEvents), values
// below 2^(kLog2Sum32 - log2 n) units, at most n of them per window pixel,
// stay below 2^32; rows with larger values add to the global image
// directly.  Each item's unit puts a typical value of its window
// kItemHeadroomBits below that cap (2^12.9 units at n = 1024: values up to
// 512 times the typical one take the window; a pixel's N rounding errors,
// independent and of mean zero, leave its sum within ~4e-5 / sqrt(N) of the
// typical value's relative size).
constexpr float kLog2Sum32 = 31.9f;
constexpr float kItemHeadroomBits = 9.0f;
constexpr int kLdsHead = 128;        // LDS bytes before the windows (two
                                     // floats per wave)
// LDS per workgroup at which two 16-wave workgroups share a CU's 160 KB
// (less a 512-byte allocation granule)
constexpr size_t kTwoPerCuLds = 80 * 1024 - 512;

// n_box and n_ell per event from the row intervals of the fused walks
// (valid when rows are hole-free intervals); scan = 1 forces the scanning
// path for every row.
hipError_t launch_row_counts(hipStream_t st, const StripConst& sg, const EventGeom* geom,
                             int64_t n, int scan, int32_t* nbox, int32_t* nell);

// The row table for n sorted events, stride entries per event (a multiple
// of 64 at least the tallest box); scan as for launch_row_counts.
hipError_t launch_row_table(hipStream_t st, const StripConst& sg, const StripConst& sk,
                            float s_pix, float du, const EventGeom* sgeom, int64_t n, int scan,
                            int64_t stride, RowEntry* tab);

// LDS bytes per workgroup for windows of at most max_ww columns with column
// stride col_stride; band = 2 (pass 2 of the banded form) holds the rho
// window only for MODE_MLEM.
size_t fused_lds_bytes(int max_ww, int col_stride, int mode, int band = 0);
// Column stride for windows of at most max_wh rows: the next multiple of
// 64, or 0 when above kMaxColStride.
int fused_col_stride(int max_wh);
// The items of an unbanded launch in ranges, each of one box height class
// and one window width class: pack = 1, 2 or 4 for boxes taller than 32
// rows, 17-32 rows or at most 16 (MLEM, MLEM_K, MLEM_KA and GRAD then walk
// two or four of their events per wave at a time; the other modes one),
// max_ww the range's widest window, which sizes its LDS (ranges of narrow
// windows can run two workgroups per CU).  The ranges cover items 0 ..
// n_items - 1 in order.
constexpr int kMaxRanges = kHeightClasses * kWidthClasses;
struct ItemRanges {
  int n = 0;
  int n_items = 0;
  int begin[kMaxRanges] = {}, end[kMaxRanges] = {}, pack[kMaxRanges] = {},
      max_ww[kMaxRanges] = {};
};
// MODE_FORWARD and MODE_GRAD also add their five scalars to args.scalars,
// through args.scal_part (n_items * kScalStride doubles).
hipError_t launch_strip_fused(hipStream_t st, int mode, int col_stride, int compensated,
                              const FusedArgs& args, const ItemRanges& ranges);
// Banded form (column stride kBandRows): band = 1 runs pass 1 into the band
// partials (mode MODE_FORWARD with sigma sums, else lambda only), band = 2
// runs pass 2 of mode (MLEM, GRAD or BACKWARD) with the combined lambda.
hipError_t launch_strip_banded(hipStream_t st, int mode, int band, const FusedArgs& args,
                               int n_items, size_t lds_bytes);
// Band partials -> lambda (sorted and original order); sigma = 1 also adds
// the log-likelihood terms, weighted sigma derivatives, max and sum of |g| K
// to the scalars (as the FORWARD mode does), through args.scal_part
// (kCombineBlocks * kScalStride doubles).
hipError_t launch_band_combine(hipStream_t st, const FusedArgs& args, int64_t n_sorted,
                               int sigma);

// out = acc * unit (* mul, when given) + bias (acc read as kind; 64-bit
// values as signed), the accumulator cleared; *overflow (nullable) set to 1
// if a 64-bit value is negative where only non-negative sums are expected
// (check_sign); *clear (nullable) set to 0.
hipError_t launch_finalize(hipStream_t st, void* acc, int kind, double unit,
                           float bias, int check_sign, int* overflow, float* out,
                           int64_t n, const float* mul = nullptr,
                           unsigned long long* clear = nullptr);

}  // namespace petstrip
