// Fused MLEM iteration over spatially binned events (strip_fused.hip).
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
  int32_t pad;
};

enum { ACC_LDS_F32 = 0, ACC_LDS_FIXED64 = 1 };
enum { GLOBAL_F32 = 0, GLOBAL_F64 = 1, GLOBAL_U64 = 2 };

hipError_t launch_bin_count(hipStream_t st, const EventGeom* geom, int64_t n,
                            int bin_w, int bin_h, int nbx, int32_t* bin,
                            int32_t* count);
hipError_t launch_bin_scatter(hipStream_t st, const EventGeom* geom, int64_t n,
                              const int32_t* bin, int32_t* cursor,
                              EventGeom* sgeom, int32_t* sidx);
hipError_t launch_item_windows(hipStream_t st, const EventGeom* sgeom,
                               WorkItem* items, int n_items);
// LDS bytes per workgroup of `waves` waves (4, 8 or 16).
size_t fused_lds_bytes(int max_window_px, int acc_mode, int max_rows, int waves);
hipError_t launch_mlem_fused(hipStream_t st, const StripConst& sg,
                             const StripConst& sk, const EventGeom* sgeom,
                             const int32_t* sidx, const WorkItem* items,
                             int n_items, size_t lds_bytes, const float* rho,
                             float* lambda, void* acc_global,
                             int acc_global_kind, int acc_mode, int max_rows,
                             int compensated, int waves);
hipError_t launch_finalize(hipStream_t st, void* acc, int kind, float* out,
                           int64_t n);

}  // namespace petstrip
