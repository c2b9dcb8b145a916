// Harness around the reference's own 3D hybrid reconstruction
// (PET3D::Hybrid::Reconstruction, src/3d/hybrid/reconstruction.h), built
// from the reference tree by oracle/Makefile into _ref/libref_hybrid.so.
// Test infrastructure: the tests call it to pin the C restatement in
// hybrid_oracle.c; nothing in the product links it.
//
// The geometry is handed over as the arrays of a GeometrySOA (one LOR
// segment, width and pixel-info range per LOR; pixel, position and weight
// per pixel info); the harness builds the reference's Geometry from them,
// runs calculate_weight / calculate_sensitivity (or set_sensitivity_to_one)
// and normalize_geometry_weights as 3d_hybrid_reconstruction does
// (reconstruction_cmd.cpp:500-521), loads the responses through the
// reference's text stream reader (translate_to_frame and the grid filter) and
// runs the iterations (one OpenMP thread unless asked for more: with one,
// the accumulation order is the reference's sequential event order; the
// iterations' wall time is returned for the CPU baseline).
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <sstream>
#include <vector>

#if _OPENMP
#include <omp.h>
#endif

#include "2d/barrel/generic_scanner.h"
#include "2d/barrel/sparse_matrix.h"
#include "common/types.h"
#include "2d/barrel/geometry.h"
#include "2d/barrel/scanner_builder.h"
#include "2d/barrel/square_detector.h"
#include "2d/strip/gaussian_kernel.h"
#include "3d/hybrid/reconstruction.h"
#include "3d/hybrid/scanner.h"

namespace {
using F = float;
using S = short;
using Scanner2D = PET2D::Barrel::GenericScanner<PET2D::Barrel::SquareDetector<F>, S>;
using Scanner = PET3D::Hybrid::Scanner<Scanner2D>;
using Kernel = PET2D::Strip::GaussianKernel<F>;
using Reconstruction = PET3D::Hybrid::Reconstruction<Scanner, Kernel>;
using Geometry = PET2D::Barrel::Geometry<F, S>;
using LOR = PET2D::Barrel::LOR<S>;
using Point2D = PET2D::Point<F>;
using Pixel = PET2D::Pixel<S>;
using Matrix = PET2D::Barrel::SparseMatrix<Pixel, LOR, Hit>;
}  // namespace

extern "C" {

// Frame event fields out: lor first, second, up, right, tan, sec, pixel
// info begin, end, plane begin, end (10 doubles per kept event).
int ref_hybrid_run(int n_detectors, int n_columns, int n_rows, float pixel_size,
                   float lower_left_x, float lower_left_y, float z_left, int n_planes,
                   const float* segments, const float* widths, const int64_t* info_begin,
                   const int64_t* info_end, const int16_t* pixels, const float* positions,
                   const float* weights, float sigma_z, float sigma_dl, float length,
                   int sensitivity_one, const int16_t* lors, const float* zzd, int64_t n_events,
                   int n_iterations, float* out_weights, float* out_sensitivity,
                   int64_t* out_n_kept, double* out_frames, float* out_rho, int n_threads,
                   double* iteration_seconds, const float* sensitivity_3d) {
  try {
#if _OPENMP
    omp_set_num_threads(n_threads > 0 ? n_threads : 1);
#endif
    Geometry::Grid grid(n_columns, n_rows, pixel_size, Point2D(lower_left_x, lower_left_y));
    Geometry geometry(n_detectors, grid);
    for (S first = 0; first < n_detectors; ++first) {
      for (S second = 0; second <= first; ++second) {
        LOR lor(first, second);
        const auto i = lor.index();
        auto& lg = geometry[lor];
        lg.lor = lor;
        lg.segment = PET2D::LineSegment<F>(Point2D(segments[4 * i], segments[4 * i + 1]),
                                           Point2D(segments[4 * i + 2], segments[4 * i + 3]));
        lg.width = widths[i];
        lg.pixel_infos.clear();
        for (int64_t k = info_begin[i]; k < info_end[i]; ++k) {
          Geometry::PixelInfo pi;
          pi.pixel = PET2D::Pixel<S>(pixels[2 * k], pixels[2 * k + 1]);
          pi.t = positions[k];
          pi.distance = 0;
          pi.fill = 0;
          pi.weight = weights[k];
          lg.pixel_infos.push_back(pi);
        }
      }
    }
    Reconstruction::Geometry soa(geometry, 1);
    Scanner2D barrel = PET2D::Barrel::ScannerBuilder<Scanner2D>::build_single_ring(
        F(1), 4, F(0.1), F(0.1));
    Scanner scanner(barrel, length);
    scanner.set_sigmas(sigma_z, sigma_dl);
    Reconstruction::Grid grid3d(grid, z_left, (S)n_planes);
    // an external volume: n_planes / 2 planes of W x H, as the command's
    // --sensitivity reads it (reconstruction_cmd.cpp:499-519)
    Reconstruction rec(scanner, grid3d, soa, sensitivity_3d != nullptr);
    rec.calculate_weight();
    if (sensitivity_3d)
      std::memcpy(rec.sensitivity.data, sensitivity_3d,
                  sizeof(F) * (size_t)n_columns * n_rows * (n_planes / 2));
    else if (sensitivity_one)
      rec.set_sensitivity_to_one();
    else
      rec.calculate_sensitivity();
    rec.normalize_geometry_weights();
    std::memcpy(out_weights, soa.pixel_weights, sizeof(F) * soa.n_pixel_infos);
    std::memcpy(out_sensitivity, rec.sensitivity.data,
                sizeof(F) * (size_t)n_columns * n_rows);
    // responses through the reference's stream reader (text: "first second
    // z_up z_dn dl", nine significant digits, which read back to the same
    // floats)
    std::stringstream text;
    text.precision(9);
    for (int64_t e = 0; e < n_events; ++e)
      text << lors[2 * e] << ' ' << lors[2 * e + 1] << ' ' << zzd[3 * e] << ' '
           << zzd[3 * e + 1] << ' ' << zzd[3 * e + 2] << '\n';
    rec << text;
    *out_n_kept = rec.n_events();
    for (int i = 0; i < rec.n_events(); ++i) {
      const auto ev = rec.frame_event(i);
      double* o = out_frames + 10 * (size_t)i;
      o[0] = ev.lor.first;
      o[1] = ev.lor.second;
      o[2] = ev.up;
      o[3] = ev.right;
      o[4] = ev.tan;
      o[5] = ev.sec;
      o[6] = (double)ev.pixel_info_begin;
      o[7] = (double)ev.pixel_info_end;
      o[8] = ev.plane_begin;
      o[9] = ev.plane_end;
    }
    const size_t n_vox = (size_t)n_columns * n_rows * n_planes;
    double seconds = 0;
    for (int it = 0; it < n_iterations; ++it) {
      const auto t0 = std::chrono::steady_clock::now();
      rec();
      seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (out_rho) std::memcpy(out_rho + (size_t)it * n_vox, rec.rho.data, sizeof(F) * n_vox);
    }
    if (iteration_seconds) *iteration_seconds = seconds;
    return 0;
  } catch (...) {
    return -1;
  }
}

// The reference's --system paths (reconstruction_cmd.cpp:241-450):
//   mode 0: a geometry (arrays as ref_hybrid_run) with n_matrices system
//           matrix files, one or n_planes / 2, loaded into the weight planes
//           (GeometrySOA(geometry, n_matrices),
//           load_weights_from_matrix_file);
//   mode 1: the geometry of one matrix file and the detector centres of a
//           scanner description file (ScannerBuilder::deserialize), LORs of
//           the inactive detectors left out (GeometrySOA(matrix, centres,
//           grid, inactive)); the grid is the matrix's pixels of
//           pixel_size, centred.
// Then, as run_reconstruction does with --system (no calculate_weight): the
// sensitivity (sens_mode 0 calculated, 1 one, 2 the given volume of
// n_planes / 2 planes), normalize_geometry_weights, the responses through
// the stream reader and n_iterations iterations with one thread.
// Outputs: the geometry arrays (mode 1; out_n_infos = the reference's
// n_pixel_infos, which counts the matrix's elements, skipped ones
// included), the weights of every plane, the sensitivity (depth planes),
// frames and images as ref_hybrid_run.
int ref_hybrid_run_system(int mode, int n_detectors, int n_columns, int n_rows,
                          float pixel_size, float lower_left_x, float lower_left_y,
                          const float* segments, const float* widths,
                          const int64_t* info_begin, const int64_t* info_end,
                          const int16_t* pixels, const float* positions,
                          const char* const* matrix_paths, int n_matrices,
                          const char* detector_file, const int16_t* inactive, int n_inactive,
                          float z_left, int n_planes, float sigma_z, float sigma_dl,
                          float length, int sens_mode, const float* sensitivity_3d,
                          const int16_t* lors, const float* zzd, int64_t n_events,
                          int n_iterations, int64_t* out_n_infos, int32_t* out_depth,
                          float* out_segments, int64_t* out_begin, int64_t* out_end,
                          int16_t* out_pixels, float* out_positions, float* out_weights,
                          float* out_sensitivity, int64_t* out_n_kept, double* out_frames,
                          float* out_rho) {
  try {
#if _OPENMP
    omp_set_num_threads(1);
#endif
    Reconstruction::Geometry* soa = nullptr;
    Geometry::Grid* grid = nullptr;
    Scanner2D barrel = PET2D::Barrel::ScannerBuilder<Scanner2D>::build_single_ring(
        F(1), 4, F(0.1), F(0.1));
    if (mode == 0) {
      grid = new Geometry::Grid(n_columns, n_rows, pixel_size,
                                Point2D(lower_left_x, lower_left_y));
      Geometry geometry(n_detectors, *grid);
      for (S first = 0; first < n_detectors; ++first) {
        for (S second = 0; second <= first; ++second) {
          LOR lor(first, second);
          const auto i = lor.index();
          auto& lg = geometry[lor];
          lg.lor = lor;
          lg.segment = PET2D::LineSegment<F>(
              Point2D(segments[4 * i], segments[4 * i + 1]),
              Point2D(segments[4 * i + 2], segments[4 * i + 3]));
          lg.width = widths[i];
          lg.pixel_infos.clear();
          for (int64_t k = info_begin[i]; k < info_end[i]; ++k) {
            Geometry::PixelInfo pi;
            pi.pixel = Pixel(pixels[2 * k], pixels[2 * k + 1]);
            pi.t = positions[k];
            pi.distance = 0;
            pi.fill = 0;
            pi.weight = 0;
            lg.pixel_infos.push_back(pi);
          }
        }
      }
      soa = new Reconstruction::Geometry(geometry, (size_t)n_matrices);
      for (int p = 0; p < n_matrices; ++p)
        soa->load_weights_from_matrix_file<Hit>(matrix_paths[p], (size_t)p);
    } else {
      util::ibstream in_matrix(matrix_paths[0]);
      if (!in_matrix.is_open()) return -1;
      Matrix matrix(in_matrix);
      std::ifstream in_dets(detector_file);
      barrel = PET2D::Barrel::ScannerBuilder<Scanner2D>::deserialize(in_dets);
      grid = new Geometry::Grid(matrix.n_pixels_in_row(), matrix.n_pixels_in_row(),
                                pixel_size);
      soa = new Reconstruction::Geometry(matrix, barrel.detector_centers(), *grid,
                                         n_inactive ? inactive : nullptr,
                                         (size_t)n_inactive);
      // The constructor sizes its arrays by the matrix's element count but
      // fills one slot per element of an active LOR (geometry_soa.h:81-115);
      // the slots of skipped elements keep default-initialised pixels
      // (Pixel() = default under new Pixel[n]) that normalize_geometry_weights
      // (reconstruction.h:253-272) then uses as sensitivity indices. Those
      // slots are given pixel (0, 0) and weight 0 here so that the run is
      // defined; no LOR's range reaches them.
      size_t filled = 0;
      for (size_t l = 0; l < soa->n_lors; ++l)
        filled = std::max(filled, (size_t)soa->lor_pixel_info_end[l]);
      for (size_t i = filled; i < soa->n_pixel_infos; ++i) {
        soa->pixels[i] = Pixel(0, 0);
        soa->pixel_positions[i] = 0;
        soa->pixel_weights[i] = 0;
      }
    }
    *out_n_infos = (int64_t)soa->n_pixel_infos;
    if (mode == 1 && out_segments) {
      for (size_t l = 0; l < soa->n_lors; ++l) {
        out_segments[4 * l] = soa->lor_line_segments[l].start.x;
        out_segments[4 * l + 1] = soa->lor_line_segments[l].start.y;
        out_segments[4 * l + 2] = soa->lor_line_segments[l].end.x;
        out_segments[4 * l + 3] = soa->lor_line_segments[l].end.y;
        out_begin[l] = (int64_t)soa->lor_pixel_info_begin[l];
        out_end[l] = (int64_t)soa->lor_pixel_info_end[l];
      }
      for (size_t i = 0; i < soa->n_pixel_infos; ++i) {
        out_pixels[2 * i] = soa->pixels[i].x;
        out_pixels[2 * i + 1] = soa->pixels[i].y;
        out_positions[i] = soa->pixel_positions[i];
      }
    }
    if (!out_weights) {  // sizes only
      delete soa;
      delete grid;
      return 0;
    }
    Scanner scanner(barrel, length);
    scanner.set_sigmas(sigma_z, sigma_dl);
    Reconstruction::Grid grid3d(*grid, z_left, (S)n_planes);
    Reconstruction rec(scanner, grid3d, *soa, sensitivity_3d != nullptr);
    const size_t np = (size_t)grid->n_columns * grid->n_rows;
    const int depth = rec.sensitivity.depth;
    *out_depth = depth;
    if (sens_mode == 2 && sensitivity_3d)
      std::memcpy(rec.sensitivity.data, sensitivity_3d, sizeof(F) * np * depth);
    else if (sens_mode == 1)
      rec.set_sensitivity_to_one();
    else
      rec.calculate_sensitivity();
    rec.normalize_geometry_weights();
    std::memcpy(out_weights, soa->pixel_weights,
                sizeof(F) * soa->n_pixel_infos * soa->n_planes_half);
    std::memcpy(out_sensitivity, rec.sensitivity.data, sizeof(F) * np * depth);
    std::stringstream text;
    text.precision(9);
    for (int64_t e = 0; e < n_events; ++e)
      text << lors[2 * e] << ' ' << lors[2 * e + 1] << ' ' << zzd[3 * e] << ' '
           << zzd[3 * e + 1] << ' ' << zzd[3 * e + 2] << '\n';
    rec << text;
    *out_n_kept = rec.n_events();
    for (int i = 0; i < rec.n_events(); ++i) {
      const auto ev = rec.frame_event(i);
      double* o = out_frames + 10 * (size_t)i;
      o[0] = ev.lor.first;
      o[1] = ev.lor.second;
      o[2] = ev.up;
      o[3] = ev.right;
      o[4] = ev.tan;
      o[5] = ev.sec;
      o[6] = (double)ev.pixel_info_begin;
      o[7] = (double)ev.pixel_info_end;
      o[8] = ev.plane_begin;
      o[9] = ev.plane_end;
    }
    const size_t n_vox = np * (size_t)n_planes;
    for (int it = 0; it < n_iterations; ++it) {
      rec();
      std::memcpy(out_rho + (size_t)it * n_vox, rec.rho.data, sizeof(F) * n_vox);
    }
    delete soa;
    delete grid;
    return 0;
  } catch (...) {
    return -1;
  }
}

// The reference's own reader of a binary geometry file (Geometry(ibstream),
// geometry.h:55-83) into GeometrySOA arrays (geometry_soa.h:178-199):
// sizes first (call with null arrays), then the arrays.
int ref_hybrid_read_geometry(const char* path, int* n_detectors, int* n_columns,
                             int* n_rows, float* grid3, int64_t* n_pixel_infos,
                             float* segments, float* widths, int64_t* info_begin,
                             int64_t* info_end, int16_t* pixels, float* positions,
                             float* weights) {
  try {
    util::ibstream in(path);
    if (!in.is_open()) return -1;
    Geometry geometry(in);
    Reconstruction::Geometry soa(geometry, 1);
    *n_detectors = geometry.n_detectors;
    *n_columns = geometry.grid.n_columns;
    *n_rows = geometry.grid.n_rows;
    grid3[0] = geometry.grid.pixel_size;
    grid3[1] = geometry.grid.lower_left.x;
    grid3[2] = geometry.grid.lower_left.y;
    *n_pixel_infos = (int64_t)soa.n_pixel_infos;
    if (!segments) return 0;
    for (size_t l = 0; l < soa.n_lors; ++l) {
      segments[4 * l] = soa.lor_line_segments[l].start.x;
      segments[4 * l + 1] = soa.lor_line_segments[l].start.y;
      segments[4 * l + 2] = soa.lor_line_segments[l].end.x;
      segments[4 * l + 3] = soa.lor_line_segments[l].end.y;
      widths[l] = soa.lor_widths[l];
      info_begin[l] = (int64_t)soa.lor_pixel_info_begin[l];
      info_end[l] = (int64_t)soa.lor_pixel_info_end[l];
    }
    for (size_t i = 0; i < soa.n_pixel_infos; ++i) {
      pixels[2 * i] = soa.pixels[i].x;
      pixels[2 * i + 1] = soa.pixels[i].y;
      positions[i] = soa.pixel_positions[i];
      weights[i] = soa.pixel_weights[i];
    }
    return 0;
  } catch (...) {
    return -1;
  }
}

}  // extern "C"
