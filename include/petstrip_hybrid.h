/* 3D hybrid reconstruction on the GPU: the strip kernel walked per LOR of a
 * barrel scanner over a voxel grid (libpetstrip_hip.so).
 *
 * Replaces PET3D::Hybrid::GPU::Reconstruction::run
 * (grey277/j-pet-mlem src/3d/hybrid/cuda/reconstruction.h:31-58,
 * cuda/reconstruction.cu:31-160), the --gpu path of 3d_hybrid_reconstruction
 * (src/3d/hybrid/reconstruction_cmd.cpp:596-640).  The iteration is the
 * reference's CPU one (src/3d/hybrid/reconstruction.h:267-445): per event,
 * the voxels of its pixel-info range along the LOR times its plane range,
 * weight = strip kernel / frame sensitivity x LOR-pixel weight x rho,
 * denominator = sum of weight x sensitivity, rho' = sum over events of
 * weight / denominator.  The host-side preparation the reference does before
 * its GPU call is available here too: the LOR-pixel weights and sensitivity
 * (calculate_weight, calculate_sensitivity, normalize_geometry_weights,
 * reconstruction.h:212-265) on the host, the frame events (translate_to_frame
 * and the grid filter, reconstruction.h:97-161) on the device.
 *
 * Errors: every call returns 0 or a negative PETSTRIP_E* code with the
 * message in petstrip_last_error() (include/petstrip.h). */
#ifndef PETSTRIP_HYBRID_H
#define PETSTRIP_HYBRID_H

#include <stdint.h>

#include "petstrip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The reference's GeometrySOA (src/2d/barrel/geometry_soa.h:22-48,178-199)
 * and voxel grid (src/3d/geometry/voxel_grid.h).  LOR (first, second) with
 * first >= second has index first (first + 1) / 2 + second; its pixel infos
 * are [info_begin, info_end) with positions ascending.  The arrays are copied
 * at create. */
typedef struct petstrip_hybrid_geometry {
  int32_t n_detectors;
  int32_t n_columns, n_rows, n_planes;
  double pixel_size;
  double lower_left_x, lower_left_y; /* the pixel grid's lower-left corner */
  double z_left;                     /* z of the first plane's lower face   */
  int64_t n_pixel_infos;
  const float* segments;    /* n_lors x 4: start x, y, end x, y             */
  const float* widths;      /* n_lors                                        */
  const int64_t* info_begin;/* n_lors                                        */
  const int64_t* info_end;  /* n_lors                                        */
  const int16_t* pixels;    /* n_pixel_infos x 2: x, y                        */
  const float* positions;   /* n_pixel_infos: t along the LOR (0 at start)   */
  const float* weights;     /* n_pixel_infos, read with
                               PETSTRIP_HYBRID_GEOMETRY_WEIGHTS (may be NULL
                               otherwise)                                    */
} petstrip_hybrid_geometry;

enum {
  /* keep the geometry's weights (the reference's --system matrices) instead
   * of the Gaussian in the distance to the LOR (calculate_weight) */
  PETSTRIP_HYBRID_GEOMETRY_WEIGHTS = 1,
  /* sensitivity one (--sens-to-one) instead of the sum of the weights */
  PETSTRIP_HYBRID_SENSITIVITY_ONE = 2
};

/* A frame event (reconstruction.h:53-63).  end may exceed the LOR's last
 * pixel info by one, as the reference's binary search leaves it
 * (reconstruction.h:133-138); it is clamped to n_pixel_infos. */
typedef struct petstrip_hybrid_event {
  int16_t first, second;
  float up, right, tan, sec;
  int64_t begin, end;
  int16_t plane_begin, plane_end;
} petstrip_hybrid_event;

typedef struct petstrip_hybrid_stats {
  int64_t n_events;          /* frame events kept                          */
  int64_t n_voxels;          /* sum over events of pixel range x planes    */
  double last_iteration_ms;  /* device time of the last iteration          */
  double last_kernel_ms;     /* device time of its event kernel            */
} petstrip_hybrid_stats;

typedef struct petstrip_hybrid petstrip_hybrid;

/* Copies the geometry, computes the LOR-pixel weights and the W x H
 * sensitivity map as flags say (with the reference's float expressions, on
 * the host), allocates the voxel image (rho = 1).  sigma_z, sigma_dl: the
 * strip kernel's widths; length: the scintillator length of the frame
 * sensitivity. */
int petstrip_hybrid_create(const petstrip_hybrid_geometry* geometry, double sigma_z,
                           double sigma_dl, double length, int32_t flags, int32_t device,
                           petstrip_hybrid** out);
void petstrip_hybrid_destroy(petstrip_hybrid* h);

/* Responses (LOR, z_up, z_dn, dl): lors n x 2 (first, second, first >
 * second), zzd n x 3.  Frame events are formed and filtered on the device
 * and kept in input order; previous events are replaced. */
int petstrip_hybrid_set_responses(petstrip_hybrid* h, const int16_t* lors, const float* zzd,
                                  int64_t n);
/* Frame events formed by the caller (the reference's host code). */
int petstrip_hybrid_set_events(petstrip_hybrid* h, const petstrip_hybrid_event* events,
                               int64_t n);
int petstrip_hybrid_get_events(petstrip_hybrid* h, petstrip_hybrid_event* events);

/* Replaces the weights (n_pixel_infos) and/or the sensitivity map (rows x
 * columns), e.g. normalised by the caller; either may be NULL.  A map
 * replaces an external volume set below. */
int petstrip_hybrid_set_weights(petstrip_hybrid* h, const float* weights,
                                const float* sensitivity);
int petstrip_hybrid_get_weights(petstrip_hybrid* h, float* weights, float* sensitivity);
/* External 3D sensitivity, the command's --sensitivity
 * (src/3d/hybrid/reconstruction_cmd.cpp:162,505-519): depth = n_planes / 2
 * planes of rows x columns, voxel (x, y, p) at x + W (y + H p).  The
 * geometry weights are normalised by plane 0 (reconstruction.h:255-271) and
 * the denominator takes plane |iz - depth| for plane iz (reconstruction.h:
 * 357-362; at iz = 0, one plane past the reference's volume, plane depth - 1). */
int petstrip_hybrid_set_sensitivity_3d(petstrip_hybrid* h, const float* sensitivity,
                                       int32_t depth);

/* Per-plane LOR-pixel weights, the reference's multi-plane geometry
 * (GeometrySOA with n_planes_half weight planes, src/2d/barrel/
 * geometry_soa.h:44-58; one system matrix per half plane,
 * src/3d/hybrid/reconstruction_cmd.cpp:251-270).  raw_weights holds
 * n_planes_half planes of n_pixel_infos weights (plane-major).  The
 * sensitivity volume of n_planes_half planes is then
 *   sensitivity_mode 0: per (pixel, plane) the sum of that plane's weights
 *                       (calculate_sensitivity, reconstruction.h:236-250);
 *                    1: one (--sens-to-one);
 *                    2: the given volume (--sensitivity, n_planes_half
 *                       planes of rows x columns, x + W (y + H p)),
 * each plane's weights are divided by it (normalize_geometry_weights,
 * reconstruction.h:252-271), and the iteration takes for voxel plane iz the
 * weights and sensitivity of plane |iz - n_planes_half|
 * (reconstruction.h:345-356; |iz - n_planes_half| = n_planes_half, at
 * iz = 0, lies one plane past the reference's arrays: plane
 * n_planes_half - 1 is used).  The frame events' grid filter keeps reading
 * plane 0's weights (reconstruction.h:152-161).  n_planes_half = 1 sets the
 * geometry weights and the 2D map instead. */
int petstrip_hybrid_set_plane_weights(petstrip_hybrid* h, const float* raw_weights,
                                      int32_t n_planes_half, int32_t sensitivity_mode,
                                      const float* sensitivity);
/* The normalised weights and the sensitivity the iteration uses, as the
 * reference holds them after normalize_geometry_weights (GeometrySOA::
 * pixel_weights, plane-major; Reconstruction::sensitivity, planes of rows x
 * columns): n_weight_planes = the weight planes (1 for one set),
 * n_sensitivity_planes = the volume's depth (1 for the 2D map).  Either
 * array may be NULL (call with both NULL for the counts). */
int petstrip_hybrid_get_plane_weights(petstrip_hybrid* h, float* weights, float* sensitivity,
                                      int32_t* n_weight_planes,
                                      int32_t* n_sensitivity_planes);

/* --- system matrices (the --system option; host only) --------------------- */

/* A 2D barrel system matrix file (src/2d/barrel/sparse_matrix.h:43-67,
 * 163-238), full forms (PETP, TOFP); triangular forms fail as they do in the
 * reference's 3D hybrid command.  Elements are kept sorted by LOR, TOF
 * position and pixel (sort_by_lor_n_pixel). */
typedef struct petstrip_system_matrix petstrip_system_matrix;
int petstrip_system_matrix_read(const char* path, petstrip_system_matrix** out);
void petstrip_system_matrix_free(petstrip_system_matrix* m);
int petstrip_system_matrix_info(const petstrip_system_matrix* m, int32_t* n_pixels_in_row,
                                int32_t* n_detectors, int64_t* n_emissions,
                                int64_t* n_elements);
/* GeometrySOA::load_weights_from_matrix_file (geometry_soa.h:210-267): each
 * element's hits / emissions added to the geometry's pixel info of the same
 * LOR and pixel (none: dropped); weights (n_pixel_infos) are overwritten. */
int petstrip_system_matrix_weights(const petstrip_system_matrix* m,
                                   const petstrip_hybrid_geometry* g, float* weights);
/* Detector centres from a scanner description file, one detector per line
 * of "x y" vertex pairs (ScannerBuilder::deserialize, src/2d/barrel/
 * scanner_builder.h:21-35; Polygon::center, src/2d/geometry/polygon.h:56-68).
 * centres (nullable) receives 2 floats per detector, up to capacity. */
int petstrip_read_detector_centres(const char* path, float* centres, int32_t capacity,
                                   int32_t* n_detectors);
/* The geometry of a system matrix and detector centres (the GeometrySOA
 * constructors of geometry_soa.h:71-182): pixel infos of each LOR from its
 * matrix elements with weight hits / emissions, LORs of inactive detectors
 * empty, segment from the second detector's centre to the first's, pixel
 * positions along it (projection_scaled of the pixel centre on the grid
 * with the given lower-left corner), sorted.  Call with segments = NULL for
 * *n_pixel_infos, then with arrays of that many pixel infos and
 * n_detectors (n_detectors + 1) / 2 LORs. */
int petstrip_hybrid_geometry_from_matrix(const petstrip_system_matrix* m, const float* centres,
                                         int32_t n_columns, int32_t n_rows, double pixel_size,
                                         double lower_left_x, double lower_left_y,
                                         const int16_t* inactive, int32_t n_inactive,
                                         float* segments, float* widths, int64_t* info_begin,
                                         int64_t* info_end, int16_t* pixels, float* positions,
                                         float* weights, int64_t* n_pixel_infos);

/* The voxel image, planes x rows x columns (voxel index x + W (y + H z)). */
int petstrip_hybrid_set_rho(petstrip_hybrid* h, const float* rho);
int petstrip_hybrid_get_rho(petstrip_hybrid* h, float* rho);

/* n MLEM iterations. */
int petstrip_hybrid_iterate(petstrip_hybrid* h, int32_t n);

/* The reference's run(): n_blocks x n_iterations_in_block iterations from
 * rho_inout, skipping those below start_iteration; output(iteration, image)
 * after iterations 1-5, 10, 15, 20, 30, 50 and 100 of the first block and
 * after every block (image: planes x rows x columns, also left in
 * rho_inout); progress before and after every iteration and once at the end
 * (cuda/reconstruction.cu:103-181). */
typedef int (*petstrip_hybrid_output_cb)(void* user, int iteration, const float* image,
                                         int32_t width, int32_t height, int32_t depth);
int petstrip_hybrid_run(petstrip_hybrid* h, float* rho_inout, int32_t n_blocks,
                        int32_t n_iterations_in_block, int32_t start_iteration,
                        petstrip_hybrid_output_cb output, petstrip_progress_cb progress,
                        void* user);

int petstrip_hybrid_stats_get(petstrip_hybrid* h, petstrip_hybrid_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* PETSTRIP_HYBRID_H */
