/* CPU oracle of the 3D hybrid reconstruction (test infrastructure only:
 * tests/, smoke() and bench.py's CPU leg call it as the checker; nothing in
 * the product links it).  A restatement of PET3D::Hybrid::Reconstruction
 * (grey277/j-pet-mlem src/3d/hybrid/reconstruction.h) in its default mode:
 * LOR-pixel weights from the Gaussian in the distance to the LOR, the 2D
 * sensitivity map from those weights (or one), normalised weights, frame
 * events from (LOR, z_up, z_dn, dl) responses, and the list-mode MLEM
 * iteration over the voxels of each event's pixel range and plane range.
 * The FP32 functions follow the reference's float expressions without
 * contraction (the reference harness is built with -ffp-contract=off) and
 * are bitwise equal to it; the FP64 twin computes the same in doubles. */
#ifndef HYBRID_ORACLE_H
#define HYBRID_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* GeometrySOA arrays (src/2d/barrel/geometry_soa.h): per LOR index
 * first (first + 1) / 2 + second: segment (start x, y, end x, y), width,
 * pixel-info range; per pixel info: pixel (x, y), position t along the LOR.
 * The voxel grid: n_columns x n_rows pixels of pixel_size from the lower
 * left corner, n_planes planes from z_left. */
typedef struct {
  int32_t n_detectors, n_columns, n_rows, n_planes;
  double pixel_size, lower_left_x, lower_left_y, z_left;
  int64_t n_pixel_infos;
  const float* segments;
  const float* widths;
  const int64_t* info_begin;
  const int64_t* info_end;
  const int16_t* pixels;
  const float* positions;
} oracle_hybrid_geometry;

/* A frame event (reconstruction.h:53-63). */
typedef struct {
  int16_t first, second;
  float up, right, tan, sec;
  int64_t begin, end;
  int16_t plane_begin, plane_end;
} oracle_hybrid_event;

/* calculate_weight, calculate_sensitivity (or set_sensitivity_to_one) and
 * normalize_geometry_weights (reconstruction.h:212-265): weights
 * (n_pixel_infos) and the W x H sensitivity map.  sensitivity_one = 2 takes
 * the map as given (plane 0 of an external sensitivity volume) and only
 * normalises the weights by it. */
void oracle_hybrid_weights_f32(const oracle_hybrid_geometry* g, int sensitivity_one,
                               float* weights, float* sensitivity);
void oracle_hybrid_weights_f64(const oracle_hybrid_geometry* g, int sensitivity_one,
                               double* weights, double* sensitivity);

/* translate_to_frame and the grid filter (reconstruction.h:97-161) for n
 * responses (lors: first, second; zzd: z_up, z_dn, dl); returns the number of
 * events kept, written to out in order. */
int64_t oracle_hybrid_frames_f32(const oracle_hybrid_geometry* g, double sigma_z,
                                 double sigma_dl, const float* weights, const int16_t* lors,
                                 const float* zzd, int64_t n, oracle_hybrid_event* out);

/* n_iter iterations (reconstruction.h:267-445) from rho (W x H x planes);
 * the image after every iteration goes to out (n_iter x voxels).
 * sens_depth = 0: sensitivity is the W x H map of the pixel (reconstruction.h
 * :364-367); sens_depth = D > 0: it is an external volume of D planes of
 * W x H (D = planes / 2) read at plane |iz - D| (reconstruction.h:357-362),
 * where plane D (iz = 0, one plane past the volume) reads plane D - 1. */
void oracle_hybrid_mlem_f32(const oracle_hybrid_geometry* g, double sigma_z,
                            double sigma_dl, double length, const float* weights,
                            const float* sensitivity, int sens_depth,
                            const oracle_hybrid_event* ev, int64_t n, int n_iter,
                            const float* rho, float* out);
void oracle_hybrid_mlem_f64(const oracle_hybrid_geometry* g, double sigma_z,
                            double sigma_dl, double length, const double* weights,
                            const double* sensitivity, int sens_depth,
                            const oracle_hybrid_event* ev, int64_t n, int n_iter,
                            const double* rho, double* out);

/* The iteration with per-plane weights (weight_planes x n_pixel_infos,
 * plane-major), voxel plane iz reading weight plane |iz - D| as it reads the
 * sensitivity volume's (D = sens_depth = weight_planes; the reference's
 * multi-plane geometry, reconstruction.h:345-356); weight_planes = 1 is
 * oracle_hybrid_mlem_*. */
void oracle_hybrid_mlem_planes_f32(const oracle_hybrid_geometry* g, double sigma_z,
                                   double sigma_dl, double length, const float* weights,
                                   int weight_planes, const float* sensitivity,
                                   int sens_depth, const oracle_hybrid_event* ev, int64_t n,
                                   int n_iter, const float* rho, float* out);
void oracle_hybrid_mlem_planes_f64(const oracle_hybrid_geometry* g, double sigma_z,
                                   double sigma_dl, double length, const double* weights,
                                   int weight_planes, const double* sensitivity,
                                   int sens_depth, const oracle_hybrid_event* ev, int64_t n,
                                   int n_iter, const double* rho, double* out);

/* Sensitivity volume and normalised weights of per-plane weights
 * (reconstruction.h:236-271 over `planes` planes): raw and weights are
 * planes x n_pixel_infos; sensitivity planes x rows x columns is computed
 * (sens_mode 0: sums of each plane's weights per pixel in pixel-info order;
 * 1: one) or given (2). */
void oracle_hybrid_plane_weights_f32(const oracle_hybrid_geometry* g, const float* raw,
                                     int planes, int sens_mode, float* weights,
                                     float* sensitivity);
void oracle_hybrid_plane_weights_f64(const oracle_hybrid_geometry* g, const double* raw,
                                     int planes, int sens_mode, double* weights,
                                     double* sensitivity);

#ifdef __cplusplus
}
#endif

#endif /* HYBRID_ORACLE_H */
