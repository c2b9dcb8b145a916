// File formats of the strip reconstruction command line, written from the
// format descriptions of the reference (grey277/j-pet-mlem):
//   responses: binary float32 triples (z_u, z_d, dl) until EOF, or text with
//              3 (or, with --3d-response, 5) numbers per line
//              (src/2d/strip/reconstruction.h:156-196)
//   images:    raw float32 row-major, text matrix or "x y value" tuples,
//              NRRD header, 8-bit inverted grayscale PNG
//              (src/util/png_writer.h:31-54, nrrd_writer.h:18-27,
//               src/2d/geometry/pixel_map.h:134-157, reconstruction.h:202-213)
//   names:     <base>_<index> zero-padded to the digits of the maximum
//              (src/util/cmdline_types.h:35-55)
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace strip_io {

// Path helpers with the reference's rules: the extension starts at the last
// '.' after the last '/'.
std::string without_ext(const std::string& p);
std::string ext(const std::string& p);
std::string without_path(const std::string& p);
std::string add_index(const std::string& base, int index, int max_value);

// Responses (z_u, z_d, dl) appended to out.  Throws std::runtime_error.
void load_binary_responses(const std::string& fn, std::vector<float>& out);
// Text responses through the library's reader (responses_io.cpp, linked
// with libpetstrip_hip): 3 fields per line, or 5 with is_3d, whose first
// two (the LOR's detectors) go to lors when it is not null.
void load_text_responses(const std::string& fn, bool is_3d, std::vector<float>& out,
                         std::vector<int16_t>* lors = nullptr);

// Images: width columns x height rows, row-major.
void write_raw(const std::string& fn, const float* img, size_t n);
void write_text_matrix(const std::string& fn, const float* img, int width, int height);
void write_text_tuples(const std::string& fn, const float* img, int width, int height);
void write_nrrd(const std::string& fn, const std::string& raw_name, int width,
                int height, bool ascii);
// NRRD header of any type ("int", "float") and dimension (sizes fastest
// first), and raw bytes (src/util/nrrd_writer.h:18-27, bstream.h).
void write_nrrd_typed(const std::string& fn, const std::string& raw_name, const char* type,
                      const std::vector<size_t>& sizes, bool ascii);
void write_raw_bytes(const std::string& fn, const void* data, size_t bytes);
// scale_max = 0: the image maximum maps to black (inverted grayscale).
void write_png(const std::string& fn, const float* img, int width, int height,
               double scale_max);

// Reads a raw float32 image of n values (for --rho); with as_int the file
// holds int32 values and 1000 is added to each (reconstruction_cmd.cpp:155-160).
void load_raw_image(const std::string& fn, float* out, size_t n, bool as_int);

}  // namespace strip_io
