// 3d_hybrid_reconstruction: the 3D hybrid list-mode reconstruction of a barrel
// scanner on an AMD Instinct GPU (libpetstrip_hip.so, include/
// petstrip_hybrid.h), with the reference tool's command line and files
// (grey277/j-pet-mlem src/3d/hybrid/reconstruction_cmd.cpp:147-640):
//   3d_hybrid_reconstruction --geometry geom.bin [--system m.bin[,m2.bin...]]
//                            [options] response...
//   3d_hybrid_reconstruction --system m.bin --detector-file dets.txt
//                            [--inactive i,j,...] [options] response...
// With --geometry the LOR-pixel weights are the geometry's Gaussian weights
// (calculate_weight), or with --system those of one system matrix or of one
// per half plane (the multi-plane geometry, reconstruction_cmd.cpp:281-365;
// without --system the reference indexes its empty matrix list,
// reconstruction_cmd.cpp:350-357, where this build takes the geometry's
// weights).  With --system alone the geometry is the matrix's (full form,
// PETP or TOFP) with the detector centres of --detector-file, the matrix's
// .cfg beside it loaded as options, --n-pixels the matrix's pixels per row
// and --s-pixel required (reconstruction_cmd.cpp:203-225,367-455).  The
// geometry file is 2d_barrel_geometry's binary output (PETg); responses
// are text lines "d1 d2 z_up z_dn dl" (3d_tool_convert's output) or binary
// records (int16 d1, d2, float z_up, z_dn, dl).  -i blocks of -I iterations
// run on the GPU; with -o base the naive image (base + .nrrd, int voxel
// counts) and each output iteration (base_NNN + .nrrd and raw or text
// floats, planes x rows x columns, optionally cropped) are written as the
// reference's --gpu path writes them.  -G is accepted (the GPU always
// runs).  --sensitivity reads an external volume of n_planes / 2 planes of
// raw floats (the reference's binary VoxelMap, voxel_map.h:175-178); the
// reference's text reader assigns to copies of the voxels and leaves the
// volume unset (voxel_map.h:153-158), so a .txt volume is refused.  With
// inactive detectors the reference sizes its pixel-info arrays by the
// matrix's element count and reads the slots of skipped elements, which it
// never fills, in calculate_sensitivity (geometry_soa.h:81-115,
// reconstruction.h:240-251); here the geometry holds the filled slots only.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/petstrip_hybrid.h"
#include "cmdline.h"
#include "strip_io.h"

namespace {

using namespace strip_cli;

const std::vector<OptSpec> kOptions = {
    {"config", 'c', true, "", true, kDontSave, "load config file"},
    {"system", 's', true, "", true, kSave, "system matrix file"},
    {"geometry", 0, true, "", true, kSave, "geometry information file"},
    {"n-planes", 0, true, "0", true, kSave, "number of voxels in z direction"},
    {"z-left", 0, true, "0", true, kSave, "left extent in z direction"},
    {"length", 'l', true, "2", true, kSave, "length of the detector"},
    {"s-z", 0, true, "0.015", true, kAlwaysSave, "TOF sigma along z axis"},
    {"s-dl", 0, true, "0.060", true, kAlwaysSave, "TOF sigma axis"},
    {"sens-to-one", 0, false, "", true, kSave, "sets sensitivity to one"},
    {"sensitivity", 0, true, "", true, kSave, "external 3D sensitivity"},
    {"inactive", 0, true, "", true, kSave, "list of inactive detectors"},
    {"crop", 0, true, "0", true, kSave, "numer of pixel to crop the output image"},
    {"crop-x", 0, true, "0", true, kSave, "crop origin pixel x"},
    {"crop-y", 0, true, "0", true, kSave, "crop origin pixel y"},
    {"crop-z", 0, true, "0", true, kSave, "crop origin pixel z"},
    {"n-pixels", 'n', true, "256", true, kSave, "number of pixels in one dimension"},
    {"s-pixel", 'p', true, "", true, kSave, "pixel size"},
    {"shape", 0, true, "square", true, kSave, "detector shape"},
    {"fov-radius", 0, true, "", true, kSave, "field of view radius"},
    {"m-pixel", 0, true, "0", true, kSave, "starting pixel for partial matrix"},
    {"n-emissions", 'e', true, "0", false, kSave, "emissions per pixel"},
    {"output", 'o', true, "", false, kDontSave, "output base name"},
    {"full", 'f', false, "", true, kSave, "full sparse system matrix"},
    {"png", 0, true, "", false, kDontSave, "output lor to png"},
    {"from", 0, true, "-1", false, kDontSave, "lor start detector to output"},
    {"to", 0, true, "-1", false, kDontSave, "lor end detector to output"},
    {"pos", 0, true, "-1", false, kDontSave, "position to output"},
    {"print", 0, false, "", true, kSave, "print triangular sparse system matrix"},
    {"stats", 0, false, "", true, kSave, "show stats"},
    {"wait", 0, false, "", true, kSave, "wait before exit"},
    {"verbose", 'v', false, "", true, kSave, "prints the iterations information on std::out"},
    {"seed", 'S', true, "0", true, kDontSave, "random number generator seed"},
    {"gpu", 'G', false, "", true, kSave, "run on the GPU (always)"},
    {"cuda-device", 'D', true, "0", true, kSave, "GPU device"},
    {"cuda-blocks", 'B', true, "", true, kSave, "launch hint (ignored)"},
    {"cuda-threads", 'W', true, "", true, kSave, "launch hint (ignored)"},
    {"n-threads", 0, true, "", true, kSave, "host threads (ignored)"},
    {"blocks", 'i', true, "0", true, kSave, "number of iteration blocks"},
    {"iterations", 'I', true, "1", true, kSave, "number of iterations (per block)"},
    {"z-position", 'z', true, "0", true, kSave, "position of the z plane"},
    {"rho", 0, true, "", true, kSave, "start rho (eg. existing iteration)"},
    {"detector-file", 0, true, "", true, kSave, "detector description file"},
    {"detector-file-sym", 0, true, "", true, kSave,
     "detector symmetries description file (unused)"},
    {"help", '?', false, "", false, kDontSave, "print this message"},
};

// A geometry file (src/2d/barrel/geometry.h:55-104, lor_geometry.h:46-88):
// magic "PETg", int16 n_detectors, the grid, then LOR records.
struct Geometry {
  int32_t n_detectors = 0, n_columns = 0, n_rows = 0;
  float pixel_size = 0, llx = 0, lly = 0;
  std::vector<float> segments, widths, positions, weights;
  std::vector<int64_t> begin, end;
  std::vector<int16_t> pixels;
};

Geometry read_geometry(const std::string& fn) {
  std::ifstream in(fn, std::ios::binary);
  if (!in.is_open()) throw std::runtime_error("cannot open geometry file: " + fn);
  auto rd = [&](void* p, size_t n) {
    in.read(static_cast<char*>(p), (std::streamsize)n);
    return (bool)in;
  };
  char magic[4];
  if (!rd(magic, 4) || std::memcmp(magic, "PETg", 4) != 0)
    throw std::runtime_error("unknown binary geometry file");
  int16_t nd, nc, nr;
  float grid[3];
  if (!rd(&nd, 2) || !rd(&nc, 2) || !rd(&nr, 2) || !rd(grid, 12) || nd < 1)
    throw std::runtime_error("truncated geometry file");
  Geometry g;
  g.n_detectors = nd;
  g.n_columns = nc;
  g.n_rows = nr;
  g.pixel_size = grid[0];
  g.llx = grid[1];
  g.lly = grid[2];
  const int64_t n_lors = (int64_t)nd * (nd + 1) / 2;
  g.segments.assign(4 * n_lors, 0.0f);
  g.widths.assign(n_lors, 0.0f);
  std::vector<std::vector<uint8_t>> infos(n_lors);
  for (;;) {
    int16_t lor[2];
    float seg[4], width;
    uint64_t n;
    if (!rd(lor, 4) || !rd(seg, 16) || !rd(&width, 4) || !rd(&n, 8)) break;
    if (lor[1] < 0 || lor[0] < lor[1] || lor[0] >= nd)
      throw std::runtime_error("geometry file: LOR out of range");
    const int64_t li = (int64_t)lor[0] * (lor[0] + 1) / 2 + lor[1];
    std::memcpy(&g.segments[4 * li], seg, 16);
    g.widths[li] = width;
    infos[li].resize(n * 20);
    if (n && !rd(infos[li].data(), n * 20)) throw std::runtime_error("truncated geometry file");
  }
  // GeometrySOA layout: pixel infos in LOR index order (geometry_soa.h:178-199)
  g.begin.resize(n_lors);
  g.end.resize(n_lors);
  int64_t total = 0;
  for (int64_t li = 0; li < n_lors; ++li) {
    g.begin[li] = total;
    const size_t n = infos[li].size() / 20;
    for (size_t k = 0; k < n; ++k) {
      const uint8_t* r = infos[li].data() + 20 * k;
      int16_t xy[2];
      float f[4];  // t, distance, fill, weight
      std::memcpy(xy, r, 4);
      std::memcpy(f, r + 4, 16);
      g.pixels.push_back(xy[0]);
      g.pixels.push_back(xy[1]);
      g.positions.push_back(f[0]);
      g.weights.push_back(f[3]);
    }
    total += (int64_t)n;
    g.end[li] = total;
  }
  return g;
}

void check(int rc);

struct Matrix {
  petstrip_system_matrix* m = nullptr;
  int32_t n_pixels_in_row = 0, n_detectors = 0;
  int64_t n_emissions = 0, n_elements = 0;
  explicit Matrix(const std::string& fn) {
    if (petstrip_system_matrix_read(fn.c_str(), &m) != PETSTRIP_OK)
      throw std::runtime_error(petstrip_last_error());
    petstrip_system_matrix_info(m, &n_pixels_in_row, &n_detectors, &n_emissions, &n_elements);
  }
  ~Matrix() { petstrip_system_matrix_free(m); }
  Matrix(const Matrix&) = delete;
  Matrix& operator=(const Matrix&) = delete;
};

petstrip_hybrid_geometry c_geometry(const Geometry& g, int n_planes, double z_left) {
  return petstrip_hybrid_geometry{g.n_detectors, g.n_columns, g.n_rows, n_planes,
                                  g.pixel_size, g.llx, g.lly, z_left,
                                  (int64_t)g.positions.size(), g.segments.data(),
                                  g.widths.data(), g.begin.data(), g.end.data(),
                                  g.pixels.data(), g.positions.data(), g.weights.data()};
}

std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  std::istringstream in(s);
  std::string item;
  while (std::getline(in, item, sep)) out.push_back(item);
  return out;
}

// The geometry of a matrix and the detector centres on the matrix's centred
// grid of pixel_size (GeometrySOA(matrix, centres, grid, inactive),
// geometry_soa.h:71-182; PixelGrid(n, n, s) has its lower left at
// (-s n / 2, -s n / 2) in float, pixel_grid.h:33-37).
Geometry geometry_from_matrix(const Matrix& mx, const std::vector<float>& centres,
                              double s_pixel, const std::vector<int16_t>& inactive) {
  Geometry g;
  g.n_detectors = mx.n_detectors;
  g.n_columns = g.n_rows = mx.n_pixels_in_row;
  g.pixel_size = (float)s_pixel;
  g.llx = g.lly = -g.pixel_size * (float)(int16_t)mx.n_pixels_in_row / 2;
  const int16_t* ina = inactive.empty() ? nullptr : inactive.data();
  int64_t ni = 0;
  check(petstrip_hybrid_geometry_from_matrix(mx.m, centres.data(), g.n_columns, g.n_rows,
                                             g.pixel_size, g.llx, g.lly, ina,
                                             (int32_t)inactive.size(), nullptr, nullptr,
                                             nullptr, nullptr, nullptr, nullptr, nullptr, &ni));
  const int64_t n_lors = (int64_t)g.n_detectors * (g.n_detectors + 1) / 2;
  g.segments.resize(4 * n_lors);
  g.widths.resize(n_lors);
  g.begin.resize(n_lors);
  g.end.resize(n_lors);
  g.pixels.resize(2 * std::max<int64_t>(ni, 1));
  g.positions.resize(std::max<int64_t>(ni, 1));
  g.weights.resize(std::max<int64_t>(ni, 1));
  check(petstrip_hybrid_geometry_from_matrix(
      mx.m, centres.data(), g.n_columns, g.n_rows, g.pixel_size, g.llx, g.lly, ina,
      (int32_t)inactive.size(), g.segments.data(), g.widths.data(), g.begin.data(),
      g.end.data(), g.pixels.data(), g.positions.data(), g.weights.data(), &ni));
  g.pixels.resize(2 * ni);
  g.positions.resize(ni);
  g.weights.resize(ni);
  return g;
}

// Responses: text "d1 d2 z_up z_dn dl" per line, read as the reference's
// fast text parser reads them (reconstruction.h:165-188, through the
// library's reader), or binary int16 d1, d2, float z_up, z_dn, dl records
// (3d/hybrid/scanner.h:80-88).
void load_responses(const std::string& fn, std::vector<int16_t>& lors, std::vector<float>& zzd) {
  if (strip_io::ext(fn) == ".txt") {
    strip_io::load_text_responses(fn, true, zzd, &lors);
    return;
  }
  std::ifstream in(fn, std::ios::binary);
  if (!in.is_open()) throw std::runtime_error("cannot open input response file: " + fn);
  for (;;) {
    int16_t l[2];
    float v[3];
    in.read(reinterpret_cast<char*>(l), 4);
    in.read(reinterpret_cast<char*>(v), 12);
    if (!in) break;
    lors.push_back(l[0]);
    lors.push_back(l[1]);
    zzd.insert(zzd.end(), v, v + 3);
  }
}

struct Ctx {
  std::string base, oext;
  bool txt = false;
  int n_iterations = 0;
  int crop = 0, crop_x = 0, crop_y = 0, crop_z = 0;
  int verbose = 0;
};

void write_volume(const Ctx& c, const std::string& fn, const float* img, int W, int H, int D) {
  std::vector<float> cropped;
  const float* out = img;
  int w = W, h = H, d = D;
  if (c.crop > 0) {
    // VoxelMap::copy: a crop^3 block centred, shifted by the crop origin
    // (3d/geometry/voxel_map.h:111-124)
    const int tx = (W - c.crop) / 2 + c.crop_x, ty = (H - c.crop) / 2 + c.crop_y,
              tz = (D - c.crop) / 2 + c.crop_z;
    if (tx < 0 || ty < 0 || tz < 0 || tx + c.crop > W || ty + c.crop > H || tz + c.crop > D)
      throw std::runtime_error("crop outside the image");
    w = h = d = c.crop;
    cropped.resize((size_t)w * h * d);
    for (int z = 0; z < d; ++z)
      for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
          cropped[((size_t)z * h + y) * w + x] =
              img[((size_t)(z + tz) * H + (y + ty)) * W + (x + tx)];
    out = cropped.data();
  }
  strip_io::write_nrrd_typed(fn + ".nrrd", strip_io::without_path(fn) + c.oext, "float",
                             {(size_t)w, (size_t)h, (size_t)d}, c.txt);
  if (c.txt)
    strip_io::write_text_matrix(fn + ".txt", out, w, h * d);
  else
    strip_io::write_raw_bytes(fn + c.oext, out, sizeof(float) * (size_t)w * h * d);
}

int on_output(void* user, int iteration, const float* img, int32_t W, int32_t H, int32_t D) {
  auto* c = static_cast<Ctx*>(user);
  if (c->base.empty()) return 0;
  const std::string fn = iteration >= 0 ? strip_io::add_index(c->base, iteration, c->n_iterations)
                                        : c->base + "_sensitivity";
  write_volume(*c, fn, img, W, H, D);
  return 0;
}

void on_progress(void* user, int completed, int finished) {
  auto* c = static_cast<Ctx*>(user);
  if (c->verbose)
    std::cerr << (finished ? "  finished " : "  iteration ") << completed << std::endl;
}

void check(int rc) {
  if (rc != PETSTRIP_OK) throw std::runtime_error(petstrip_last_error());
}

int run(int argc, char** argv) {
  Args a;
  a.specs = kOptions;
  std::vector<std::string> args(argv + 1, argv + argc);
  parse(a, args, false);
  if (a.exist("help") || argc == 1) {
    usage(a, "3d_hybrid_reconstruction --geometry=file [options] response ...",
          "note: All length options below should be expressed in meters.");
    return 0;
  }
  const bool with_geometry = a.exist("geometry");
  if (!with_geometry && !a.exist("system"))
    throw std::runtime_error("either --geometry or --system matrix must be specified");
  const int verbose = a.flag("verbose");
  Geometry g;
  std::vector<std::unique_ptr<Matrix>> matrices;
  int n_pixels;
  double s_pixel;
  if (with_geometry) {
    g = read_geometry(a.get("geometry"));
    // --n-pixels and --s-pixel from the geometry unless given, then the
    // planes (reconstruction_cmd.cpp:287-315)
    n_pixels = a.exist("n-pixels") ? (int)to_long(a.get("n-pixels"))
                                   : std::max(g.n_columns, g.n_rows);
    // the reference passes the geometry's pixel size through the option
    // parser as text with six significant digits
    if (a.exist("s-pixel")) {
      s_pixel = to_double(a.get("s-pixel"));
    } else {
      std::ostringstream ps;
      ps << g.pixel_size;
      s_pixel = to_double(ps.str());
    }
    if (a.exist("system"))
      for (const auto& fn : split(a.get("system"), ','))
        matrices.emplace_back(new Matrix(fn));
  } else {
    // the matrix, then its .cfg as options (reconstruction_cmd.cpp:203-215)
    const std::string fn = a.get("system");
    matrices.emplace_back(new Matrix(fn));
    const std::string cfg = strip_io::without_ext(fn) + ".cfg";
    if (std::ifstream(cfg).is_open()) load_config(a, cfg, a.flag("verbose"));
    const Matrix& mx = *matrices[0];
    // the triangular forms are refused by the reader
    n_pixels = a.exist("n-pixels") ? (int)to_long(a.get("n-pixels")) : mx.n_pixels_in_row;
    if (!a.exist("s-pixel"))
      throw std::runtime_error("--s-pixel is required with --system alone (or in the "
                               "matrix's .cfg)");
    s_pixel = to_double(a.get("s-pixel"));
    // the detector centres of the scanner description file
    // (reconstruction_cmd.cpp:127-146,409-411)
    const std::string df = a.exist("detector-file") ? a.get("detector-file") : std::string();
    std::vector<float> centres;
    int32_t nd = 0;
    if (petstrip_read_detector_centres(df.c_str(), nullptr, 0, &nd) != PETSTRIP_OK) {
      std::cerr << "cannot open detector description file `" << df << "'\n";
      nd = 0;
    } else {
      centres.resize(2 * (size_t)std::max(nd, 1));
      check(petstrip_read_detector_centres(df.c_str(), centres.data(), nd, &nd));
    }
    if (nd != mx.n_detectors) throw std::runtime_error("n_detectors mismatch");
    std::vector<int16_t> inactive;
    if (a.exist("inactive"))
      for (const auto& d : split(a.get("inactive"), ',')) {
        const long v = to_long(d);
        if (v < 0 || v >= nd) throw std::runtime_error("inactive detector out of range: " + d);
        inactive.push_back((int16_t)v);
      }
    g = geometry_from_matrix(mx, centres, s_pixel, inactive);
  }
  const int n_planes = a.exist("n-planes") ? (int)to_long(a.get("n-planes")) : n_pixels;
  const double z_left = a.exist("z-left") ? to_double(a.get("z-left")) : -n_planes * s_pixel / 2;
  if (with_geometry && !matrices.empty() && matrices.size() != 1 &&
      matrices.size() != (size_t)(n_planes / 2))
    throw std::runtime_error("you must specify either one or half no. planes matrices");
  if (verbose)
    std::cerr << "3D hybrid reconstruction w/" << (with_geometry ? "geometry" : "system matrix")
              << ":\n    detectors = " << g.n_detectors << "\n   pixel grid = " << g.n_columns
              << " x " << g.n_rows << " / " << g.pixel_size << "\n  planes = " << n_planes
              << ", z-left = " << z_left << std::endl;

  // reconstruction_cmd.cpp:505-519: an external volume takes precedence over
  // --sens-to-one and normalises the geometry weights; read before any
  // device work
  std::vector<float> vol;
  if (a.exist("sensitivity")) {
    const std::string sn = a.get("sensitivity");
    if (strip_io::ext(sn) == ".txt")
      throw std::runtime_error("a text sensitivity volume is not read (give raw floats): " + sn);
    const int depth = n_planes / 2;
    if (depth < 1) throw std::runtime_error("--sensitivity needs at least two planes");
    vol.resize((size_t)g.n_columns * g.n_rows * depth);
    std::ifstream in(sn, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error("cannot open sensitivity file: " + sn);
    in.read(reinterpret_cast<char*>(vol.data()), (std::streamsize)(vol.size() * sizeof(float)));
    if ((size_t)in.gcount() != vol.size() * sizeof(float))
      throw std::runtime_error("sensitivity file shorter than " + std::to_string(depth) +
                               " planes of " + std::to_string(g.n_columns) + " x " +
                               std::to_string(g.n_rows) + " floats: " + sn);
  }

  petstrip_hybrid_geometry hg = c_geometry(g, n_planes, z_left);
  // with --geometry and matrices: each matrix's weights on the geometry's
  // pixel infos (load_weights_from_matrix_file), one plane per matrix
  std::vector<float> raw;
  if (with_geometry && !matrices.empty()) {
    const size_t ni = g.positions.size();
    raw.resize(matrices.size() * ni);
    for (size_t p = 0; p < matrices.size(); ++p) {
      if (verbose) std::cerr << "system matrix = " << split(a.get("system"), ',')[p] << std::endl;
      check(petstrip_system_matrix_weights(matrices[p]->m, &hg, raw.data() + p * ni));
    }
  }
  // without --geometry the weights are the matrix's; with --geometry and no
  // matrix, the Gaussian weights computed from the geometry
  const int32_t flags = (a.exist("sens-to-one") ? PETSTRIP_HYBRID_SENSITIVITY_ONE : 0) |
                        (with_geometry ? 0 : PETSTRIP_HYBRID_GEOMETRY_WEIGHTS);
  petstrip_hybrid* h = nullptr;
  check(petstrip_hybrid_create(&hg, to_double(a.get("s-z")), to_double(a.get("s-dl")),
                               to_double(a.get("length")), flags,
                               (int32_t)to_long(a.get("cuda-device")), &h));
  struct Guard {
    petstrip_hybrid* h;
    ~Guard() { petstrip_hybrid_destroy(h); }
  } guard{h};

  // the sensitivity: the external volume, one, or calculated
  // (reconstruction_cmd.cpp:505-521)
  if (!raw.empty()) {
    const int mode = !vol.empty() ? 2 : a.exist("sens-to-one") ? 1 : 0;
    check(petstrip_hybrid_set_plane_weights(h, raw.data(), (int32_t)matrices.size(), mode,
                                            vol.empty() ? nullptr : vol.data()));
  } else if (!vol.empty()) {
    check(petstrip_hybrid_set_sensitivity_3d(h, vol.data(), n_planes / 2));
  }

  const size_t nv = (size_t)g.n_columns * g.n_rows * n_planes;
  std::vector<float> rho(nv, 1.0f);
  int start_iteration = 0;
  if (a.exist("rho")) {
    const std::string rn = a.get("rho");
    // the index suffix of the name is the iteration to restart at
    const std::string stem = strip_io::without_ext(rn);
    size_t d = stem.size();
    while (d > 0 && std::isdigit((unsigned char)stem[d - 1])) --d;
    start_iteration = d < stem.size() ? std::stoi(stem.substr(d)) : 0;
    // a text rho is read into copies of the voxels and so leaves rho at one
    // (voxel_map.h:153-158); a binary one is read
    if (strip_io::ext(rn) != ".txt") {
      std::ifstream in(rn, std::ios::binary);
      if (!in.is_open()) throw std::runtime_error("cannot open rho file: " + rn);
      in.read(reinterpret_cast<char*>(rho.data()), (std::streamsize)(nv * sizeof(float)));
    }
    if (start_iteration && verbose) std::cerr << "restarting at = " << start_iteration << std::endl;
  }

  std::vector<int16_t> lors;
  std::vector<float> zzd;
  for (const auto& fn : a.rest) {
    if (verbose) std::cerr << "     response = " << fn << std::endl;
    load_responses(fn, lors, zzd);
  }
  const int64_t n = (int64_t)lors.size() / 2;
  for (int64_t e = 0; e < n; ++e)
    if (lors[2 * e] < lors[2 * e + 1])
      throw std::runtime_error("response with its first detector below its second");
  check(petstrip_hybrid_set_responses(h, lors.data(), zzd.data(), n));
  petstrip_hybrid_stats st{};
  check(petstrip_hybrid_stats_get(h, &st));
  if (verbose)
    std::cerr << "       events = " << st.n_events << "\n  voxels/event = "
              << (st.n_events ? (double)st.n_voxels / st.n_events : 0.0) << std::endl;

  Ctx c;
  if (a.exist("output")) {
    const std::string o = a.get("output");
    c.base = strip_io::without_ext(o);
    c.oext = strip_io::ext(o);
    c.txt = c.oext == ".txt";
  }
  const int n_blocks = (int)to_long(a.get("blocks"));
  const int n_iter = (int)to_long(a.get("iterations"));
  c.n_iterations = n_blocks * n_iter;
  c.crop = (int)to_long(a.get("crop"));
  c.crop_x = (int)to_long(a.get("crop-x"));
  c.crop_y = (int)to_long(a.get("crop-y"));
  c.crop_z = (int)to_long(a.get("crop-z"));
  c.verbose = verbose;

  if (!c.base.empty()) {
    // the naive image: each kept event's centre point counted in its voxel
    // (reconstruction.h:449-466), written where --output points
    std::vector<petstrip_hybrid_event> ev(st.n_events);
    check(petstrip_hybrid_get_events(h, ev.data()));
    std::vector<int32_t> naive(nv, 0);
    for (const auto& e : ev) {
      const int64_t li = (int64_t)e.first * (e.first + 1) / 2 + e.second;
      const float* s = &g.segments[4 * li];
      const float mx = (s[0] + s[2]) / 2.0f, my = (s[1] + s[3]) / 2.0f;
      const float vx = s[2] - s[0], vy = s[3] - s[1];
      const float l = std::sqrt(vx * vx + vy * vy);
      const float px = mx + (vx / l) * e.up, py = my + (vy / l) * e.up;
      const int x = (int)std::floor((px - g.llx) / g.pixel_size);
      const int y = (int)std::floor((py - g.lly) / g.pixel_size);
      const int z = (int)std::floor((e.right - (float)z_left) / g.pixel_size);
      if (x >= 0 && x < g.n_columns && y >= 0 && y < g.n_rows && z >= 0 && z < n_planes)
        ++naive[((size_t)z * g.n_rows + y) * g.n_columns + x];
    }
    strip_io::write_nrrd_typed(c.base + ".nrrd", strip_io::without_path(a.get("output")), "int",
                               {(size_t)g.n_columns, (size_t)g.n_rows, (size_t)n_planes}, c.txt);
    if (c.txt) {
      std::ofstream t(a.get("output"));
      for (size_t i = 0; i < nv; ++i)
        t << naive[i] << (((i + 1) % g.n_columns) ? " " : "\n");
    } else {
      strip_io::write_raw_bytes(a.get("output"), naive.data(), nv * sizeof(int32_t));
    }
  }
  check(petstrip_hybrid_run(h, rho.data(), n_blocks, n_iter, start_iteration, on_output,
                            on_progress, &c));
  check(petstrip_hybrid_stats_get(h, &st));
  std::cerr << "  event count = " << st.n_events << std::endl;
  std::cerr << "  voxel count = " << st.n_voxels << "("
            << (st.n_events ? (double)st.n_voxels / st.n_events : 0.0) << " / event)"
            << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
}
