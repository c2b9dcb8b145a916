// 2d_strip_phantom: list-mode responses of the 2D strip PET scanner from a
// phantom, generated on an AMD Instinct GPU (libpetstrip_hip.so).
//
// Keeps the command line and file conventions of the reference tool
// (grey277/j-pet-mlem src/2d/strip/phantom_cmd.cpp:136-444,
// options.cpp:13-104):
//   2d_strip_phantom [options] phantom_description...
// Phantom descriptions hold one region per line:
//   ellipse x y a b angle intensity | rectangle x y a b intensity |
//   point x y intensity
// (src/2d/geometry/phantom.h:137-181; lengths of ellipses and rectangles
// scaled by --scale).  -e emissions are made (--detected: -e detected
// events); the detected ones are written as float32 (z_u, z_d, dl) triples
// (text with a .txt output), with <base>.cfg, the emitted / detected / naive
// images (raw int32 + NRRD + PNG, row 0 at the bottom) and, with
// --tan-bins, the tan-bins map and the analytic-kernel map.  The random
// stream is the library's counter-based one (--seed), so the events are the
// same for any thread count or device.
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/petstrip.h"
#include "cmdline.h"
#include "strip_io.h"

namespace {

using namespace strip_cli;

const std::vector<OptSpec> kOptions = {
    {"config", 'c', true, "", true, kDontSave, "load config file"},
    {"radius", 'r', true, "0.5", true, kSave, "R distance between scintillator"},
    {"length", 'l', true, "1", true, kSave, "scintillator length"},
    {"s-pixel", 'p', true, "0.01", true, kSave, "pixel size"},
    {"n-pixels", 'n', true, "", true, kDontSave, "number of pixels"},
    {"n-z-pixels", 0, true, "", true, kSave, "number of z pixels"},
    {"n-y-pixels", 0, true, "", true, kSave, "number of y pixels"},
    {"s-z", 0, true, "0.015", true, kAlwaysSave, "TOF sigma along z axis"},
    {"s-dl", 0, true, "0.06", true, kAlwaysSave, "TOF sigma delta-l"},
    {"n-emissions", 'e', true, "0", false, kSave, "number of emissions"},
    {"verbose", 'v', false, "", true, kDontSave, "print progress information (-v)"},
    {"n-threads", 'T', true, "", true, kDontSave, "ignored (generation runs on the GPU)"},
    {"scale", 0, true, "1", true, kAlwaysSave, "Scale phantom with given constant"},
    {"additive", 0, false, "", true, kDontSave, "phantom regions are additive, not disjunctive"},
    {"tan-bins", 0, true, "0", true, kDontSave,
     "Bin reconstructed events using given number of tangent bins"},
    {"kernel-point", 0, true, "", true, kDontSave, "Point to evaluate kernel in"},
    {"full", 0, false, "", true, kDontSave, "Emit additional full information"},
    {"no-responses", 0, false, "", true, kDontSave, "Do not emit responses"},
    {"point", 0, true, "", true, kSave, "Add point phantom x,y[,intensity]"},
    {"output", 'o', true, "phantom.bin", false, kSave, "responses file"},
    {"detected", 0, false, "", true, kDontSave, "collects detected emissions"},
    {"image-origin", 0, true, "", true, kDontSave, "Image origin x,y (default 0,0)"},
    {"seed", 0, true, "1", true, kSave, "seed of the counter-based random stream"},
    {"cuda-device", 'D', true, "0", true, kDontSave, "GPU device (negative: host threads)"},
    {"help", 'h', false, "", false, kDontSave, "print this message"},
};

void check(int rc) {
  if (rc != PETSTRIP_OK) throw std::runtime_error(petstrip_last_error());
}

// Region septuples (kind, x, y, a, b, angle_deg, intensity) from a phantom
// description; a line that does not parse ends the file, an unknown type is
// an error (phantom.h:137-181).
void read_description(const std::string& fn, double scale, std::vector<double>& regions) {
  std::ifstream in(fn);
  if (!in.is_open()) throw std::runtime_error("cannot open phantom description: " + fn);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    std::string type;
    iss >> type;
    if (type == "ellipse") {
      double x, y, a, b, angle, intensity;
      if (!(iss >> x >> y >> a >> b >> angle >> intensity)) break;
      regions.insert(regions.end(),
                     {0, x * scale, y * scale, a * scale, b * scale, angle, intensity});
    } else if (type == "rectangle") {
      double x, y, a, b, intensity;
      if (!(iss >> x >> y >> a >> b >> intensity)) break;
      regions.insert(regions.end(), {1, x * scale, y * scale, a * scale, b * scale, 0, intensity});
    } else if (type == "point") {
      double x, y, intensity;
      if (!(iss >> x >> y >> intensity)) break;
      regions.insert(regions.end(), {2, x, y, 0, 0, 0, intensity});
    } else {
      throw std::runtime_error("unknown phantom type: " + type);
    }
  }
}

// Scanner::sensitivity at a point (src/2d/strip/scanner.h:147-171).
double sensitivity(double R, double L, double z, double y) {
  double hl = L / 2, lp = hl + z, lm = hl - z, rp = R + y, rm = R - y;
  if (rp <= 0) rp = 1;
  if (rm <= 0) rm = 1;
  double mn = std::min(lm / rm, lp / rp), mx = std::max(-lp / rm, -lm / rp);
  return (std::atan(mn) - std::atan(mx)) / M_PI;
}

void write_responses(const std::string& fn, bool txt, const float* r, int64_t n, int width) {
  if (txt) {
    std::ofstream out(fn);
    if (!out.is_open()) throw std::runtime_error("cannot write " + fn);
    for (int64_t i = 0; i < n; ++i) {
      for (int k = 0; k < width; ++k) out << (k ? " " : "") << r[width * i + k];
      out << '\n';
    }
  } else {
    strip_io::write_raw_bytes(fn, r, (size_t)n * width * sizeof(float));
  }
}

void write_image(const std::string& base, const std::string& suffix,
                 const std::vector<int32_t>& img, int w, int h) {
  const std::string raw = base + suffix;
  strip_io::write_raw_bytes(raw, img.data(), img.size() * sizeof(int32_t));
  strip_io::write_nrrd_typed(raw + ".nrrd", strip_io::without_path(raw), "int",
                             {(size_t)w, (size_t)h}, false);
  std::vector<float> f(img.begin(), img.end());
  strip_io::write_png(raw + ".png", f.data(), w, h, 0);
}

int run(int argc, char** argv) {
  Args a;
  a.specs = kOptions;
  std::vector<std::string> args(argv + 1, argv + argc);
  parse(a, args, false);
  if (a.exist("help")) {
    usage(a, "2d_strip_phantom [options] phantom_description ...", "");
    return 0;
  }
  derive_pixel_counts(a);
  if (a.rest.empty() && !a.exist("point")) {
    if (argc == 1) {
      usage(a, "2d_strip_phantom [options] phantom_description ...", "");
      return 0;
    }
    throw std::runtime_error("at least one input phantom description expected, consult --help");
  }
  const int64_t n = to_long(a.get("n-emissions"));
  const int verbose = a.flag("verbose");

  petstrip_scanner sc{};
  sc.radius = to_doubles(a.get("radius")).at(0);
  sc.scintillator_length = to_double(a.get("length"));
  sc.n_y_pixels = (int32_t)to_long(a.get("n-y-pixels"));
  sc.n_z_pixels = (int32_t)to_long(a.get("n-z-pixels"));
  sc.pixel_height = sc.pixel_width = to_double(a.get("s-pixel"));
  sc.sigma_z = to_double(a.get("s-z"));
  sc.sigma_dl = to_double(a.get("s-dl"));

  std::vector<double> regions;
  const double scale = to_double(a.get("scale"));
  for (auto& fn : a.rest) read_description(fn, scale, regions);
  if (a.exist("point")) {
    auto pt = to_doubles(a.get("point"));
    if (pt.size() < 2 || pt.size() > 3)
      throw std::runtime_error("--point must specifcy 2 or 3 values x,y[,intensity]");
    regions.insert(regions.end(), {2, pt[0], pt[1], 0, 0, 0, pt.size() == 3 ? pt[2] : 1.0});
  }
  if (regions.empty()) throw std::runtime_error("must specify at least one region");

  const int W = sc.n_z_pixels, H = sc.n_y_pixels;
  const double s = sc.pixel_width;
  if (verbose) {
    std::cerr << "    image size: " << W << " x " << H << std::endl
              << "    pixel size: " << s << " x " << s << std::endl
              << "    space size: " << W * s << " x " << H * s << std::endl
              << "space top left: " << -W * s / 2 + s / 2 << " x " << H * s / 2 - s / 2
              << std::endl;
  }

  const std::string output = a.get("output");
  const std::string base = strip_io::without_ext(output), ext = strip_io::ext(output);
  const bool txt = ext == ".txt", full = a.exist("full");
  const bool responses = !base.empty() && !a.exist("no-responses");

  double ox = 0, oy = 0;
  if (a.exist("image-origin")) {
    auto o = to_doubles(a.get("image-origin"));
    if (o.size() != 2) throw std::runtime_error("--image-origin must specify two values for x,y");
    ox = o[0];
    oy = o[1];
  }
  const int tan_bins = (int)to_long(a.get("tan-bins"));
  // the reference's (F)1.1 * L / (2 R) in float
  const double max_tan = (float)(1.1f * (float)sc.scintillator_length / (2 * (float)sc.radius));
  double kpx = 0, kpy = 0;
  if (a.exist("kernel-point")) {
    auto k = to_doubles(a.get("kernel-point"));
    if (k.size() != 2) throw std::runtime_error("--kernel-point must supply exactly 2 values for x, y");
    kpx = k[0];
    kpy = k[1];
  }
  petstrip_phantom_grid grid{W, H, s, -s * W / 2 + ox, -s * H / 2 + oy, std::max(tan_bins, 0),
                             max_tan};

  const size_t np = (size_t)W * H;
  std::vector<float> resp((size_t)std::max<int64_t>(n, 0) * 3);
  std::vector<float> exact(full ? resp.size() : 0), events(full ? (size_t)n * 4 : 0);
  std::vector<int32_t> emitted(np), detected(np), naive(np);
  std::vector<int32_t> tanmap(tan_bins > 0 ? np * tan_bins : 0);
  petstrip_phantom_output out{};
  out.responses = resp.data();
  out.exact = full ? exact.data() : nullptr;
  out.events = full ? events.data() : nullptr;
  out.emitted = emitted.data();
  out.detected = detected.data();
  out.naive = naive.data();
  out.tan_bins_map = tan_bins > 0 ? tanmap.data() : nullptr;
  const int32_t flags = (a.exist("additive") ? PETSTRIP_PHANTOM_ADDITIVE : 0) |
                        (a.exist("detected") ? PETSTRIP_PHANTOM_DETECTED : 0);
  check(petstrip_phantom_generate(&sc, regions.data(), (int32_t)(regions.size() / 7), flags,
                                  (uint64_t)to_long(a.get("seed")), n, &grid, &out,
                                  (int)to_long(a.get("cuda-device"))));
  if (verbose)
    std::cerr << " emitted: " << out.n_emitted << " events" << std::endl
              << "detected: " << out.n_responses << " events" << std::endl;

  if (responses) {
    write_responses(output, txt, resp.data(), out.n_responses, 3);
    if (full) {
      write_responses(base + "_wo_error" + ext, txt, exact.data(), out.n_responses, 3);
      write_responses(base + "_full_response" + ext, txt, exact.data(), out.n_responses, 3);
      write_responses(base + "_events" + ext, txt, events.data(), out.n_responses, 4);
    }
  }
  if (!base.empty()) {
    write_config(a, base + ".cfg");
    write_image(base, "_detected", detected, W, H);
    write_image(base, "_emitted", emitted, W, H);
    write_image(base, "_naive", naive, W, H);
    if (tan_bins > 0) {
      const std::string tb = base + "_tan_bins";
      strip_io::write_raw_bytes(tb, tanmap.data(), tanmap.size() * sizeof(int32_t));
      strip_io::write_nrrd_typed(tb + ".nrrd", strip_io::without_path(tb), "int",
                                 {(size_t)W, (size_t)H, (size_t)tan_bins}, false);
      if (verbose) std::cerr << "generating kernel map..." << std::endl;
      const double inv_sens =
          a.exist("detected")
              ? 1 / sensitivity(sc.radius, sc.scintillator_length, kpx, kpy)
              : 1.0;
      std::vector<float> kmap(np * tan_bins);
      check(petstrip_phantom_tan_kernel_map(&sc, &grid, kpx, kpy, inv_sens, kmap.data(),
                                            (int)to_long(a.get("cuda-device"))));
      const std::string tk = base + "_tan_kernel";
      strip_io::write_raw_bytes(tk, kmap.data(), kmap.size() * sizeof(float));
      strip_io::write_nrrd_typed(tk + ".nrrd", strip_io::without_path(tk), "float",
                                 {(size_t)W, (size_t)H, (size_t)tan_bins}, false);
    }
  }
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
