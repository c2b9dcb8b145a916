// 3d_tool_convert: list-mode data of the J-PET prototypes into the responses
// the strip and hybrid reconstructions read ("d1 d2 z_u z_d dl" text lines,
// the --3d-response format), with the reference tool's options and output
// (grey277/j-pet-mlem src/3d/tool/convert_cmd.cpp:55-210,
// src/3d/tool/options.cpp:32-66, src/2d/barrel/options.cpp:16-41,225-360):
//   3d_tool_convert [options] input... > responses.txt
//
// --warsaw (the default): each input line holds the hits of both gammas,
//   x y z t (cm, ps) twice, then extra columns and the scattering type last.
//   A hit belongs to the barrel detector whose rectangle contains it (within
//   0.1 mm); pairs with both hits on detectors and |z| inside the
//   scintillator are written with the higher detector first, z in metres
//   plus N(0, s-z) and dl = (t1 - t2) c plus N(0, s-dl).
// --tbednarski: strip numbers (1-based) and the times at both ends of both
//   strips; z from the time differences at the scintillator's light speed,
//   dl from the mean times (--relative-time: the times less their minimum).
//
// The barrel is built as the reference's ScannerBuilder builds rings of
// square detectors (src/2d/barrel/scanner_builder.h:36-93,94-228,
// square_detector.h:18-24), in single precision as there: a w x h rectangle
// moved out to the ring radius and turned to each detector's angle, the
// rings after the first turned further by rotation / n.  The random errors
// come from std::mt19937_64 seeded with --seed through two
// std::normal_distribution<double> (z and dl), drawn per written pair in the
// order of the output line: z_u, z_d, dl.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cmdline.h"
#include "strip_io.h"

namespace {

using namespace strip_cli;

const std::vector<OptSpec> kOptions = {
    {"config", 'c', true, "", true, kDontSave, "load config file"},
    {"shape", 0, true, "square", true, kSave,
     "detector shape (square, circle, triangle, hexagon)"},
    {"w-detector", 'w', true, "", true, kSave, "detector width"},
    {"h-detector", 'h', true, "", true, kSave, "detector height"},
    {"d-detector", 0, true, "", true, kSave, "inscribe detector shape into circle of given diameter"},
    {"radius", 'r', true, "", true, kSave, "inner detector ring radius(es)"},
    {"rotation", 0, true, "", true, kSave, "ring rotation(s) (0-1)"},
    {"n-detectors", 'd', true, "", true, kSave, "number of detectors in ring(s)"},
    {"fov-radius", 0, true, "", true, kSave, "field of view radius"},
    {"s-dl", 0, true, "0.06", true, kAlwaysSave, "TOF sigma delta-l"},
    {"s-z", 0, true, "0.015", true, kAlwaysSave, "TOF sigma along z axis"},
    {"length", 'l', true, "1", true, kSave, "scintillator length"},
    {"verbose", 'v', false, "", true, kSave, "prints the iterations information on std::out"},
    {"seed", 'S', true, "0", true, kDontSave, "random number generator seed"},
    {"warsaw", 0, false, "", true, kDontSave, "Warsaw data format"},
    {"tbednarski", 0, false, "", true, kDontSave, "T.Bednarski data format"},
    {"only", 0, true, "0", true, kSave, "limit event (scattering) type, eg. 1 for non-scattered"},
    {"relative-time", 0, false, "", true, kDontSave,
     "output relative time instead z_u, z_d, dl"},
    {"help", '?', false, "", false, kDontSave, "print this message"},
};

constexpr double kCm = 0.01;
constexpr double kLightMPerPs = 299792458.0e-12;      // in vacuum
constexpr double kScintLightMPerPs = 126000000.0e-12;  // in the scintillator
constexpr size_t kMaxDetectors = 400;  // the reference's default build limit

std::vector<int> to_ints(const std::string& s) {
  std::vector<int> out;
  for (double v : to_doubles(s)) out.push_back((int)v);
  return out;
}

// One detector: a rectangle's corners in the reference's order.
struct Detector {
  float x[4], y[4];

  void rotate(float phi) {
    const float s = std::sin(phi), c = std::cos(phi);
    for (int k = 0; k < 4; ++k) {
      const float tx = x[k] * c - y[k] * s;
      const float ty = x[k] * s + y[k] * c;
      x[k] = tx;
      y[k] = ty;
    }
  }

  // inside or within tol of every edge (edges taken from the last corner)
  bool contains(float px, float py, float tol) const {
    int prev = 3;
    for (int k = 0; k < 4; ++k) {
      const float ex = x[k] - x[prev], ey = y[k] - y[prev];
      // clockwise normal of the edge, dotted with the point's offset
      const float side = ey * (px - x[prev]) + (-ex) * (py - y[prev]);
      if (side < -tol) return false;
      prev = k;
    }
    return true;
  }
};

// One ring of n detectors of w x h (h = w when not given) at radius r, the
// first centred on -y; scanner_builder.h:36-93.
std::vector<Detector> ring_of(float r, int n, float w, float h, float d) {
  if (n > (int)kMaxDetectors)
    throw std::runtime_error("buld single ring: too many detectors: " + std::to_string(n));
  if (r <= 0) throw std::runtime_error("invalid radius");
  if (w > 0 && h == 0) h = w;
  if (w <= 0 || h <= 0) throw std::runtime_error("invalid detector size");
  if (n > 4 && n % 4 != 0) throw std::runtime_error("number of detectors must be multiple of 4");
  Detector base;
  const float hw = w / 2, hh = h / 2;
  const float shift = r + (d > 0 ? d : h) / 2;
  const float cx[4] = {hw, hw, -hw, -hw}, cy[4] = {hh, -hh, -hh, hh};
  for (int k = 0; k < 4; ++k) {
    base.x[k] = cx[k] + 0.0f;
    base.y[k] = cy[k] + shift;
  }
  std::vector<Detector> ring;
  for (int i = 0; i < n; ++i) {
    Detector det = base;
    det.rotate(2 * float(M_PI) * i / n - float(M_PI_2));
    ring.push_back(det);
  }
  return ring;
}

// All rings, inner first; scanner_builder.h:94-228.
std::vector<Detector> build_barrel(std::vector<float> radius, std::vector<float> rotation,
                                   std::vector<int> n, float w, float h, float d) {
  if (radius.empty()) throw std::runtime_error("must specify at least one radius");
  if (n.empty()) throw std::runtime_error("must specify at least one number of detectors");
  if (n.size() < rotation.size()) throw std::runtime_error("too many rotation values");
  if (radius.size() < n.size() || radius.size() < rotation.size())
    throw std::runtime_error("must specify all radiuses");
  for (size_t i = 0; i < radius.size(); ++i) {
    if (i && radius[i] < radius[i - 1])
      throw std::runtime_error("radiuses must be specified in incrementing order");
    if (i && n.size() == i) n.push_back(n[i - 1]);
    if (rotation.size() == i) rotation.push_back(0);
  }
  for (float rot : rotation)
    if (std::abs(rot) >= FLT_EPSILON && std::abs(rot - 0.5f) > FLT_EPSILON)
      throw std::runtime_error("broken symmetry");
  std::vector<Detector> all = ring_of(radius[0], n[0], w, h, d);
  for (size_t i = 1; i < radius.size(); ++i) {
    if (!radius[i]) break;
    if (!n[i]) n[i] = n[i - 1];
    if (n[i] + all.size() > kMaxDetectors)
      throw std::runtime_error("build multiple rings: too many detectors");
    std::vector<Detector> ring = ring_of(radius[i], n[i], w, h, d);
    const float turn = 2 * float(M_PI) * rotation[i] / ring.size();
    for (auto& det : ring) {
      det.rotate(turn);
      all.push_back(det);
    }
  }
  return all;
}

// Defaults of the barrel options (src/2d/barrel/options.cpp:225-360).
void barrel_defaults(Args& a, int argc) {
  if (!a.exist("w-detector") && !a.exist("d-detector") && !a.exist("n-detectors")) {
    if (argc == 1) {
      usage(a, "3d_tool_convert [options] means ...", nullptr);
      std::cerr << "requires one of following options specified:\n"
                << "  --w-detector, --d-detector, --n-detectors\n";
      std::exit(0);
    }
    throw std::runtime_error("need to specify either --w-detector, --d-detector, --n-detectors");
  }
  std::ostringstream assumed;
  std::vector<double> radius;
  if (a.exist("radius")) radius = to_doubles(a.get("radius"));
  if (radius.empty()) {
    radius.push_back(M_SQRT2);
    std::ostringstream v;
    v.precision(17);
    v << M_SQRT2;
    a.value["radius"] = v.str();
    assumed << "--radius=" << radius[0] << "\n";
  }
  if (!a.exist("fov-radius")) {
    const double fov = radius[0] / M_SQRT2;
    assumed << "--fov-radius=" << fov << "\n";
  }
  std::vector<int> n;
  if (a.exist("n-detectors")) n = to_ints(a.get("n-detectors"));
  const double d = a.exist("d-detector") ? to_double(a.get("d-detector")) : 0.0;
  double w = a.exist("w-detector") ? to_double(a.get("w-detector")) : 0.0;
  if (!a.exist("w-detector")) {
    if (!n.empty()) {
      w = 2 * M_PI * .9 * radius[0] / n[0];
    } else if (a.exist("d-detector")) {
      const std::string shape = a.get("shape");
      if (shape == "circle") {
        w = d;
      } else {
        double mult = 1., sides = 0.;
        if (shape == "triangle")
          sides = 3.;
        else if (shape == "square")
          sides = 4.;
        else if (shape == "hexagon")
          sides = 6., mult = 2.;
        else
          throw std::runtime_error("cannot determine detector width for given shape");
        w = d * std::sin(M_PI / sides) * mult;
      }
    }
    std::ostringstream v;
    v.precision(17);
    v << w;
    a.value["w-detector"] = v.str();
    assumed << "--w-detector=" << w << "\n";
  }
  if (n.empty()) {
    const double ang = a.exist("d-detector") ? std::atan2(d, 2 * radius[0] + d / 2)
                                             : std::atan2(w, 2 * radius[0]);
    n.push_back(((int)std::floor(M_PI / ang) / 4) * 4);
    a.value["n-detectors"] = std::to_string(n[0]);
    assumed << "--n-detectors=" << n[0] << "\n";
  }
  if (a.exist("verbose") && !assumed.str().empty())
    std::cerr << "assumed:\n" << assumed.str();
}

std::vector<float> to_floats(const std::string& s) {
  std::vector<float> out;
  for (double v : to_doubles(s)) out.push_back((float)v);
  return out;
}

void convert_warsaw(const Args& a) {
  const float sigma_z = (float)to_double(a.get("s-z"));
  const float sigma_dl = (float)to_double(a.get("s-dl"));
  const float length = (float)to_double(a.get("length"));
  const int verbose = a.flag("verbose");
  std::normal_distribution<double> z_error(0, sigma_z);
  std::normal_distribution<double> dl_error(0, sigma_dl);
  std::mt19937_64 rng((std::mt19937_64::result_type)to_long(a.get("seed")));
  const std::vector<Detector> barrel = build_barrel(
      to_floats(a.get("radius")),
      a.exist("rotation") ? to_floats(a.get("rotation")) : std::vector<float>{},
      to_ints(a.get("n-detectors")), (float)to_double(a.get("w-detector")),
      a.exist("h-detector") ? (float)to_double(a.get("h-detector")) : 0.0f,
      a.exist("d-detector") ? (float)to_double(a.get("d-detector")) : 0.0f);
  const int only = (int)to_long(a.get("only"));
  const float half_length = length / 2;
  auto detector_of = [&](double x, double y) {
    const float px = (float)(x * kCm), py = (float)(y * kCm);
    int found = -1;
    for (size_t i = 0; i < barrel.size(); ++i)
      if (barrel[i].contains(px, py, 0.0001f)) found = (int)i;
    return found;
  };
  for (const auto& fn : a.rest) {
    std::ifstream in(fn);
    if (!in.is_open()) throw std::runtime_error("cannot open input means file: " + fn);
    std::string line;
    size_t n_columns = 0;
    while (std::getline(in, line)) {
      if (!n_columns) {
        // the first line fixes the column count: the first eight and the
        // last column are used
        std::istringstream cols(line);
        std::string tok;
        while (cols >> tok) ++n_columns;
        if (verbose) std::cerr << "reading " << n_columns << " column: " << fn << std::endl;
        if (!n_columns) break;
        if (n_columns < 9)
          throw std::runtime_error("expected at least 9 columns in " + fn);
      }
      std::istringstream in_line(line);
      double x1 = 0, y1 = 0, z1 = 0, t1 = 0, x2 = 0, y2 = 0, z2 = 0, t2 = 0;
      int scatter = 0;
      if (!(in_line >> x1 >> y1 >> z1 >> t1 >> x2 >> y2 >> z2 >> t2)) continue;
      std::string skip;
      for (size_t i = 0; i < n_columns - 9; ++i) in_line >> skip;
      in_line >> scatter;
      if (std::abs(z1 * kCm) > half_length || std::abs(z2 * kCm) > half_length) continue;
      if (only != 0 && only != scatter) continue;
      int d1 = detector_of(x1, y1), d2 = detector_of(x2, y2);
      if (d1 >= 0 && d2 >= 0) {
        if (d1 < d2) {
          std::swap(d1, d2);
          std::swap(z1, z2);
          std::swap(t1, t2);
        }
        const double dl = (t1 - t2) * kLightMPerPs;
        const double e_zu = z_error(rng);
        const double e_zd = z_error(rng);
        const double e_dl = dl_error(rng);
        std::cout << d1 << " " << d2 << " " << z1 * kCm + e_zu << " " << z2 * kCm + e_zd
                  << " " << dl + e_dl << "\n";
      } else if (verbose) {
        std::cerr << "point (" << (float)(x1 * kCm) << ' ' << (float)(y1 * kCm) << ") or ("
                  << (float)(x2 * kCm) << ' ' << (float)(y2 * kCm)
                  << ") outside detector - skiping\n";
      }
    }
  }
}

void convert_tbednarski(const Args& a) {
  const bool relative = a.exist("relative-time");
  for (const auto& fn : a.rest) {
    std::ifstream in(fn);
    if (!in.is_open()) throw std::runtime_error("cannot open input means file: " + fn);
    std::string line;
    while (std::getline(in, line)) {
      std::istringstream in_line(line);
      int du = 0, dd = 0;
      double tul = 0, tur = 0, tdl = 0, tdr = 0;
      if (!(in_line >> du >> dd >> tul >> tur >> tdl >> tdr)) continue;
      if (relative) {
        const double t0 = std::min(std::min(tur, tul), std::min(tdr, tdl));
        std::cout << du << ' ' << dd << ' ' << tul - t0 << ' ' << tur - t0 << ' '
                  << tdl - t0 << ' ' << tdr - t0 << '\n';
        continue;
      }
      if (du < dd) {
        std::swap(du, dd);
        std::swap(tul, tdl);
        std::swap(tur, tdr);
      }
      const double zu = 0.5 * kScintLightMPerPs * (tul - tur);
      const double zd = 0.5 * kScintLightMPerPs * (tdl - tdr);
      const double dl = 0.5 * kLightMPerPs * ((tul + tur) - (tdl + tdr));
      std::cout << du - 1 << ' ' << dd - 1 << ' ' << zu << ' ' << zd << ' ' << dl << '\n';
    }
  }
}

int run(int argc, char** argv) {
  Args a;
  a.specs = kOptions;
  std::vector<std::string> args(argv + 1, argv + argc);
  parse(a, args, false);
  if (a.exist("help")) {
    usage(a, "3d_tool_convert [options] means ...", nullptr);
    return 0;
  }
  // <input without extension>.cfg (src/util/cmdline_hooks.cpp:110-123)
  for (const auto& fn : a.rest) {
    const std::string cfg = strip_io::without_ext(fn) + ".cfg";
    std::ifstream probe(cfg);
    if (probe.is_open()) {
      probe.close();
      load_config(a, cfg, a.exist("verbose"));
    }
  }
  barrel_defaults(a, argc);
  if (a.exist("tbednarski") && !a.exist("warsaw"))
    convert_tbednarski(a);
  else
    convert_warsaw(a);
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
