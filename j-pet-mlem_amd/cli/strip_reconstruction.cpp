// 2d_strip_reconstruction: list-mode MLEM for the 2D strip PET scanner on
// an AMD Instinct GPU (libpetstrip_hip.so).
//
// Keeps the command line and file conventions of the reference tool
// (grey277/j-pet-mlem src/2d/strip/reconstruction_cmd.cpp:105-301,
// options.cpp:13-124, common/options.cpp:5-23):
//   2d_strip_reconstruction [options] response...
// Every run executes on the GPU.  With -G/--gpu the outputs follow the
// reference's GPU branch (sensitivity map, images at iterations 1-5, 10,
// 15, 20, 30, 50, 100 of the first block and at every block end, .txt as a
// matrix, always .nrrd); without it they follow the CPU branch (images at
// block ends, --png-max, .txt as "x y value" tuples).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <set>
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
    {"config", 'c', true, "", true, kSave, "load config file"},
    {"radius", 'r', true, "0.5", true, kSave, "R distance between scintillator"},
    {"length", 'l', true, "1", true, kSave, "scintillator length"},
    {"s-pixel", 'p', true, "0.01", true, kSave, "pixel size"},
    {"n-pixels", 'n', true, "", true, kSave, "number of pixels"},
    {"n-z-pixels", 0, true, "", true, kSave, "number of z pixels"},
    {"n-y-pixels", 0, true, "", true, kSave, "number of y pixels"},
    {"s-z", 0, true, "0.015", true, kSave, "TOF sigma along z axis"},
    {"s-dl", 0, true, "0.06", true, kSave, "TOF sigma delta-l"},
    {"n-emissions", 'e', true, "0", false, kSave, "number of emissions"},
    {"verbose", 'v', false, "", true, kSave, "print progress information (-v) or benchmark (-vv)"},
    {"n-threads", 'T', true, "", true, kSave, "number of host threads"},
    {"3d-response", 0, false, "", true, kSave, "input response contains additional LOR information"},
    {"dl-is-time", 0, false, "", true, kSave, "delta is time, convert using speed of light"},
    {"speed-of-light", 0, true, "299792458", true, kSave, "speed of light"},
    {"scale-length", 0, true, "1", true, kSave, "scale input length values"},
    {"scale-time", 0, true, "1", true, kSave, "scale input time values"},
    {"blocks", 'i', true, "0", true, kSave, "number of iteration blocks"},
    {"iterations", 'I', true, "1", true, kSave, "number of iterations (per block)"},
    {"output", 'o', true, "", false, kSave, "output files prefix (png)"},
    {"png-max", 0, true, "0", true, kSave, "maximum value mapped to 255 in PNG"},
    {"gpu", 'G', false, "", true, kSave, "GPU output conventions (always runs on the GPU)"},
    {"cuda-device", 'D', true, "0", true, kSave, "GPU device"},
    {"cuda-blocks", 'B', true, "64", true, kSave, "ignored (kept for compatibility)"},
    {"cuda-threads", 'W', true, "512", true, kSave, "ignored (kept for compatibility)"},
    {"rho", 0, true, "", true, kSave, "start rho (eg. existing iteration)"},
    {"rho-is-int", 0, false, "", true, kSave, "start rho uses integer representation (not float)"},
    {"accumulation", 0, true, "fp32", true, kSave, "fp32 | compensated | fixed64"},
    {"update", 0, true, "cpu", true, kSave,
     "cpu (Gaussian kernel, the CPU update) | gpu (analytic kernel, 5-point "
     "sensitivity in the denominator: the reference --gpu iteration)"},
    {"help", 'h', false, "", false, kSave, "print this message"},
};


// Benchmark / estimate progress in the reference's format (src/util/progress.h).
struct Progress {
  int verbosity;
  int total;
  std::chrono::steady_clock::time_point t0, start;
  double total_ms = 0;
  Progress(int v, int total) : verbosity(v), total(total) {
    t0 = start = std::chrono::steady_clock::now();
    if (verbosity >= 2) std::cout << "# it     time (ms)" << std::endl;
  }
  void operator()(int completed, bool finished) {
    if (!verbosity) return;
    auto now = std::chrono::steady_clock::now();
    if (verbosity >= 2) {
      if (!finished) {
        start = now;
        return;
      }
      double ms = std::chrono::duration<double, std::milli>(now - start).count();
      total_ms += ms;
      std::cout << std::fixed << std::setprecision(2) << std::setw(4) << completed + 1 << ' '
                << std::setw(10) << ms << std::endl;
      if (completed == total - 1)
        std::cout << "# av " << std::setw(10) << total_ms / total << "   total " << total_ms
                  << std::endl;
      std::cout.unsetf(std::ios::fixed);
      return;
    }
    if (finished && completed < total - 1) return;
    int done = finished ? completed + 1 : completed;
    double el = std::chrono::duration<double>(now - t0).count();
    std::cerr << ' ' << std::lround(100.0 * done / std::max(total, 1)) << "% " << done << "/"
              << total;
    if (done > 0 && el > 0) {
      double per = done / el;
      auto fmt = [](long s) {
        char b[32];
        std::snprintf(b, sizeof b, "%2ld:%02ld:%02ld", s / 3600, (s / 60) % 60, s % 60);
        return std::string(b);
      };
      std::cerr << " " << fmt(std::lround((total - done) / per)) << " left, "
                << fmt(std::lround(total / per)) << " total  ";
    }
    std::cerr << ((finished && done == total) ? "\n" : "\r");
  }
};

struct RunCtx {
  bool gpu_style;
  std::string base, oext;
  bool txt;
  int n_iterations;
  double png_max;
  Progress* progress;
};

void write_outputs(const RunCtx& c, int iteration, const float* img, int W, int H) {
  if (c.base.empty()) return;
  if (c.gpu_style) {
    // reference GPU branch (reconstruction_cmd.cpp:232-252)
    std::string fn = iteration >= 0 ? strip_io::add_index(c.base, iteration, c.n_iterations)
                                    : c.base + "_sensitivity";
    strip_io::write_png(fn + ".png", img, W, H, 0);
    if (c.txt)
      strip_io::write_text_matrix(fn + ".txt", img, W, H);
    else if (c.oext != ".png")
      strip_io::write_raw(fn + c.oext, img, (size_t)W * H);
    strip_io::write_nrrd(fn + ".nrrd", strip_io::without_path(fn) + c.oext, W, H, c.txt);
    return;
  }
  // reference CPU branch (reconstruction_cmd.cpp:265-291)
  std::string fn = strip_io::add_index(c.base, iteration, c.n_iterations);
  strip_io::write_png(fn + ".png", img, W, H, c.png_max);
  if (c.txt) {
    strip_io::write_text_tuples(fn + ".txt", img, W, H);
  } else if (c.oext != ".png") {
    strip_io::write_raw(fn + c.oext, img, (size_t)W * H);
    strip_io::write_nrrd(fn + ".nrrd", strip_io::without_path(fn) + c.oext, W, H, false);
  }
}

int on_output(void* user, int iteration, const float* img, int32_t W, int32_t H) {
  // only the GPU-branch conventions go through petstrip_run's schedule
  write_outputs(*static_cast<RunCtx*>(user), iteration, img, W, H);
  return 0;
}

void on_progress(void* user, int completed, int finished) {
  auto* c = static_cast<RunCtx*>(user);
  (*c->progress)(completed, finished != 0);
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
    usage(a, "2d_strip_reconstruction [options] response ...",
          "note: lengths are in the units of the responses (normally metres).");
    return 0;
  }
  if (a.rest.empty()) throw std::runtime_error("at least one responses input file expected, consult --help");
  // derived pixel counts (options.cpp:106-124), before the accompanying config
  derive_pixel_counts(a);
  // <response without extension>.cfg (cmdline_hooks.cpp:110-123)
  for (auto& fn : a.rest) {
    std::string cfg = strip_io::without_ext(fn) + ".cfg";
    std::ifstream probe(cfg);
    if (probe.is_open()) {
      probe.close();
      load_config(a, cfg, a.exist("verbose"));
    }
  }
  const int verbose = a.count.count("verbose") ? a.count["verbose"] : 0;

  petstrip_scanner sc{};
  sc.radius = to_doubles(a.get("radius")).at(0);
  sc.scintillator_length = to_double(a.get("length"));
  sc.n_y_pixels = (int32_t)to_long(a.get("n-y-pixels"));
  sc.n_z_pixels = (int32_t)to_long(a.get("n-z-pixels"));
  sc.pixel_height = sc.pixel_width = to_double(a.get("s-pixel"));
  sc.sigma_z = to_double(a.get("s-z"));
  sc.sigma_dl = to_double(a.get("s-dl"));
  const int W = sc.n_z_pixels, H = sc.n_y_pixels;
  if (verbose) std::cout << "# image: " << H << "x" << W << std::endl;

  std::vector<float> rho((size_t)W * H, 1.0f);
  if (a.value.count("rho")) {
    std::string rn = a.get("rho");
    if (strip_io::ext(rn) == ".txt") {
      // the reference's text --rho reads into copies and leaves rho at 1
      // (src/2d/geometry/pixel_map.h:127-131)
    } else {
      strip_io::load_raw_image(rn, rho.data(), rho.size(), a.exist("rho-is-int"));
    }
  }

  const bool is_3d = a.exist("3d-response");
  std::vector<float> resp;
  for (auto& fn : a.rest) {
    if (strip_io::ext(fn) == ".txt") {
      strip_io::load_text_responses(fn, is_3d, resp);
    } else {
      if (is_3d) throw std::runtime_error("3D input must have .txt extension");
      strip_io::load_binary_responses(fn, resp);
    }
    if (verbose) std::cerr << "# read " << resp.size() / 3 << " responsess from " << fn << std::endl;
    const bool dl_is_time = a.exist("dl-is-time");
    if (a.value.count("scale-length")) {
      double scale = to_double(a.get("scale-length"));
      for (size_t i = 0; i < resp.size(); i += 3) {
        resp[i] = (float)(resp[i] * scale);
        resp[i + 1] = (float)(resp[i + 1] * scale);
        if (!dl_is_time) resp[i + 2] = (float)(resp[i + 2] * scale);
      }
    }
    if (dl_is_time) {
      double scale = to_double(a.get("scale-time")), c = to_double(a.get("speed-of-light"));
      for (size_t i = 2; i < resp.size(); i += 3) resp[i] = (float)(resp[i] * c * scale);
    }
  }

  const int n_blocks = (int)to_long(a.get("blocks"));
  const int n_iter = (int)to_long(a.get("iterations"));
  const std::string out_name = a.get("output");
  RunCtx ctx;
  ctx.gpu_style = a.exist("gpu");
  ctx.base = strip_io::without_ext(out_name);
  ctx.oext = strip_io::ext(out_name);
  ctx.txt = ctx.oext == ".txt";
  ctx.n_iterations = n_blocks * n_iter;
  ctx.png_max = to_double(a.get("png-max"));
  Progress progress(verbose, ctx.n_iterations);
  ctx.progress = &progress;

  petstrip_options opts{};
  std::string accm = a.get("accumulation");
  opts.accumulation = accm == "fixed64" ? PETSTRIP_ACCUM_FIXED64
                      : accm == "compensated" ? PETSTRIP_ACCUM_COMPENSATED
                                              : PETSTRIP_ACCUM_FP32;
  std::string upd = a.get("update");
  if (upd == "gpu") {
    opts.update = PETSTRIP_UPDATE_SENSITIVITY_DENOMINATOR;
    opts.kernel = PETSTRIP_KERNEL_ANALYTIC;
  } else if (upd != "cpu") {
    throw std::runtime_error("--update must be cpu or gpu");
  }
  petstrip_handle* h = nullptr;
  check(petstrip_create(&sc, (int)to_long(a.get("cuda-device")), &opts, &h));
  struct Guard {
    petstrip_handle* h;
    ~Guard() { petstrip_destroy(h); }
  } guard{h};
  if (ctx.gpu_style && verbose) {
    char name[256];
    petstrip_device_name(h, name, sizeof name);
    std::cerr << "# device: " << name << std::endl;
  }
  check(petstrip_set_events(h, resp.data(), (int64_t)(resp.size() / 3)));

  if (ctx.gpu_style) {
    check(petstrip_run(h, rho.data(), n_blocks, n_iter, on_output, on_progress, &ctx));
  } else {
    if (!ctx.base.empty()) {
      std::vector<float> sens((size_t)W * H);
      check(petstrip_sensitivity(h, 0, sens.data()));
      strip_io::write_png(ctx.base + "_sensitivity.png", sens.data(), W, H, 0);
    }
    check(petstrip_set_rho(h, rho.data()));
    for (int block = 0; block < n_blocks; ++block) {
      for (int i = 0; i < n_iter; ++i) {
        progress(block * n_iter + i, false);
        check(petstrip_iterate(h, 1));
        progress(block * n_iter + i, true);
      }
      check(petstrip_get_rho(h, rho.data()));
      write_outputs(ctx, (block + 1) * n_iter, rho.data(), W, H);
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
