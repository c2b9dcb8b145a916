// File formats of the strip reconstruction command line (see strip_io.h).
#include "strip_io.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace strip_io {

namespace {

size_t sep_pos(const std::string& p) { return p.find_last_of("\\/"); }

size_t ext_pos(const std::string& p) {
  size_t e = p.find_last_of('.');
  size_t s = sep_pos(p);
  if (s != std::string::npos && e != std::string::npos && s > e) return std::string::npos;
  return e;
}

// CRC-32 (IEEE 802.3, reflected 0xEDB88320) for PNG chunks.
uint32_t crc32_update(uint32_t crc, const unsigned char* p, size_t n) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xff] ^ (crc >> 8);
  return crc;
}

void put_be32(std::vector<unsigned char>& v, uint32_t x) {
  v.push_back((unsigned char)(x >> 24));
  v.push_back((unsigned char)(x >> 16));
  v.push_back((unsigned char)(x >> 8));
  v.push_back((unsigned char)x);
}

void put_chunk(std::ofstream& out, const char* type, const std::vector<unsigned char>& data) {
  std::vector<unsigned char> buf;
  put_be32(buf, (uint32_t)data.size());
  size_t start = buf.size();
  buf.insert(buf.end(), type, type + 4);
  buf.insert(buf.end(), data.begin(), data.end());
  uint32_t crc = crc32_update(0xffffffffu, buf.data() + start, buf.size() - start) ^ 0xffffffffu;
  put_be32(buf, crc);
  out.write(reinterpret_cast<const char*>(buf.data()), (std::streamsize)buf.size());
}

}  // namespace

std::string without_ext(const std::string& p) { return p.substr(0, ext_pos(p)); }

std::string ext(const std::string& p) {
  size_t e = ext_pos(p);
  return e == std::string::npos ? std::string() : p.substr(e);
}

std::string without_path(const std::string& p) {
  size_t s = sep_pos(p);
  return s == std::string::npos ? p : p.substr(s + 1);
}

std::string add_index(const std::string& base, int index, int max_value) {
  int digits = 1;
  if (max_value > 0) {
    long lim = 10;
    while (lim <= max_value) {
      ++digits;
      lim *= 10;
    }
  } else {
    digits = max_value;
  }
  char buf[64];
  std::snprintf(buf, sizeof buf, "_%0*d", std::max(digits, 1), index);
  return base + buf;
}

void load_binary_responses(const std::string& fn, std::vector<float>& out) {
  std::ifstream in(fn, std::ios::binary);
  if (!in.is_open()) throw std::runtime_error("cannot open phantom responses file: " + fn);
  in.seekg(0, std::ios::end);
  std::streamoff size = in.tellg();
  in.seekg(0, std::ios::beg);
  size_t n = (size_t)size / (3 * sizeof(float));  // a trailing partial triple is ignored
  size_t old = out.size();
  out.resize(old + 3 * n);
  in.read(reinterpret_cast<char*>(out.data() + old), (std::streamsize)(3 * n * sizeof(float)));
  if ((size_t)in.gcount() != 3 * n * sizeof(float))
    throw std::runtime_error("short read: " + fn);
}

void write_raw(const std::string& fn, const float* img, size_t n) {
  std::ofstream out(fn, std::ios::binary);
  if (!out.is_open()) throw std::runtime_error("cannot write " + fn);
  out.write(reinterpret_cast<const char*>(img), (std::streamsize)(n * sizeof(float)));
}

void write_text_matrix(const std::string& fn, const float* img, int width, int height) {
  std::ofstream out(fn);
  if (!out.is_open()) throw std::runtime_error("cannot write " + fn);
  size_t n = (size_t)width * height;
  for (size_t i = 0; i < n; ++i) {
    out << img[i];
    out << (((i + 1) % (size_t)width == 0) ? "\n" : " ");
  }
}

void write_text_tuples(const std::string& fn, const float* img, int width, int height) {
  std::ofstream out(fn);
  if (!out.is_open()) throw std::runtime_error("cannot write " + fn);
  out.precision(12);
  out << std::fixed;
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      float v = img[(size_t)y * width + x];
      if (v >= 1e-12f) out << x << ' ' << y << ' ' << v << '\n';
    }
}

void write_nrrd(const std::string& fn, const std::string& raw_name, int width, int height,
                bool ascii) {
  std::ofstream out(fn);
  if (!out.is_open()) throw std::runtime_error("cannot write " + fn);
  out << "NRRD0001\n"
      << "type: float\n"
      << "endian: little\n"
      << "dimension: 2\n"
      << "sizes: " << width << ' ' << height << '\n'
      << "data file: ./" << raw_name << '\n'
      << "encoding: " << (ascii ? "ascii" : "raw") << '\n';
}

void write_nrrd_typed(const std::string& fn, const std::string& raw_name, const char* type,
                      const std::vector<size_t>& sizes, bool ascii) {
  std::ofstream out(fn);
  if (!out.is_open()) throw std::runtime_error("cannot write " + fn);
  out << "NRRD0001\n"
      << "type: " << type << '\n'
      << "endian: little\n"
      << "dimension: " << sizes.size() << '\n'
      << "sizes:";
  for (size_t v : sizes) out << ' ' << v;
  out << '\n'
      << "data file: ./" << raw_name << '\n'
      << "encoding: " << (ascii ? "ascii" : "raw") << '\n';
}

void write_raw_bytes(const std::string& fn, const void* data, size_t bytes) {
  std::ofstream out(fn, std::ios::binary);
  if (!out.is_open()) throw std::runtime_error("cannot write " + fn);
  out.write(reinterpret_cast<const char*>(data), (std::streamsize)bytes);
}

void write_png(const std::string& fn, const float* img, int width, int height,
               double scale_max) {
  std::ofstream out(fn, std::ios::binary);
  if (!out.is_open()) throw std::runtime_error("cannot write " + fn);
  double vmax = scale_max;
  size_t n = (size_t)width * height;
  if (vmax == 0)
    for (size_t i = 0; i < n; ++i) vmax = std::max(vmax, (double)img[i]);
  double gain = 255.0 / vmax;
  // raw scanlines: filter byte 0 + width pixels
  std::vector<unsigned char> raw;
  raw.reserve((size_t)height * (width + 1));
  for (int y = 0; y < height; ++y) {
    raw.push_back(0);
    for (int x = 0; x < width; ++x) {
      double v = 255.0 - gain * img[(size_t)y * width + x];
      v = (v > 0.0) ? v : 0.0;  // max(0, .), NaN -> 0
      raw.push_back((unsigned char)std::min(v, 255.0));
    }
  }
  static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
  out.write(reinterpret_cast<const char*>(sig), 8);
  std::vector<unsigned char> ihdr;
  put_be32(ihdr, (uint32_t)width);
  put_be32(ihdr, (uint32_t)height);
  ihdr.push_back(8);  // bit depth
  ihdr.push_back(0);  // grayscale
  ihdr.push_back(0);
  ihdr.push_back(0);
  ihdr.push_back(0);
  put_chunk(out, "IHDR", ihdr);
  // zlib stream of stored deflate blocks
  std::vector<unsigned char> z;
  z.push_back(0x78);
  z.push_back(0x01);
  size_t pos = 0;
  do {
    size_t len = std::min<size_t>(65535, raw.size() - pos);
    bool last = pos + len == raw.size();
    z.push_back(last ? 1 : 0);
    z.push_back((unsigned char)(len & 0xff));
    z.push_back((unsigned char)(len >> 8));
    z.push_back((unsigned char)(~len & 0xff));
    z.push_back((unsigned char)((~len >> 8) & 0xff));
    z.insert(z.end(), raw.begin() + (std::ptrdiff_t)pos, raw.begin() + (std::ptrdiff_t)(pos + len));
    pos += len;
  } while (pos < raw.size());
  uint32_t a = 1, b = 0;  // Adler-32
  for (unsigned char c : raw) {
    a = (a + c) % 65521;
    b = (b + a) % 65521;
  }
  put_be32(z, (b << 16) | a);
  put_chunk(out, "IDAT", z);
  put_chunk(out, "IEND", {});
}

void load_raw_image(const std::string& fn, float* out, size_t n, bool as_int) {
  std::ifstream in(fn, std::ios::binary);
  if (!in.is_open()) throw std::runtime_error("cannot open rho file: " + fn);
  if (as_int) {
    std::vector<int32_t> v(n);
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * 4));
    for (size_t i = 0; i < n; ++i) out[i] = (float)(v[i] + 1000);
  } else {
    in.read(reinterpret_cast<char*>(out), (std::streamsize)(n * 4));
  }
}

}  // namespace strip_io
