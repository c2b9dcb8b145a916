// System matrices and the geometry built from them for the 3D hybrid
// reconstruction's --system options (host code; no device work).
//
// The reference's 3d_hybrid_reconstruction takes the LOR-pixel weights from
// 2D barrel system matrices (src/3d/hybrid/reconstruction_cmd.cpp:147-370):
//   - with --geometry, one matrix file, or one per half plane, whose hits /
//     emissions replace the geometry's weights plane by plane
//     (GeometrySOA::load_weights_from_matrix_file, src/2d/barrel/
//     geometry_soa.h:210-267);
//   - without it, the whole geometry comes from one matrix and the detector
//     centres of the scanner description file: pixel infos per LOR in the
//     matrix's order, positions along the segment between the two detector
//     centres, sorted by position, inactive detectors' LORs left empty
//     (geometry_soa.h:71-182, reconstruction_cmd.cpp:341-450).
// The arithmetic follows the reference's float expressions without
// contraction, and its sorts use the same comparisons, so the arrays equal
// the reference's own.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "petstrip_hybrid.h"

namespace petstrip {
int set_error(int code, const char* msg);
}
using petstrip::set_error;

struct petstrip_system_matrix {
  struct Element {
    int16_t first, second;  // LOR, first >= second
    uint32_t position;      // TOF position (0 without TOF)
    int16_t x, y;           // pixel
    int32_t hits;
  };
  uint32_t n_pixels_in_row = 0, n_emissions = 0, n_detectors = 0, n_tof = 1;
  std::vector<Element> elements;  // sorted by LOR, position, pixel y, x
};

namespace {

constexpr uint32_t fourcc(const char* s) {
  return (uint32_t)s[0] | ((uint32_t)s[1] << 8) | ((uint32_t)s[2] << 16) |
         ((uint32_t)s[3] << 24);
}

bool lor_less(int16_t af, int16_t as, int16_t bf, int16_t bs) {
  return af < bf || (af == bf && as < bs);
}

// SortByLORNPositionNPixel (sparse_matrix.h:584-592)
bool element_less(const petstrip_system_matrix::Element& a,
                  const petstrip_system_matrix::Element& b) {
  if (lor_less(a.first, a.second, b.first, b.second)) return true;
  if (lor_less(b.first, b.second, a.first, a.second)) return false;
  if (a.position < b.position) return true;
  if (a.position > b.position) return false;
  if (a.y < b.y) return true;
  if (a.y > b.y) return false;
  return a.x < b.x;
}

int64_t lor_index(int first, int second) { return (int64_t)first * (first + 1) / 2 + second; }

struct Seg {
  float sx, sy, ex, ey, dx, dy, len;
};

// LineSegment(start, end) (line_segment.h:27-34): direction (end - start)
// normalised by two divisions, length sqrt(x x + y y).
Seg make_segment(float sx, float sy, float ex, float ey) {
  Seg s{sx, sy, ex, ey, 0, 0, 0};
  const float vx = ex - sx, vy = ey - sy;
  const float x2 = vx * vx, y2 = vy * vy;
  const float len = std::sqrt(x2 + y2);
  s.dx = vx / len;
  s.dy = vy / len;
  s.len = len;
  return s;
}

}  // namespace

extern "C" {

int petstrip_system_matrix_read(const char* path, petstrip_system_matrix** out) {
  if (!path || !out) return set_error(PETSTRIP_EINVAL, "bad argument");
  *out = nullptr;
  std::ifstream in(path, std::ios::binary);
  if (!in.is_open())
    return set_error(PETSTRIP_EINVAL, (std::string("cannot open system matrix file: ") + path).c_str());
  auto rd32 = [&](uint32_t& v) { return (bool)in.read(reinterpret_cast<char*>(&v), 4); };
  auto rd16 = [&](uint16_t& v) { return (bool)in.read(reinterpret_cast<char*>(&v), 2); };
  uint32_t magic = 0;
  if (!rd32(magic)) return set_error(PETSTRIP_EINVAL, "system matrix: empty file");
  const bool full = magic == fourcc("PETP") || magic == fourcc("TOFP");
  const bool tof = magic == fourcc("TOFp") || magic == fourcc("TOFP");
  const bool known = full || tof || magic == fourcc("PETp") || magic == fourcc("PETt") ||
                     magic == fourcc("PETs");
  if (!known) return set_error(PETSTRIP_EINVAL, "invalid file type format");
  if (!full)
    return set_error(PETSTRIP_EINVAL,
                     "matrix must be in full form, convert using 2d_barrel_matrix --full option");
  auto* m = new petstrip_system_matrix();
  if (!rd32(m->n_pixels_in_row) || !rd32(m->n_emissions) || !rd32(m->n_detectors) ||
      (tof && !rd32(m->n_tof))) {
    delete m;
    return set_error(PETSTRIP_EINVAL, "system matrix: truncated header");
  }
  if (m->n_detectors < 1 || m->n_detectors > 32767 || m->n_pixels_in_row > 32767) {
    delete m;
    return set_error(PETSTRIP_EINVAL, "system matrix: bad sizes");
  }
  for (;;) {
    uint16_t a, b;
    if (!rd16(a) || !rd16(b)) break;  // end of the file
    uint32_t count;
    if (!rd32(count)) break;
    const int16_t fa = (int16_t)a, fb = (int16_t)b;
    for (uint32_t i = 0; i < count; ++i) {
      petstrip_system_matrix::Element e{};
      uint16_t x, y;
      uint32_t hits, position = 0;
      if ((tof && !rd32(position)) || !rd16(x) || !rd16(y) || !rd32(hits)) {
        delete m;
        return set_error(PETSTRIP_EINVAL, "system matrix: truncated element");
      }
      e.first = std::max(fa, fb);
      e.second = std::min(fa, fb);
      e.position = position;
      e.x = (int16_t)x;
      e.y = (int16_t)y;
      e.hits = (int32_t)hits;
      if (e.second < 0 || e.first >= (int)m->n_detectors) {
        delete m;
        return set_error(PETSTRIP_EINVAL, "system matrix: LOR beyond its detectors");
      }
      m->elements.push_back(e);
    }
  }
  // full matrices are read unsorted; the reference sorts them
  // (sort_by_lor_n_pixel, sparse_matrix.h:449-454)
  std::sort(m->elements.begin(), m->elements.end(), element_less);
  *out = m;
  return PETSTRIP_OK;
}

void petstrip_system_matrix_free(petstrip_system_matrix* m) { delete m; }

int petstrip_system_matrix_info(const petstrip_system_matrix* m, int32_t* n_pixels_in_row,
                                int32_t* n_detectors, int64_t* n_emissions,
                                int64_t* n_elements) {
  if (!m) return set_error(PETSTRIP_EINVAL, "bad argument");
  if (n_pixels_in_row) *n_pixels_in_row = (int32_t)m->n_pixels_in_row;
  if (n_detectors) *n_detectors = (int32_t)m->n_detectors;
  if (n_emissions) *n_emissions = (int64_t)m->n_emissions;
  if (n_elements) *n_elements = (int64_t)m->elements.size();
  return PETSTRIP_OK;
}

int petstrip_system_matrix_weights(const petstrip_system_matrix* m,
                                   const petstrip_hybrid_geometry* g, float* weights) {
  if (!m || !g || !weights || !g->info_begin || !g->info_end || !g->pixels)
    return set_error(PETSTRIP_EINVAL, "bad argument");
  const int64_t n_lors = (int64_t)(g->n_detectors + 1) * g->n_detectors / 2;
  std::memset(weights, 0, sizeof(float) * (size_t)g->n_pixel_infos);
  const float n_emissions = (float)m->n_emissions;
  int64_t cur = -1, pi = 0, pi_end = 0;
  std::vector<int64_t> idx;
  for (const auto& e : m->elements) {
    const int64_t l = lor_index(e.first, e.second);
    if (l >= n_lors)
      return set_error(PETSTRIP_EINVAL, "system matrix has more detectors than the geometry");
    if (l != cur) {
      cur = l;
      const int64_t b = g->info_begin[l], en = g->info_end[l];
      pi = 0;
      pi_end = en - b;
      idx.resize((size_t)pi_end);
      for (int64_t i = 0; i < pi_end; ++i) idx[(size_t)i] = i + b;
      // the LOR's pixel infos by pixel (y, then x), Pixel::operator<
      std::sort(idx.begin(), idx.end(), [g](int64_t a, int64_t c) {
        const int16_t ay = g->pixels[2 * a + 1], cy = g->pixels[2 * c + 1];
        return ay < cy || (ay == cy && g->pixels[2 * a] < g->pixels[2 * c]);
      });
    }
    auto less_than_elem = [&](int64_t i) {
      const int16_t py = g->pixels[2 * i + 1], px = g->pixels[2 * i];
      return py < e.y || (py == e.y && px < e.x);
    };
    while (pi < pi_end && less_than_elem(idx[(size_t)pi])) ++pi;
    if (pi == pi_end) continue;
    const int64_t i = idx[(size_t)pi];
    const int16_t py = g->pixels[2 * i + 1], px = g->pixels[2 * i];
    if (e.y < py || (e.y == py && e.x < px)) continue;  // no such pixel in the LOR
    weights[i] += (float)e.hits / n_emissions;
  }
  return PETSTRIP_OK;
}

int petstrip_read_detector_centres(const char* path, float* centres, int32_t capacity,
                                   int32_t* n_detectors) {
  if (!path || !n_detectors) return set_error(PETSTRIP_EINVAL, "bad argument");
  std::ifstream in(path);
  if (!in.is_open())
    return set_error(PETSTRIP_EINVAL,
                     (std::string("cannot open detector description file `") + path + "'")
                         .c_str());
  std::string line;
  int32_t n = 0;
  while (std::getline(in, line)) {
    std::istringstream sin(line);
    float x, y, cx = 0.0f, cy = 0.0f;
    int k = 0;
    while (sin >> x >> y) {
      cx += x;
      cy += y;
      ++k;
    }
    // Polygon::center (src/2d/geometry/polygon.h:56-68): float sums over
    // the vertices, divided by their count
    if (k > 0) {
      cx /= (float)k;
      cy /= (float)k;
    }
    if (centres && n < capacity) {
      centres[2 * n] = cx;
      centres[2 * n + 1] = cy;
    }
    ++n;
  }
  *n_detectors = n;
  return PETSTRIP_OK;
}

int petstrip_hybrid_geometry_from_matrix(const petstrip_system_matrix* m, const float* centres,
                                         int32_t n_columns, int32_t n_rows, double pixel_size,
                                         double lower_left_x, double lower_left_y,
                                         const int16_t* inactive, int32_t n_inactive,
                                         float* segments, float* widths, int64_t* info_begin,
                                         int64_t* info_end, int16_t* pixels, float* positions,
                                         float* weights, int64_t* n_pixel_infos) {
  if (!m || !n_pixel_infos || n_inactive < 0 || (n_inactive > 0 && !inactive))
    return set_error(PETSTRIP_EINVAL, "bad argument");
  const int nd = (int)m->n_detectors;
  std::vector<char> off((size_t)nd, 0);
  for (int32_t k = 0; k < n_inactive; ++k)
    if (inactive[k] >= 0 && inactive[k] < nd) off[(size_t)inactive[k]] = 1;
  auto skipped = [&](const petstrip_system_matrix::Element& e) {
    return n_inactive > 0 && (off[(size_t)e.first] || off[(size_t)e.second]);
  };
  int64_t n = 0;
  for (const auto& e : m->elements) n += skipped(e) ? 0 : 1;
  if (!segments) {  // sizes only
    *n_pixel_infos = n;
    return PETSTRIP_OK;
  }
  if (!centres || !widths || !info_begin || !info_end || !pixels || !positions || !weights)
    return set_error(PETSTRIP_EINVAL, "bad argument");
  const int64_t n_lors = (int64_t)(nd + 1) * nd / 2;
  for (int64_t l = 0; l < n_lors; ++l) {
    info_begin[l] = info_end[l] = 0;
    widths[l] = 0.0f;  // unused with system weights
  }
  // pixel infos in the matrix's order (LOR, position, pixel); an inactive
  // detector's LOR keeps an empty range (geometry_soa.h:84-119)
  const float n_emissions = (float)m->n_emissions;
  int64_t index = 0, cur = -1;
  for (const auto& e : m->elements) {
    const int64_t l = lor_index(e.first, e.second);
    if (l != cur) {
      if (cur >= 0) info_end[cur] = index;
      cur = l;
      info_begin[l] = index;
    }
    if (skipped(e)) continue;
    pixels[2 * index] = e.x;
    pixels[2 * index + 1] = e.y;
    weights[index] = (float)e.hits / n_emissions;
    ++index;
  }
  if (cur >= 0) info_end[cur] = index;
  // LOR segments from the second detector's centre to the first's, pixel
  // positions along them (projection_scaled of the pixel centre), sorted
  // (geometry_soa.h:137-181); PixelGrid::center_at (pixel_grid.h:54-57)
  const float ps = (float)pixel_size;
  const float llcx = (float)lower_left_x + ps / 2, llcy = (float)lower_left_y + ps / 2;
  std::vector<int64_t> idx;
  std::vector<int16_t> sp;
  std::vector<float> sw, st;
  for (int first = 0; first < nd; ++first) {
    for (int second = 0; second <= first; ++second) {
      const int64_t l = lor_index(first, second);
      const Seg s = make_segment(centres[2 * second], centres[2 * second + 1],
                                 centres[2 * first], centres[2 * first + 1]);
      segments[4 * l] = s.sx;
      segments[4 * l + 1] = s.sy;
      segments[4 * l + 2] = s.ex;
      segments[4 * l + 3] = s.ey;
      const int64_t b = info_begin[l], en = info_end[l], k = en - b;
      if (k <= 0) continue;
      idx.resize((size_t)k);
      for (int64_t i = 0; i < k; ++i) {
        const int64_t pi = b + i;
        const float cx = llcx + (float)pixels[2 * pi] * ps;
        const float cy = llcy + (float)pixels[2 * pi + 1] * ps;
        const float ax = (cx - s.sx) * s.dx, ay = (cy - s.sy) * s.dy;
        positions[pi] = (ax + ay) / s.len;
        idx[(size_t)i] = pi;
      }
      std::sort(idx.begin(), idx.end(),
                [positions](int64_t a, int64_t c) { return positions[a] < positions[c]; });
      sp.resize(2 * (size_t)k);
      sw.resize((size_t)k);
      st.resize((size_t)k);
      for (int64_t i = 0; i < k; ++i) {
        const int64_t j = idx[(size_t)i];
        sp[2 * i] = pixels[2 * j];
        sp[2 * i + 1] = pixels[2 * j + 1];
        sw[(size_t)i] = weights[j];
        st[(size_t)i] = positions[j];
      }
      for (int64_t i = 0; i < k; ++i) {
        pixels[2 * (b + i)] = sp[2 * i];
        pixels[2 * (b + i) + 1] = sp[2 * i + 1];
        weights[b + i] = sw[(size_t)i];
        positions[b + i] = st[(size_t)i];
      }
    }
  }
  (void)n_columns;
  (void)n_rows;
  *n_pixel_infos = index;
  return PETSTRIP_OK;
}

}  // extern "C"
