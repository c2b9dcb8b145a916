// Text response files, read with the reference's fast text reader semantics
// (host code; no device work).
//
// The reference commands read `.txt` responses through
// util::text_parser (src/util/text_parser.h:29-175), from
// Strip::Reconstruction::fast_load_txt_events
// (src/2d/strip/reconstruction.h:169-196; 3 fields, or "i j z_u z_d dl"
// with --3d-response) and from the 3D hybrid Reconstruction's
// (src/3d/hybrid/reconstruction.h:165-188; "lor.first lor.second z_up z_dn
// dl").  Its floats are not the correctly rounded values of their decimal
// strings, so the events it produces differ from a strtod reader's in the
// last bits, and a pixel decision near the 3-sigma boundary can flip.  This
// file restates its rules:
//   - lines end at '\n'; a last line without one is not read, and reading
//     stops at a line of 16384 or more bytes (the reader's buffer);
//   - fields are separated by ' ' only; characters after the last field
//     read are ignored;
//   - an optional sign, then decimal digits summed in a 32-bit int (which
//     wraps), for the whole part, the fraction and the exponent alike;
//   - a float is float(whole) + float(fraction) * N(d) with d fraction
//     digits, in one fused multiply-add (the reference is built with GCC at
//     -O3 -march=native, which contracts that expression on FMA hosts),
//     where N(0) = 1 and N(d) = 0.1f * N(d - 1) in float;
//   - an exponent e > 0 multiplies by 10 * N(e - 1) (so "1e2" reads as 1,
//     as in the reference), e < 0 by N(-e); |e| must stay below 300;
//   - the result is multiplied by the sign last;
//   - errors: a number without digits before or after its period, one
//     with neither a whole part nor a period, an exponent mark without
//     digits, an exponent out of range.  Integer fields (the LOR indices)
//     take no digits without error, as the reference's do.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "petstrip.h"

namespace petstrip {
int set_error(int code, const char* msg);
}

namespace {

constexpr int kMaxExp = 300;            // the reader's power tables
constexpr size_t kLineBuffer = 16384;   // the reader's line buffer

// N(d) = 0.1f^d by repeated float multiplication, and P(e) = 10 N(e - 1).
struct Powers {
  float neg[kMaxExp];
  float pos[kMaxExp];
  Powers() {
    neg[0] = 1.0f;
    pos[0] = 1.0f;
    for (int e = 1; e < kMaxExp; ++e) {
      neg[e] = 0.1f * neg[e - 1];
      pos[e] = 10.0f * neg[e - 1];
    }
  }
};

const Powers& powers() {
  static const Powers p;
  return p;
}

struct ParseError {
  const char* what;
};

inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

// Decimal digits summed in a wrapping 32-bit int.
inline int32_t digits(const char*& p, int& n) {
  uint32_t sum = 0;
  n = 0;
  while (is_digit(*p)) {
    sum = sum * 10u + (uint32_t)(*p - '0');
    ++p;
    ++n;
  }
  return (int32_t)sum;
}

inline int sign(const char*& p) {
  if (*p == '-') {
    ++p;
    return -1;
  }
  if (*p == '+') ++p;
  return 1;
}

inline void skip_spaces(const char*& p) {
  while (*p == ' ') ++p;
}

int32_t read_int(const char*& p) {
  skip_spaces(p);
  const int s = sign(p);
  int n;
  return (int32_t)((uint32_t)s * (uint32_t)digits(p, n));
}

float read_float(const char*& p) {
  skip_spaces(p);
  const Powers& pw = powers();
  const int s = sign(p);
  int n_whole;
  const int32_t whole = digits(p, n_whole);
  float v = (float)whole;
  if (*p == '.') {
    ++p;
    int n_frac;
    const int32_t frac = digits(p, n_frac);
    if (n_whole == 0 && n_frac == 0) throw ParseError{"period with no digits either before or after"};
    if (n_frac >= kMaxExp) throw ParseError{"too many fractional digits"};
    if (n_frac > 0) v = std::fmaf((float)frac, pw.neg[n_frac], v);
  } else if (n_whole == 0) {
    throw ParseError{"neither whole part nor period"};
  }
  if (*p == 'e' || *p == 'E') {
    ++p;
    const int es = sign(p);
    int n_exp;
    const int32_t e = (int32_t)((uint32_t)es * (uint32_t)digits(p, n_exp));
    if (n_exp == 0) throw ParseError{"found exponent char but no exponent"};
    if (e >= kMaxExp || e <= -kMaxExp) throw ParseError{"exponent out of range"};
    if (e > 0)
      v *= pw.pos[e];
    else if (e < 0)
      v *= pw.neg[-e];
  }
  return v * (float)s;
}

// Calls line(ptr) for every '\n'-terminated line of the file (the newline
// replaced by 0) with the reader's rules for the last line and for long
// lines.  Returns false (errno set) if the file cannot be read.
template <typename F>
bool for_each_line(const char* path, F&& line) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::vector<char> buf(1 << 20);
  size_t have = 0;
  bool ok = true;
  for (;;) {
    const size_t got = std::fread(buf.data() + have, 1, buf.size() - have, f);
    if (got == 0) {
      ok = !std::ferror(f);
      break;
    }
    char* const end = buf.data() + have + got;
    char* ln = buf.data();
    char* nl;
    bool stop = false;
    while ((nl = static_cast<char*>(std::memchr(ln, '\n', (size_t)(end - ln))))) {
      if ((size_t)(nl - ln) >= kLineBuffer) {
        stop = true;
        break;
      }
      *nl = 0;
      line(ln);
      ln = nl + 1;
    }
    have = (size_t)(end - ln);
    if (stop || have >= kLineBuffer) break;  // a line the reader cannot hold
    std::memmove(buf.data(), ln, have);
  }
  std::fclose(f);
  return ok;
}

}  // namespace

extern "C" int petstrip_read_text_responses(const char* path, int32_t columns,
                                            float* responses, int16_t* lors,
                                            int64_t capacity, int64_t* n_out) {
  using petstrip::set_error;
  if (!path || !n_out) return set_error(PETSTRIP_EINVAL, "path and n_out are required");
  if (columns != 3 && columns != 5)
    return set_error(PETSTRIP_EINVAL, "columns must be 3 or 5");
  if (responses && capacity < 0) return set_error(PETSTRIP_EINVAL, "negative capacity");
  int64_t n = 0;
  std::string bad, why;
  const bool ok = for_each_line(path, [&](const char* text) {
    if (!bad.empty()) return;
    if (!responses) {
      ++n;
      return;
    }
    if (n >= capacity) {
      bad = text;
      why = "more lines than the capacity";
      return;
    }
    const char* p = text;
    try {
      int32_t i = 0, j = 0;
      if (columns == 5) {
        i = read_int(p);
        j = read_int(p);
      }
      const float zu = read_float(p);
      const float zd = read_float(p);
      const float dl = read_float(p);
      responses[3 * n] = zu;
      responses[3 * n + 1] = zd;
      responses[3 * n + 2] = dl;
      if (lors) {
        lors[2 * n] = (int16_t)i;
        lors[2 * n + 1] = (int16_t)j;
      }
      ++n;
    } catch (const ParseError& e) {
      bad = text;
      why = e.what;
    }
  });
  if (!ok) {
    return set_error(PETSTRIP_EINVAL, (std::string("cannot read input file: ") + path + ": " +
                                       std::strerror(errno)).c_str());
  }
  if (!bad.empty())
    return set_error(PETSTRIP_EINVAL, ("error line: " + bad + " (" + why + ")").c_str());
  *n_out = n;
  return PETSTRIP_OK;
}

extern "C" int petstrip_parse_text_floats(const char* text, int32_t n, float* out,
                                          int32_t* consumed) {
  using petstrip::set_error;
  if (!text || !out || n < 0) return set_error(PETSTRIP_EINVAL, "bad argument");
  const char* p = text;
  try {
    for (int32_t k = 0; k < n; ++k) out[k] = read_float(p);
  } catch (const ParseError& e) {
    return set_error(PETSTRIP_EINVAL, e.what);
  }
  if (consumed) *consumed = (int32_t)(p - text);
  return PETSTRIP_OK;
}
