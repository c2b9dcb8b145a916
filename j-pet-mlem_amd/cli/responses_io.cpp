// Text responses for the commands that link libpetstrip_hip: the library's
// reader follows the reference's fast text parser (src/util/text_parser.h:
// 29-175, src/2d/strip/reconstruction.h:169-196, src/3d/hybrid/
// reconstruction.h:165-188), so these commands reconstruct the floats the
// reference reads, not strtod's.
#include <stdexcept>

#include "../../include/petstrip.h"
#include "strip_io.h"

namespace strip_io {

void load_text_responses(const std::string& fn, bool is_3d, std::vector<float>& out,
                         std::vector<int16_t>* lors) {
  const int cols = is_3d ? 5 : 3;
  int64_t n = 0;
  if (petstrip_read_text_responses(fn.c_str(), cols, nullptr, nullptr, 0, &n) != 0)
    throw std::runtime_error(petstrip_last_error());
  const size_t old = out.size();
  out.resize(old + 3 * (size_t)n);
  size_t old_l = 0;
  if (lors) {
    old_l = lors->size();
    lors->resize(old_l + 2 * (size_t)n);
  }
  int64_t got = 0;
  if (petstrip_read_text_responses(fn.c_str(), cols, out.data() + old,
                                   lors ? lors->data() + old_l : nullptr, n, &got) != 0)
    throw std::runtime_error(petstrip_last_error());
  out.resize(old + 3 * (size_t)got);
  if (lors) lors->resize(old_l + 2 * (size_t)got);
}

}  // namespace strip_io
