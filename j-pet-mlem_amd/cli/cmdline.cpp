#include "cmdline.h"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace strip_cli {

const OptSpec* Args::find(const std::string& n) const {
  for (auto& o : specs)
    if (n == o.name) return &o;
  return nullptr;
}

const OptSpec* Args::find_short(char c) const {
  for (auto& o : specs)
    if (o.short_name == c) return &o;
  return nullptr;
}

std::string Args::get(const std::string& n) const {
  auto it = value.find(n);
  if (it != value.end()) return it->second;
  auto* o = find(n);
  return o ? o->def : "";
}

long to_long(const std::string& s) {
  char* end = nullptr;
  long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str()) throw std::runtime_error("bad integer: " + s);
  if (*end) {
    if (end[1]) throw std::runtime_error("bad integer: " + s);
    switch (std::tolower(*end)) {
      case 'k': v *= 1000; break;
      case 'm': v *= 1000000; break;
      case 'g': v *= 1000000000; break;
      default: throw std::runtime_error("bad integer: " + s);
    }
  }
  return v;
}

double to_double(const std::string& s) {
  char* end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end) throw std::runtime_error("bad number: " + s);
  return v;
}

std::vector<double> to_doubles(const std::string& s) {
  std::vector<double> v;
  std::stringstream ss(s);
  std::string el;
  while (std::getline(ss, el, ',')) v.push_back(to_double(el));
  return v;
}

namespace {

void set_option(Args& a, const OptSpec* o, const std::string& val, bool from_file) {
  if (from_file && !o->from_file) return;  // not_from_file options keep their value
  if (!o->takes_value) {
    a.count[o->name] += 1;
    return;
  }
  if (std::string(o->name) == "config") {
    load_config(a, val, a.exist("verbose"));
    return;
  }
  a.value[o->name] = val;
}

// An error in a config file is a warning there (cmdline_hooks.cpp:86-95).
void problem(std::vector<std::string>* warn, const std::string& msg) {
  if (!warn) throw std::runtime_error(msg);
  warn->push_back(msg);
}

void parse_into(Args& a, const std::vector<std::string>& argv, bool from_file,
                std::vector<std::string>* warn) {
  for (size_t i = 0; i < argv.size(); ++i) {
    const std::string& s = argv[i];
    if (s.size() > 2 && s[0] == '-' && s[1] == '-') {
      size_t eq = s.find('=');
      std::string name = s.substr(2, eq == std::string::npos ? std::string::npos : eq - 2);
      const OptSpec* o = a.find(name);
      if (!o) {
        problem(warn, "undefined option: --" + name);
        continue;
      }
      if (!o->takes_value) {
        set_option(a, o, "", from_file);
      } else if (eq != std::string::npos) {
        set_option(a, o, s.substr(eq + 1), from_file);
      } else {
        if (i + 1 >= argv.size()) {
          problem(warn, "option needs value: --" + name);
          continue;
        }
        set_option(a, o, argv[++i], from_file);
      }
    } else if (s.size() > 1 && s[0] == '-' && !std::isdigit((unsigned char)s[1]) && s[1] != '.') {
      for (size_t j = 1; j < s.size(); ++j) {
        const OptSpec* o = a.find_short(s[j]);
        if (!o) {
          problem(warn, std::string("undefined short option: -") + s[j]);
          break;
        }
        if (!o->takes_value) {
          set_option(a, o, "", from_file);
          continue;
        }
        if (j + 1 < s.size()) {
          set_option(a, o, s.substr(j + 1), from_file);
        } else if (i + 1 >= argv.size()) {
          problem(warn, std::string("option needs value: -") + s[j]);
        } else {
          set_option(a, o, argv[++i], from_file);
        }
        break;
      }
    } else {
      a.rest.push_back(s);
    }
  }
}

}  // namespace

void parse(Args& a, const std::vector<std::string>& argv, bool from_file) {
  parse_into(a, argv, from_file, nullptr);
}

void load_config(Args& a, const std::string& arg, bool verbose) {
  std::string path = arg;
  if (!a.dir_stack.empty() && !path.empty() && path[0] != '/') path = a.dir_stack.back() + path;
  if (verbose) std::cout << "# load: " << path << std::endl;
  std::ifstream in(path);
  if (!in.is_open()) throw std::runtime_error("cannot open input config file: " + path);
  size_t sep = path.find_last_of('/');
  a.dir_stack.push_back(sep == std::string::npos ? std::string() : path.substr(0, sep + 1));
  std::vector<std::string> args;
  std::string line;
  while (std::getline(in, line))
    if (!line.empty() && line[0] != '#') args.push_back(line);
  size_t rest_before = a.rest.size();
  std::vector<std::string> warn;
  parse_into(a, args, true, &warn);
  a.rest.resize(rest_before);  // positional lines in config files are ignored
  a.dir_stack.pop_back();
  if (!warn.empty()) {
    std::cerr << (warn.size() == 1 ? "warning" : "warnings") << " parsing " << path << ":"
              << std::endl;
    for (auto& w : warn) std::cerr << "  " << w << std::endl;
  }
}

void usage(const Args& a, const char* synopsis, const char* note) {
  std::cerr << "usage: " << synopsis << "\noptions:\n";
  for (auto& o : a.specs) {
    std::cerr << "  ";
    if (o.short_name)
      std::cerr << '-' << o.short_name << ", ";
    else
      std::cerr << "    ";
    std::cerr << "--" << std::left << std::setw(16) << o.name << o.help;
    if (o.def[0]) std::cerr << " (=" << o.def << ")";
    std::cerr << '\n';
  }
  if (note && note[0]) std::cerr << note << '\n';
}

void derive_pixel_counts(Args& a) {
  if (a.value.count("n-pixels")) {
    a.value["n-z-pixels"] = a.value["n-pixels"];
    a.value["n-y-pixels"] = a.value["n-pixels"];
    return;
  }
  auto radius = to_doubles(a.get("radius"));
  if (radius.empty()) throw std::runtime_error("at least one --radius has to be specified");
  if (radius.size() > 1) std::cerr << "warning: more than one radius specified with --radius\n";
  double L = to_double(a.get("length")), s = to_double(a.get("s-pixel"));
  if (!a.value.count("n-z-pixels")) a.value["n-z-pixels"] = std::to_string((long)std::ceil(L / s));
  if (!a.value.count("n-y-pixels"))
    a.value["n-y-pixels"] = std::to_string((long)std::ceil(2 * radius[0] / s));
}

void write_config(const Args& a, const std::string& path) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write " + path);
  for (auto& o : a.specs) {
    if (o.save == kDontSave) continue;
    if (!o.takes_value) {
      if (a.flag(o.name)) out << "--" << o.name << '\n';
      continue;
    }
    auto it = a.value.find(o.name);
    const bool set = it != a.value.end();
    const std::string v = set ? it->second : o.def;
    if (o.save == kAlwaysSave || (set && v != o.def)) {
      if (!v.empty()) out << "--" << o.name << '=' << v << '\n';
    }
  }
}

}  // namespace strip_cli
