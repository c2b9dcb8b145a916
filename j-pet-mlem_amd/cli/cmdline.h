// Command-line parsing shared by the strip commands, with the reference
// tools' conventions (lib/cmdline/cmdline.h, src/util/cmdline_hooks.cpp,
// src/util/cmdline_types.h):
//   --name=value, --name value, -x value, -xvalue, grouped short flags;
//   integers with k/m/g suffixes; vectors as comma-separated lists;
//   -c/--config files with one argument per line ('#' lines are comments),
//   relative paths resolved against the including file's directory;
//   options read from a file that are unknown only warn, and options marked
//   not-from-file keep their command-line value.
#pragma once

#include <map>
#include <string>
#include <vector>

namespace strip_cli {

enum { kDontSave = 0, kSave = 1, kAlwaysSave = 2 };

struct OptSpec {
  const char* name;
  char short_name;
  bool takes_value;
  const char* def;  // default value ("" = none)
  bool from_file;   // may be set from a config file
  int save;         // written to a .cfg: kDontSave, kSave (when set or not the
                    // default), kAlwaysSave
  const char* help;
};

struct Args {
  std::vector<OptSpec> specs;
  std::map<std::string, std::string> value;
  std::map<std::string, int> count;  // flags and how often they were given
  std::vector<std::string> rest;
  std::vector<std::string> dir_stack;

  const OptSpec* find(const std::string& n) const;
  const OptSpec* find_short(char c) const;
  bool exist(const std::string& n) const { return value.count(n) || count.count(n); }
  std::string get(const std::string& n) const;
  int flag(const std::string& n) const {
    auto it = count.find(n);
    return it == count.end() ? 0 : it->second;
  }
};

long to_long(const std::string& s);
double to_double(const std::string& s);
std::vector<double> to_doubles(const std::string& s);

void parse(Args& a, const std::vector<std::string>& argv, bool from_file);
void load_config(Args& a, const std::string& path, bool verbose);
void usage(const Args& a, const char* synopsis, const char* note);

// n-z-pixels / n-y-pixels from n-pixels, or from length, radius and s-pixel
// (src/2d/strip/options.cpp:106-124).
void derive_pixel_counts(Args& a);

// The options as the reference's parser saves them (--name=value lines).
void write_config(const Args& a, const std::string& path);

}  // namespace strip_cli
