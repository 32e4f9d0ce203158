// Host-side inputs of the search executable: config file, LST list, Euler file, MRC images.
#pragma once
#include <map>
#include <string>
#include <variant>
#include <vector>

namespace gsio {

// `key = value` config (reference Config, include/EMReader/DataReader2.h:17-71 and
// src/EMReader/DataReader2.cpp:6-53): keys with defaults, types from the defaults, unknown keys
// ignored.  Two keys are new: nms_radius and tile_local_centers.
struct Config {
  using Node = std::variant<int, float, std::string>;
  std::map<std::string, Node> value;
  std::vector<std::string> errors;
  Config();
  bool parse(const std::string& path);
  std::vector<std::string> missing_required() const;
  int geti(const std::string& k) const { return std::get<int>(value.at(k)); }
  float getf(const std::string& k) const { return std::get<float>(value.at(k)); }
  const std::string& gets(const std::string& k) const { return std::get<std::string>(value.at(k)); }
};

// LST line "index path defocus=.. dfdiff=.. dfang=.." (src/EMReader/DataReader2.cpp:80-101).
struct LstEntry {
  int unused = 0;
  std::string rpath;
  double defocus = 0, dfdiff = 0, dfang = 0;
};
std::vector<LstEntry> load_lst(const std::string& path);

// Euler file: three floats per line in one of the reference's layouts
// (src/EMReader/DataReader2.cpp:55-78).
struct Euler {
  std::vector<float> alt, az, phi;
  size_t size() const { return alt.size(); }
};
Euler load_euler(const std::string& path);

// MRC (MRC2014) image or stack; modes 0, 1, 2, 6.  Returns false on any error.
struct Mrc {
  int nx = 0, ny = 0, nz = 0;
  std::vector<float> data;
};
bool read_mrc(const std::string& path, Mrc& out, std::string& err);

}  // namespace gsio
