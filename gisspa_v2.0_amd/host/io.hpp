// Host-side inputs of the search executable: config file, LST list, Euler file, MRC and EMAN
// HDF5 images.
#pragma once
#include <map>
#include <string>
#include <variant>
#include <vector>

namespace gsio {

// `key = value` config (reference Config, include/EMReader/DataReader2.h:17-71 and
// src/EMReader/DataReader2.cpp:6-53): keys with defaults, types from the defaults, unknown keys
// ignored.  Four keys are new: nms_radius, tile_local_centers, reference_records and resume.
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

// Image or stack as nz sections of nx x ny floats (x fastest).
struct Mrc {
  int nx = 0, ny = 0, nz = 0;
  std::vector<float> data;
};
// Every section of an MRC file (modes 0, 1, 2, 6).  Returns false on any error.
bool read_mrc(const std::string& path, Mrc& out, std::string& err);
// One image of an MRC stack or an EMAN HDF5 stack, told apart by the file signature as the
// reference's emdata::readImage(path, index) does (src/EMReader/emdata.cpp:192-227): for MRC,
// section index if 0 <= index < nz, else section 0 (the reference then reads the whole map and
// keeps its first nx x ny values); for HDF5, image index.  Micrographs are read at their LST
// index (src/utils/image.cpp:9).
bool read_image(const std::string& path, int index, Mrc& out, std::string& err);
// Images 0 .. count - 1 of a stack (Templates, src/utils/templates.cpp:11-31).
bool read_stack(const std::string& path, int count, Mrc& out, std::string& err);
// HDF5 (hdf5_io.cpp): the signature test and count images from first, EMAN1 or EMAN2 layout.
bool is_hdf5(const std::string& path);
bool read_hdf5_images(const std::string& path, int first, int count, Mrc& out, std::string& err);

}  // namespace gsio
