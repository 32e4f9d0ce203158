// Host-side inputs of the search executable (see io.hpp).
#include "io.hpp"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace gsio {

Config::Config() {
  using namespace std::string_literals;
  value = {
      {"input", ""s},       {"template", ""s},   {"template_dims", 2},  {"eulerfile", ""s},
      {"angpix", -1.0f},    {"phistep", -1.0f},  {"kk", -1.0f},         {"energy", -1.0f},
      {"cs", -1.0f},        {"Highres", -1.0f},  {"Lowres", 0.0f},      {"diameter", -1.0f},
      {"threshold", 7.0f},  {"output", ""s},     {"first", 0},          {"last", 0},
      {"GPU_ID", -1},       {"window_size", 320}, {"phase_flip", 0},    {"overlap", 0},
      {"norm_type", 0},     {"invert", 0},
      // new keys
      {"nms_radius", 0.0f}, {"tile_local_centers", 0}, {"reference_records", 0}, {"resume", 0},
  };
}

bool Config::parse(const std::string& path) {
  std::ifstream f(path);
  if (!f) {
    errors.push_back("cannot open config file " + path);
    return false;
  }
  std::string key, eq, val;
  while (f >> key >> eq >> val) {
    if (eq != "=") {
      errors.push_back("malformed line near key " + key);
      return false;
    }
    auto it = value.find(key);
    if (it == value.end()) continue;
    try {
      if (std::holds_alternative<int>(it->second))
        it->second = std::stoi(val);
      else if (std::holds_alternative<float>(it->second))
        it->second = std::stof(val);
      else
        it->second = val;
    } catch (const std::exception&) {
      errors.push_back("bad value for " + key + ": " + val);
      return false;
    }
  }
  return true;
}

std::vector<std::string> Config::missing_required() const {  // checkRequestPara
  std::vector<std::string> m;
  for (const char* k : {"input", "template", "eulerfile"})
    if (gets(k).empty()) m.push_back(k);
  for (const char* k : {"angpix", "phistep", "kk", "energy", "cs", "Highres", "Lowres", "diameter"})
    if (getf(k) < 0) m.push_back(k);
  return m;
}

std::vector<LstEntry> load_lst(const std::string& path) {
  std::vector<LstEntry> out;
  std::ifstream f(path);
  if (!f) return out;
  std::string line;
  LstEntry e;  // reused across lines like the reference, so a blank line repeats the last CTF
  while (std::getline(f, line)) {
    if (!line.empty() && line[0] == '#') continue;
    char buf[1024] = {0};
    std::sscanf(line.c_str(), "%d %1023s defocus=%lf dfdiff=%lf dfang=%lf", &e.unused, buf,
                &e.defocus, &e.dfdiff, &e.dfang);
    e.rpath = buf;
    out.push_back(e);
  }
  return out;
}

Euler load_euler(const std::string& path) {
  Euler e;
  std::ifstream f(path);
  if (!f) return e;
  std::string line;
  float alt, az, phi;
  while (std::getline(f, line)) {
    if (std::sscanf(line.c_str(), "%*f %f %f %f", &alt, &az, &phi) == 3 ||
        std::sscanf(line.c_str(), "%f %f %f", &alt, &az, &phi) == 3 ||
        std::sscanf(line.c_str(), "%*d %*s %f %f %f", &alt, &az, &phi) == 3) {
      e.alt.push_back(alt);
      e.az.push_back(az);
      e.phi.push_back(phi);
    }
  }
  return e;
}

// Sections [first, first + count) of an MRC file (count < 0: all of them).  Modes as the
// reference's readMRC (src/EMReader/emdata.cpp:412-677) converts them: 0 unsigned bytes
// v / 100 - 1.28 (emdata.cpp:540-547), 1 int16, 2 float32, and 6 uint16 (which the reference
// rejects); either byte order.
static bool read_mrc_sections(const std::string& path, int first, int count, Mrc& out, std::string& err) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    err = "cannot open " + path;
    return false;
  }
  int32_t hdr[256];
  if (!f.read(reinterpret_cast<char*>(hdr), 1024)) {
    err = "truncated MRC header: " + path;
    return false;
  }
  bool swap = false;
  auto bs = [](int32_t v) { return (int32_t)__builtin_bswap32((uint32_t)v); };
  if (hdr[3] < 0 || hdr[3] > 16) {  // mode out of range: try the other byte order
    swap = true;
    for (int i = 0; i < 256; ++i) hdr[i] = bs(hdr[i]);
  }
  const int nx = hdr[0], ny = hdr[1], nz = hdr[2];
  const int mode = hdr[3];
  const int ext = hdr[23];
  if (nx <= 0 || ny <= 0 || nz <= 0 || ext < 0) {
    err = "bad MRC dimensions: " + path;
    return false;
  }
  if (count < 0) first = 0, count = nz;
  if (first < 0 || first + count > nz) {
    err = path + " has " + std::to_string(nz) + " sections, " + std::to_string(first + count) + " needed";
    return false;
  }
  int bytes = mode == 0 ? 1 : (mode == 1 || mode == 6) ? 2 : mode == 2 ? 4 : 0;
  if (!bytes) {
    err = "unsupported MRC mode " + std::to_string(mode) + ": " + path;
    return false;
  }
  const size_t sec = (size_t)nx * ny, n = sec * count;
  // the sections must lie inside the file before anything is allocated for them (a header with
  // huge dimensions is a truncated or foreign file, not a request for terabytes)
  f.seekg(0, std::ios::end);
  const unsigned long long size = (unsigned long long)(std::streamoff)f.tellg();
  const unsigned long long head = 1024ull + (unsigned long long)ext;
  const unsigned long long per = (unsigned long long)sec * bytes;  // < 2^64: nx, ny < 2^31
  if (size < head || (unsigned long long)(first + count) > (size - head) / per) {
    err = "truncated MRC data: " + path + " (" + std::to_string(size) + " bytes for " +
          std::to_string(first + count) + " sections of " + std::to_string(per) + ")";
    return false;
  }
  f.seekg(1024 + (std::streamoff)ext + (std::streamoff)(sec * first * bytes));
  std::vector<char> raw(n * bytes);
  if (!f.read(raw.data(), raw.size())) {
    err = "truncated MRC data: " + path;
    return false;
  }
  out.nx = nx;
  out.ny = ny;
  out.nz = count;
  out.data.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const char* p = raw.data() + i * bytes;
    if (mode == 0) {
      out.data[i] = (float)((double)(uint8_t)p[0] / 100.0 - 1.28);
    } else if (mode == 1 || mode == 6) {
      uint16_t u;
      std::memcpy(&u, p, 2);
      if (swap) u = __builtin_bswap16(u);
      out.data[i] = mode == 1 ? (float)(int16_t)u : (float)u;
    } else {
      uint32_t u;
      std::memcpy(&u, p, 4);
      if (swap) u = __builtin_bswap32(u);
      float v;
      std::memcpy(&v, &u, 4);
      out.data[i] = v;
    }
  }
  return true;
}

bool read_mrc(const std::string& path, Mrc& out, std::string& err) { return read_mrc_sections(path, 0, -1, out, err); }

static int stack_size(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  int32_t hdr[4] = {0, 0, 0, 0};
  if (!f.read(reinterpret_cast<char*>(hdr), 16)) return 0;
  if (hdr[3] < 0 || hdr[3] > 16) hdr[2] = (int32_t)__builtin_bswap32((uint32_t)hdr[2]);
  return hdr[2];
}

bool read_image(const std::string& path, int index, Mrc& out, std::string& err) {
  if (is_hdf5(path)) return read_hdf5_images(path, index < 0 ? 0 : index, 1, out, err);
  const int nz = stack_size(path);
  return read_mrc_sections(path, (index >= 0 && index < nz) ? index : 0, 1, out, err);
}

bool read_stack(const std::string& path, int count, Mrc& out, std::string& err) {
  if (is_hdf5(path)) return read_hdf5_images(path, 0, count, out, err);
  return read_mrc_sections(path, 0, count, out, err);
}

}  // namespace gsio
