// No-GPU inspection modes shared by gisspa_main and the sanitizer-built gisspa_inspect.
#include "inspect.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>

#include "io.hpp"

namespace gsio {

int die(const char* prog, const std::string& msg) {
  std::fprintf(stderr, "%s: %s\n", prog, msg.c_str());
  return 1;
}

namespace {
std::string g_prog_name = "gisspa_main";
}

// Canonical dump of the parsed inputs (no GPU use): config keys, LST entries, Euler angles.
static int parse_dump(const char* path) {
  gsio::Config conf;
  if (!conf.parse(path)) return die(g_prog_name.c_str(), conf.errors.empty() ? "bad config" : conf.errors[0]);
  for (const auto& kv : conf.value) {
    if (std::holds_alternative<int>(kv.second))
      std::printf("config %s %d\n", kv.first.c_str(), std::get<int>(kv.second));
    else if (std::holds_alternative<float>(kv.second))
      std::printf("config %s %.6g\n", kv.first.c_str(), std::get<float>(kv.second));
    else
      std::printf("config %s %s\n", kv.first.c_str(), std::get<std::string>(kv.second).c_str());
  }
  for (const auto& e : gsio::load_lst(conf.gets("input")))
    std::printf("lst %d %s %.6f %.6f %.6f\n", e.unused, e.rpath.c_str(), e.defocus, e.dfdiff, e.dfang);
  auto eu = gsio::load_euler(conf.gets("eulerfile"));
  for (size_t i = 0; i < eu.size(); ++i) std::printf("euler %.6g %.6g %.6g\n", eu.alt[i], eu.az[i], eu.phi[i]);
  return 0;
}

// Dimensions and checksums of an MRC file as read (no GPU use).
static int mrc_info(const char* path) {
  gsio::Mrc m;
  std::string err;
  if (!gsio::read_mrc(path, m, err)) return die(g_prog_name.c_str(), err);
  double s = 0, s2 = 0;
  for (float v : m.data) {
    s += v;
    s2 += (double)v * v;
  }
  std::printf("mrc %d %d %d %.17g %.17g\n", m.nx, m.ny, m.nz, s, s2);
  return 0;
}

// Dimensions and checksums of one image as the search reads a micrograph (MRC section or HDF5
// image `index`; no GPU use).
static int image_info(const char* path, int index) {
  gsio::Mrc m;
  std::string err;
  if (!gsio::read_image(path, index, m, err)) return die(g_prog_name.c_str(), err);
  double s = 0, s2 = 0;
  for (float v : m.data) {
    s += v;
    s2 += (double)v * v;
  }
  std::printf("image %d %d %.17g %.17g\n", m.nx, m.ny, s, s2);
  return 0;
}

// The same for images 0 .. count - 1 of a stack as the search reads its templates.
static int stack_info(const char* path, int count) {
  gsio::Mrc m;
  std::string err;
  if (!gsio::read_stack(path, count, m, err)) return die(g_prog_name.c_str(), err);
  double s = 0, s2 = 0;
  for (float v : m.data) {
    s += v;
    s2 += (double)v * v;
  }
  std::printf("stack %d %d %d %.17g %.17g\n", m.nx, m.ny, m.nz, s, s2);
  return 0;
}

int inspect(int argc, char** argv, bool& handled) {
  handled = true;
  if (argc > 0 && argv[0]) {
    const std::string a0 = argv[0];
    g_prog_name = a0.substr(a0.find_last_of('/') == std::string::npos ? 0 : a0.find_last_of('/') + 1);
  }
  const std::string mode = argc > 1 ? argv[1] : "";
  if (mode == "--parse-dump") return argc > 2 ? parse_dump(argv[2]) : die("gisspa_main", "missing config");
  if (mode == "--mrc-info") return argc > 2 ? mrc_info(argv[2]) : die("gisspa_main", "missing file");
  if (mode == "--image-info")
    return argc > 2 ? image_info(argv[2], argc > 3 ? std::atoi(argv[3]) : 0) : die("gisspa_main", "missing file");
  if (mode == "--stack-info")
    return argc > 3 ? stack_info(argv[2], std::atoi(argv[3])) : die("gisspa_main", "missing file or count");
  handled = false;
  return 0;
}

}  // namespace gsio
