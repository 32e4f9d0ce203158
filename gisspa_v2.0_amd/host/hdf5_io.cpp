// EMAN HDF5 image stacks, read through the HDF5 C library loaded at run time (dlopen), so the
// executable has no link-time dependency on it: an HDF5 input fails with a clear message when
// no libhdf5 (1.10 or later) can be loaded.
//
// Layouts, as the reference's emdata::readImage tells them apart (src/EMReader/emdata.cpp:
// 192-220, src/EMReader/emhdf.cpp:45-115): a root attribute "num_dataset" marks the old EMAN1
// layout, whose image i is the root dataset "i" (emhdf.cpp:118-240; the reference takes
// nx = dims[0], ny = dims[1]); otherwise image i is the dataset "/MDF/images/i/image" of the
// EMAN2 layout (emhdf2.cpp:241-362), dims [ny][nx].  Data are read as native floats.
#include <dlfcn.h>

#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "io.hpp"

namespace gsio {

namespace {

using hid_t = int64_t;  // HDF5 >= 1.10
using herr_t = int;
using htri_t = int;
using hsize_t = unsigned long long;

struct H5Lib {
  bool ok = false;
  std::string why;
  herr_t (*open)() = nullptr;
  herr_t (*get_libversion)(unsigned*, unsigned*, unsigned*) = nullptr;
  herr_t (*eset_auto2)(hid_t, void*, void*) = nullptr;
  hid_t (*fopen)(const char*, unsigned, hid_t) = nullptr;
  herr_t (*fclose)(hid_t) = nullptr;
  htri_t (*aexists)(hid_t, const char*) = nullptr;
  htri_t (*lexists)(hid_t, const char*, hid_t) = nullptr;
  hid_t (*dopen2)(hid_t, const char*, hid_t) = nullptr;
  herr_t (*dclose)(hid_t) = nullptr;
  hid_t (*dget_space)(hid_t) = nullptr;
  int (*sget_ndims)(hid_t) = nullptr;
  int (*sget_dims)(hid_t, hsize_t*, hsize_t*) = nullptr;
  herr_t (*sclose)(hid_t) = nullptr;
  herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void*) = nullptr;
  const hid_t* native_float = nullptr;
};

template <typename F>
bool bind(void* lib, const char* name, F& f) {
  f = reinterpret_cast<F>(dlsym(lib, name));
  return f != nullptr;
}

H5Lib load_h5() {
  H5Lib h;
  std::vector<std::string> names;
  if (const char* env = std::getenv("GISSPA_HDF5_LIB")) names.push_back(env);
  for (const char* n : {"libhdf5.so", "libhdf5.so.103", "libhdf5.so.200", "libhdf5.so.310", "/opt/conda/lib/libhdf5.so"})
    names.push_back(n);
  void* lib = nullptr;
  for (const auto& n : names)
    if ((lib = dlopen(n.c_str(), RTLD_NOW | RTLD_LOCAL))) break;
  if (!lib) {
    h.why = "no HDF5 library found (set GISSPA_HDF5_LIB to libhdf5.so)";
    return h;
  }
  const bool all = bind(lib, "H5open", h.open) && bind(lib, "H5get_libversion", h.get_libversion) &&
                   bind(lib, "H5Eset_auto2", h.eset_auto2) && bind(lib, "H5Fopen", h.fopen) &&
                   bind(lib, "H5Fclose", h.fclose) && bind(lib, "H5Aexists", h.aexists) &&
                   bind(lib, "H5Lexists", h.lexists) && bind(lib, "H5Dopen2", h.dopen2) &&
                   bind(lib, "H5Dclose", h.dclose) && bind(lib, "H5Dget_space", h.dget_space) &&
                   bind(lib, "H5Sget_simple_extent_ndims", h.sget_ndims) &&
                   bind(lib, "H5Sget_simple_extent_dims", h.sget_dims) && bind(lib, "H5Sclose", h.sclose) &&
                   bind(lib, "H5Dread", h.dread) && bind(lib, "H5T_NATIVE_FLOAT_g", h.native_float);
  if (!all) {
    h.why = "the HDF5 library lacks a needed symbol";
    return h;
  }
  if (h.open() < 0) {
    h.why = "H5open failed";
    return h;
  }
  unsigned maj = 0, min = 0, rel = 0;
  h.get_libversion(&maj, &min, &rel);
  if (maj < 1 || (maj == 1 && min < 10)) {
    h.why = "HDF5 " + std::to_string(maj) + "." + std::to_string(min) + " is too old (1.10 or later)";
    return h;
  }
  h.eset_auto2(0, nullptr, nullptr);  // no error printing from the library; statuses are checked
  h.ok = true;
  return h;
}

const H5Lib& h5() {
  static const H5Lib h = load_h5();
  return h;
}

}  // namespace

bool is_hdf5(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  char x[8] = {};
  if (!f.read(x, 8)) return false;
  const unsigned char sig[8] = {137, 'H', 'D', 'F', '\r', '\n', 26, '\n'};
  for (int i = 0; i < 8; ++i)
    if ((unsigned char)x[i] != sig[i]) return false;
  return true;
}

bool read_hdf5_images(const std::string& path, int first, int count, Mrc& out, std::string& err) {
  const H5Lib& h = h5();
  if (!h.ok) {
    err = path + ": " + h.why;
    return false;
  }
  const hid_t f = h.fopen(path.c_str(), 0u /* read only */, 0 /* default properties */);
  if (f < 0) {
    err = "cannot open HDF5 file " + path;
    return false;
  }
  const bool eman1 = h.aexists(f, "num_dataset") > 0;
  bool ok = true;
  out.nx = out.ny = 0;
  out.nz = count;
  out.data.clear();
  for (int i = first; i < first + count && ok; ++i) {
    const std::string name = eman1 ? std::to_string(i) : "/MDF/images/" + std::to_string(i) + "/image";
    if (!eman1 && (h.lexists(f, "/MDF", 0) <= 0 || h.lexists(f, "/MDF/images", 0) <= 0 ||
                   h.lexists(f, ("/MDF/images/" + std::to_string(i)).c_str(), 0) <= 0 ||
                   h.lexists(f, name.c_str(), 0) <= 0)) {
      err = path + ": no image " + std::to_string(i);
      ok = false;
      break;
    }
    const hid_t d = h.dopen2(f, name.c_str(), 0);
    if (d < 0) {
      err = path + ": no image " + std::to_string(i);
      ok = false;
      break;
    }
    const hid_t sp = h.dget_space(d);
    hsize_t dims[3] = {0, 0, 0};
    const int nd = sp >= 0 ? h.sget_ndims(sp) : -1;
    if (nd == 2 || nd == 3) h.sget_dims(sp, dims, nullptr);
    if (nd == 2 || (nd == 3 && dims[0] == 1)) {  // one 2-D image (a single section)
      const hsize_t a = nd == 2 ? dims[0] : dims[1], b = nd == 2 ? dims[1] : dims[2];
      // EMAN2: [ny][nx]; EMAN1 as the reference reads it: nx = dims[0], ny = dims[1]
      const int nx = (int)(eman1 ? a : b), ny = (int)(eman1 ? b : a);
      if (i == first) {
        out.nx = nx;
        out.ny = ny;
        out.data.resize((size_t)nx * ny * count);
      }
      if (nx != out.nx || ny != out.ny || nx <= 0 || ny <= 0) {
        err = path + ": image " + std::to_string(i) + " has another size";
        ok = false;
      } else if (h.dread(d, *h.native_float, 0, 0, 0, out.data.data() + (size_t)(i - first) * nx * ny) < 0) {
        err = path + ": cannot read image " + std::to_string(i);
        ok = false;
      }
    } else {
      err = path + ": image " + std::to_string(i) + " is not 2-D";
      ok = false;
    }
    if (sp >= 0) h.sclose(sp);
    h.dclose(d);
  }
  h.fclose(f);
  return ok;
}

}  // namespace gsio
