// gisspa_main <config>: the config-driven search executable (reference src/main.cu:14-98).
//
// Reads the config, the LST list of micrographs with CTF parameters, the Euler file and the
// template stack (MRC or EMAN HDF5), then searches micrographs [first, last) on one GPU and
// writes one record line per pick in the reference's format (src/norm.cu:610-615).  Differences from the reference:
// GPU_ID = -1 uses device 0 (one GPU per process; shard micrographs with first/last), errors
// are reported with a non-zero exit status instead of being printed and ignored, and records
// pass through NMS when nms_radius > 0.  With reference_records = 1 (norm_type=1) the records are
// the reference's own per-chunk records, in its order and without NMS.
//
// Checkpoint and resume (the reference has neither: its output is truncated on start,
// src/main.cu:34).  Beside the output, `<output>.done` lists every finished micrograph as
// "index bytes path", bytes being the output's size once that micrograph's records are flushed.
// A run with resume = 1 cuts the output back to the last listed size (dropping the lines of a
// micrograph that was being written when a run stopped), appends to it and skips the listed
// micrographs; with resume = 0 both files start empty.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "gisspa.h"
#include "inspect.hpp"
#include "io.hpp"

static int die(const std::string& msg) { return gsio::die("gisspa_main", msg); }

int main(int argc, char** argv) {
  if (argc < 2)
    return die("usage: gisspa_main <config> | --parse-dump <config> | --mrc-info <file.mrc> | "
               "--image-info <file> [index] | --stack-info <file> <count>");
  bool handled = false;
  const int rc = gsio::inspect(argc, argv, handled);
  if (handled) return rc;
  gsio::Config conf;
  if (!conf.parse(argv[1])) return die(conf.errors.empty() ? "bad config" : conf.errors[0]);
  auto missing = conf.missing_required();
  if (!missing.empty()) {
    std::string m = "missing required config keys:";
    for (auto& k : missing) m += " " + k;
    return die(m);
  }
  auto lst = gsio::load_lst(conf.gets("input"));
  if (lst.empty()) return die("empty or unreadable LST file " + conf.gets("input"));
  auto euler = gsio::load_euler(conf.gets("eulerfile"));
  if (euler.size() == 0) return die("empty or unreadable Euler file " + conf.gets("eulerfile"));
  // micrograph range, src/main.cu:25-26
  const int first = std::max(0, std::min(conf.geti("first"), (int)lst.size() - 1));
  const int last = std::min(conf.geti("last"), std::max(0, (int)lst.size()));
  int device = conf.geti("GPU_ID");
  if (device < 0) device = 0;

  // one template per Euler line (Templates(path, euler.size()), src/main.cu:32)
  gsio::Mrc tm;
  std::string err;
  const int n = (int)euler.size();
  if (!gsio::read_stack(conf.gets("template"), n, tm, err)) return die(err);
  std::vector<float> altaz(2 * n);
  for (int i = 0; i < n; ++i) {
    altaz[2 * i] = euler.alt[i];
    altaz[2 * i + 1] = euler.az[i];
  }

  gs_params p{};
  p.angpix = conf.getf("angpix");
  p.phistep = conf.getf("phistep");
  p.kk = conf.getf("kk");
  p.energy = conf.getf("energy");
  p.cs = conf.getf("cs");
  p.highres = conf.getf("Highres");
  p.lowres = conf.getf("Lowres");
  p.diameter = conf.getf("diameter");
  p.threshold = conf.getf("threshold");
  p.window_size = conf.geti("window_size");
  p.overlap = conf.geti("overlap");
  p.norm_type = conf.geti("norm_type");
  p.invert = conf.geti("invert");
  p.phase_flip = conf.geti("phase_flip");
  p.nms_radius = conf.getf("nms_radius");
  p.reference_records = conf.geti("reference_records");
  const int tile_local = conf.geti("tile_local_centers");

  gs_engine* eng = nullptr;
  if (gs_create(&p, device, &eng) != GS_OK) return die(gs_last_error());
  if (gs_load_templates(eng, tm.data.data(), n, tm.nx, tm.ny, altaz.data()) != GS_OK) {
    std::string e = gs_last_error();
    gs_destroy(eng);
    return die(e);
  }
  const std::string out_path = conf.gets("output"), done_path = out_path + ".done";
  std::set<int> done;
  if (conf.geti("resume")) {
    unsigned long long keep = 0;
    std::ifstream mf(done_path);
    std::string line;
    while (std::getline(mf, line)) {
      std::istringstream ls(line);
      int idx;
      unsigned long long bytes;
      if (ls >> idx >> bytes) {
        done.insert(idx);
        keep = bytes;
      }
    }
    std::error_code ec;
    if (std::filesystem::exists(out_path, ec) && std::filesystem::file_size(out_path, ec) > keep)
      std::filesystem::resize_file(out_path, keep, ec);
    if (ec) {
      gs_destroy(eng);
      return die("cannot cut " + out_path + " back to its last finished micrograph: " + ec.message());
    }
  }
  const auto mode = conf.geti("resume") ? std::ios::out | std::ios::app : std::ios::out | std::ios::trunc;
  std::ofstream out(out_path, mode);
  std::ofstream manifest(done_path, mode);
  if (!out || !manifest) {
    gs_destroy(eng);
    return die("cannot write output " + out_path + " or " + done_path);
  }
  int status = 0;
  long long lines = 0;
  for (int i = first; i < last; ++i) {
    const auto& e = lst[i];
    if (done.count(i)) {
      std::printf("%s: done in an earlier run, skipped\n", e.rpath.c_str());
      continue;
    }
    gsio::Mrc img;
    if (!gsio::read_image(e.rpath, e.unused, img, err)) {  // image number = LST index
      status = die(err);
      break;
    }
    gs_ctf ctf{e.defocus, e.dfdiff, e.dfang};
    gs_record* recs = nullptr;
    size_t nrec = 0;
    auto t0 = std::chrono::steady_clock::now();
    if (gs_search(eng, img.data.data(), img.nx, img.ny, &ctf, &recs, &nrec) != GS_OK) {
      status = die(e.rpath + ": " + gs_last_error());
      break;
    }
    char buf[2048];
    for (size_t k = 0; k < nrec; ++k) {
      gs_format_record(eng, &recs[k], e.unused, e.rpath.c_str(), &ctf, tile_local, buf, sizeof buf);
      out << buf;
      ++lines;
    }
    gs_free(recs);
    out.flush();
    if (!out) {
      status = die("cannot write output " + out_path);
      break;
    }
    // the size on disk (in append mode the stream's own position is not the file's size before
    // its first write)
    std::error_code ec;
    const unsigned long long size = std::filesystem::file_size(out_path, ec);
    if (ec) {
      status = die("cannot stat output " + out_path + ": " + ec.message());
      break;
    }
    manifest << i << " " << size << " " << e.rpath << "\n";
    manifest.flush();
    auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
    std::printf("%s: %zu records in %lld ms\n", e.rpath.c_str(), nrec, (long long)ms);
  }
  std::printf("output line count: %lld\n", lines);
  gs_destroy(eng);
  return status;
}
