// Search engine: orchestration of the device pipeline and the C ABI of include/gisspa.h.
//
// One engine = one device + one HIP stream (default priority), as one SearchNorm/SearchNoNorm
// per host thread in the reference (src/norm.cu:48-123, src/nonorm.cu:50-122).  Device buffers
// grow on demand and persist across micrographs instead of being rebuilt per micrograph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <functional>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <future>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/gisspa.h"
#include "common.hpp"
#include "fft_generic.hpp"
#include "kernels.hpp"

namespace gs {

constexpr double kRefPi = 3.1415926535;  // include/utils/constants.h:3
constexpr int kTile = 720;               // include/utils/constants.h:5

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  T* ensure(size_t count) {
    if (count <= n && p) return p;
    release();
    GS_HIP(hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)));
    n = count;
    return p;
  }
};

// blocks_one_axis of src/norm.cu:68-81.
static std::vector<int> block_offsets(int length, int l, int overlap) {
  std::vector<int> o{0};
  int off = 0;
  while (off + l < length) {
    off += l - overlap;
    if (off + l >= length) off = length - l;
    o.push_back(off);
  }
  return o;
}

// The float angle loop of src/norm.cu:642 / src/nonorm.cu:582.
static std::vector<float> angle_list(float phistep) {
  std::vector<float> a;
  for (float e = 0.0f; e < 360.0f; e += phistep) a.push_back(e);
  return a;
}

// Exact radial bin, identical to the device rbin.
static int rbin_host(long long dx, long long dy) {
  long long n = dx * dx + dy * dy;
  long long k = (long long)std::sqrt((double)n);
  while (k * k > n) --k;
  while ((k + 1) * (k + 1) <= n) ++k;
  if (n > k * k + k) ++k;
  return (int)k - 1;
}

// Pixel counts per radial bin over the half plane x <= nx/2 (the rb of SQRSum_by_circle,
// k:115-125): purely geometric, so computed exactly on the host.
static std::vector<double> radial_counts(int nx, int ny) {
  const int nbins = std::max(nx, ny) / 2;
  std::vector<double> rb(nbins, 0.0);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x <= nx / 2; ++x) {
      const int r = rbin_host(std::min(x, nx - x), std::min(y, ny - y));
      if (r >= 0 && r < nbins) rb[r] += 1;
    }
  return rb;
}

// LoadImage host clipping (src/norm.cu:156-181): float sequential sums, int-truncated bounds.
static void clip_image_host(float* img, size_t N, int nx, int ny, int invert) {
#pragma clang fp contract(off)
  // The terms cur / nx / ny and cur^2 / nx / ny are formed a block at a time (independent
  // operations, which the compiler vectorises), then added in element order: the same float
  // operations in the same order as one loop, so the sums are bit-identical; the two sequential
  // add chains are what remains of the time.
  float sum = 0, sum_s2 = 0;
  constexpr size_t kB = 4096;
  float q1[kB], q2[kB];
  const float fx = (float)nx, fy = (float)ny;
  for (size_t i0 = 0; i0 < N; i0 += kB) {
    const size_t m = std::min(kB, N - i0);
    for (size_t i = 0; i < m; ++i) {
      const float cur = img[i0 + i];
      q1[i] = cur / fx / fy;
      q2[i] = cur * cur / fx / fy;
    }
    for (size_t i = 0; i < m; ++i) {
      sum += q1[i];
      sum_s2 += q2[i];
    }
  }
  // a non-finite pixel (or one whose square overflows float) makes the sums non-finite; the
  // reference then searches NaN planes, here it is an input error
  GS_REQUIRE(std::isfinite(sum) && std::isfinite(sum_s2), kInvalid,
             "image holds non-finite pixels (NaN or Inf) or values whose squares overflow float");
  const float avg = sum;
  const float var = std::sqrt(sum_s2 - avg * avg);
  // the reference's (int) casts, with the result x86 gives for a value outside int's range (the
  // "integer indefinite" INT_MIN) stated instead of left undefined; sqrt of a negative
  // difference is NaN, which also gives INT_MIN
  auto to_int = [](float v) { return (v >= -2147483648.f && v < 2147483648.f) ? (int)v : INT_MIN; };
  const int up = to_int(avg + 6 * var);
  const int low = to_int(avg - 6 * var);
  if (sum_s2 > avg * avg)
    for (size_t i = 0; i < N; ++i)
      if (img[i] > up || img[i] < low) img[i] = avg;
  if (invert)
    for (size_t i = 0; i < N; ++i) img[i] = -img[i];
}

// TileImages geometry (src/utils/tileimages.cpp:18-70), upper-left corners in x-major order.
static void tile_geometry(int W, int H, std::vector<int>& ulx, std::vector<int>& uly) {
  const int cols = (int)std::ceil((double)W / kTile), rows = (int)std::ceil((double)H / kTile);
  const int x_step = W / cols, y_step = H / rows;
  for (int j = 0; j < cols; ++j) {
    const int cx = j == 0 ? kTile / 2 : (j == cols - 1 ? W - kTile / 2 : x_step * (2 * j + 1) / 2);
    for (int i = 0; i < rows; ++i) {
      const int cy = i == 0 ? kTile / 2 : (i == rows - 1 ? H - kTile / 2 : y_step * (2 * i + 1) / 2);
      ulx.push_back(std::max(cx - kTile / 2, 0));
      uly.push_back(std::max(cy - kTile / 2, 0));
    }
  }
}

static bool record_before(const gs_record& a, const gs_record& b) {
  if (a.score != b.score) return a.score > b.score;
  if (a.y != b.y) return a.y < b.y;
  if (a.x != b.x) return a.x < b.x;
  if (a.template_id != b.template_id) return a.template_id < b.template_id;
  if (a.angle_index != b.angle_index) return a.angle_index < b.angle_index;
  if (a.block != b.block) return a.block < b.block;
  return a.pixel < b.pixel;
}

// Greedy non-maximum suppression (no reference implementation: the reference leaves it to
// external scripts, README.md:81): best first, drop anything strictly closer than radius.
static void nms(std::vector<gs_record>& v, float radius) {
  std::stable_sort(v.begin(), v.end(), record_before);
  if (radius <= 0) return;
  const double R2 = (double)radius * radius;
  std::vector<gs_record> kept;
  for (const auto& c : v) {
    bool ok = true;
    for (const auto& k : kept) {
      const double dx = (double)c.x - k.x, dy = (double)c.y - k.y;
      if (dx * dx + dy * dy < R2) {
        ok = false;
        break;
      }
    }
    if (ok) kept.push_back(c);
  }
  v.swap(kept);
}

// Event pairs around every launch of one kernel; read back once the run has synchronised, so
// timing adds no host synchronisation inside the pipeline.
struct KernelClock {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  size_t used = 0;
  double total_ms = 0;
  long long launches = 0;
  void reset() {
    total_ms = 0;
    launches = 0;
    used = 0;
  }
  void harvest() {
    for (size_t i = 0; i < used; ++i) {
      float ms = 0;
      GS_HIP(hipEventElapsedTime(&ms, ev[i].first, ev[i].second));
      total_ms += ms;
    }
    launches += (long long)used;
    used = 0;
  }
  void destroy() {
    for (auto& e : ev) {
      (void)hipEventDestroy(e.first);
      (void)hipEventDestroy(e.second);
    }
    ev.clear();
  }
};

// A captured angle loop and its executable instance.
struct CapturedGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

// Page-locked host staging buffer (hipHostMalloc), grown on demand.
struct PinnedBuf {
  float* p = nullptr;
  size_t n = 0;
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() { release(); }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
  float* ensure(size_t count) {
    if (count <= n && p) return p;
    release();
    GS_HIP(hipHostMalloc((void**)&p, sizeof(float) * std::max<size_t>(count, 1), hipHostMallocDefault));
    n = count;
    return p;
  }
};

class Engine {
 public:
  Engine(const gs_params& p, int device) : P_(p), device_(device) {
    validate();
    int count = 0;
    GS_HIP(hipGetDeviceCount(&count));
    GS_REQUIRE(device >= 0 && device < count, kInvalid, "device index out of range");
    GS_HIP(hipSetDevice(device_));
    GS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    GS_HIP(hipEventCreate(&ev_start_));
    GS_HIP(hipEventCreate(&ev_corr0_));
    GS_HIP(hipEventCreate(&ev_corr1_));
    GS_HIP(hipEventCreate(&ev_end_));
  }
  ~Engine() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graphs();
    for (hipEvent_t e : {ev_start_, ev_corr0_, ev_corr1_, ev_end_})
      if (e) (void)hipEventDestroy(e);
    clk_col_.destroy();
    clk_row_.destroy();
    if (stream2_) {
      (void)hipStreamSynchronize(stream2_);
      for (hipEvent_t e :
           {ev_zready_[0], ev_zready_[1], ev_zfree_[0], ev_zfree_[1], ev_sync_, ev_tprep_, ev_rfork_, ev_rjoin_})
        if (e) (void)hipEventDestroy(e);
      (void)hipStreamDestroy(stream2_);
    }
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  void load_templates(const float* data, int n, int w, int h, const float* alt_az) {
    GS_REQUIRE(data && n > 0 && w > 0 && h > 0, kInvalid, "empty template stack");
    const int l = P_.window_size;
    GS_REQUIRE(w <= l && h <= l, kInvalid, "templates larger than window_size");
    bind();
    n_ = n;
    w_ = w;
    h_ = h;
    white_ok_ = false;
    alt_.assign(n, 0.f);
    az_.assign(n, 0.f);
    if (alt_az)
      for (int i = 0; i < n; ++i) {
        alt_[i] = alt_az[2 * i];
        az_[i] = alt_az[2 * i + 1];
      }
    d_raw_.ensure((size_t)n * w * h);
    GS_HIP(hipMemcpyAsync(d_raw_.p, data, sizeof(float) * n * w * h, hipMemcpyHostToDevice, stream_));
    GS_HIP(hipStreamSynchronize(stream_));
  }

  void check_image(const float* img, int W, int H) const {
    GS_REQUIRE(img && W > 0 && H > 0, kInvalid, "empty image");
    const int l = P_.window_size;
    if (P_.norm_type) {
      GS_REQUIRE(W >= l && H >= l, kUnsupported, "norm_type=1 needs an image at least window_size wide and high");
    } else {
      GS_REQUIRE(W >= kTile && H >= kTile, kUnsupported, "norm_type=0 needs an image of at least 720 x 720");
    }
  }

  // Host part of LoadImage (src/norm.cu:156-181): copy into `dst` and, for norm_type=1, clip
  // and invert there.  Makes no HIP call, so it may run on a helper thread.
  void stage_image(const float* img, int W, int H, float* dst) const {
    std::memcpy(dst, img, sizeof(float) * (size_t)W * H);
    if (P_.norm_type) {
      clip_image_host(dst, (size_t)W * H, W, H, P_.invert);  // checks finiteness through its sums
    } else {
      bool finite = true;
      for (size_t i = 0, n = (size_t)W * H; i < n; ++i) finite &= std::isfinite(dst[i]);
      GS_REQUIRE(finite, kInvalid, "image holds non-finite pixels (NaN or Inf)");
    }
  }

  // Uploads a staged image (the output of stage_image).
  void upload_image(const float* staged, int W, int H, const gs_ctf& ctf) {
    bind();
    W_ = W;
    H_ = H;
    ctf_ = ctf;
    d_img_.ensure((size_t)W * H);
    GS_HIP(hipMemcpyAsync(d_img_.p, staged, sizeof(float) * W * H, hipMemcpyHostToDevice, stream_));
    GS_HIP(hipStreamSynchronize(stream_));
    have_image_ = true;
  }

  void set_image(const float* img, int W, int H, const gs_ctf& ctf) {
    check_image(img, W, H);
    bind();
    float* staged = pinned_[0].ensure((size_t)W * H);
    stage_image(img, W, H, staged);
    upload_image(staged, W, H, ctf);
  }

  // Searches a sequence of micrographs (each with its own CTF and optionally its own template
  // set) in order, as main's per-micrograph loop does (src/main.cu:37-49).  While micrograph i
  // runs on the device, a helper thread copies and clips micrograph i+1 into the other of two
  // page-locked staging buffers, and with graphs on the angle loop is replayed from a capture
  // whenever the problem shape repeats.  Records equal those of one gs_search per micrograph.
  // With refine_top > 0 (norm_type=1), the first refine_top records of each micrograph (its
  // best picks) are refined on that micrograph before the next one is searched; their results
  // go to `refined` (item order) and their counts to refined_counts.
  void search_batch(const gs_batch_item* items, int n_items, std::vector<gs_record>& all,
                    size_t* counts, int refine_top = 0, const gs_refine_opts* ropts = nullptr,
                    std::vector<gs_refined>* refined = nullptr, size_t* refined_counts = nullptr) {
    GS_REQUIRE(items && n_items > 0 && counts, kInvalid, "empty batch");
    GS_REQUIRE(refine_top <= 0 || (P_.norm_type == 1 && ropts && refined && refined_counts), kInvalid,
               "refinement in a batch needs norm_type=1, options and outputs");
    for (int i = 0; i < n_items; ++i) {
      check_image(items[i].image, items[i].width, items[i].height);
      if (items[i].templates) {
        GS_REQUIRE(items[i].n_templates > 0 && items[i].template_width > 0 && items[i].template_height > 0,
                   kInvalid, "batch item with an empty template stack");
        GS_REQUIRE(items[i].template_width <= P_.window_size && items[i].template_height <= P_.window_size,
                   kInvalid, "templates larger than window_size");
      } else {
        GS_REQUIRE(i > 0 || n_ > 0, kState, "first batch item has no templates and none are loaded");
      }
    }
    bind();
    size_t biggest = 0;
    for (int i = 0; i < n_items; ++i) biggest = std::max(biggest, (size_t)items[i].width * items[i].height);
    pinned_[0].ensure(biggest);
    pinned_[1].ensure(biggest);
    auto stage = [this, items](int i) {
      stage_image(items[i].image, items[i].width, items[i].height, pinned_[i % 2].p);
    };
    // Capturing the angle loop as a graph costs more than one eager run of it (about 0.14 s at
    // these sizes) and pays off only when the shape comes again: capture only shapes (micrograph
    // size, template count and size) that occur at least twice in the batch.
    std::vector<int> shape_count(n_items, 0);
    {
      std::map<std::array<long long, 4>, int> seen;
      std::vector<std::array<long long, 4>> keys(n_items);
      long long tn = n_, tw = w_, th = h_;
      for (int i = 0; i < n_items; ++i) {
        if (items[i].templates) tn = items[i].n_templates, tw = items[i].template_width, th = items[i].template_height;
        keys[i] = {(long long)items[i].width * 100000 + items[i].height, tn, tw, th};
        ++seen[keys[i]];
      }
      for (int i = 0; i < n_items; ++i) shape_count[i] = seen[keys[i]];
    }
    struct GraphGate {
      bool& flag;
      ~GraphGate() { flag = true; }
    } gate{graph_allowed_};
    std::future<void> next = std::async(std::launch::async, stage, 0);
    all.clear();
    try {
      for (int i = 0; i < n_items; ++i) {
        next.get();  // micrograph i is staged (rethrows a staging failure)
        const gs_batch_item& it = items[i];
        graph_allowed_ = shape_count[i] >= 2;
        if (it.templates)
          load_templates(it.templates, it.n_templates, it.template_width, it.template_height, it.alt_az);
        upload_image(pinned_[i % 2].p, it.width, it.height, it.ctf);
        // the other buffer was last read by the upload of micrograph i-1, which has completed
        if (i + 1 < n_items) next = std::async(std::launch::async, stage, i + 1);
        run();
        all.insert(all.end(), records_.begin(), records_.end());
        counts[i] = records_.size();
        if (refine_top > 0) {
          const size_t k = std::min<size_t>((size_t)refine_top, records_.size());
          refined_counts[i] = k;
          if (k) {
            std::vector<gs_peak> pk(k), outp(k);
            for (size_t q = 0; q < k; ++q) pk[q] = peak_from_record(records_[q]);
            std::vector<double> z(k), z0(k), g(4 * k);
            refine(pk.data(), k, ropts->iterations, ropts->step, ropts->scale, outp.data(), z.data(), z0.data(),
                   g.data());
            for (size_t q = 0; q < k; ++q) {
              gs_refined r{};
              r.peak = outp[q];
              r.score = z[q];
              r.score0 = z0[q];
              for (int c = 0; c < 4; ++c) r.grad[c] = g[4 * q + c];
              refined->push_back(r);
            }
          }
        }
      }
    } catch (...) {
      if (next.valid()) next.wait();
      throw;
    }
  }

  void run() {
    GS_REQUIRE(n_ > 0, kState, "no templates loaded");
    GS_REQUIRE(have_image_, kState, "no image set");
    Range whole("gs_search");
    bind();
    set_params();
    records_.clear();
    GS_HIP(hipEventRecord(ev_start_, stream_));
    if (P_.norm_type)
      run_norm();
    else
      run_nonorm();
    GS_HIP(hipEventRecord(ev_end_, stream_));
    GS_HIP(hipStreamSynchronize(stream_));
    float ms = 0;
    GS_HIP(hipEventElapsedTime(&ms, ev_start_, ev_end_));
    total_ms_ = ms;
    GS_HIP(hipEventElapsedTime(&ms, ev_corr0_, ev_corr1_));
    corr_ms_ = ms;
    clk_col_.harvest();
    clk_row_.harvest();
    Range rec("records");
    if (P_.norm_type && P_.reference_records) {
      collect_reference();  // the reference's own records, unsuppressed, in its order
    } else {
      if (P_.norm_type)
        collect_norm();
      else
        collect_nonorm();
      nms(records_, P_.nms_radius);
    }
    ran_ = true;
  }

  const std::vector<gs_record>& records() const { return records_; }

  size_t stage_bytes(int stage) const {
    const size_t l = P_.window_size, hx = l / 2 + 1;
    switch (stage) {
      case GS_STAGE_TEMPLATES:
        return P_.norm_type ? d_tprep_.n * sizeof(float) : (size_t)n_ * hx * l * sizeof(float2);
      case GS_STAGE_IMAGE: return P_.norm_type ? (size_t)W_ * H_ * sizeof(float) : 0;
      case GS_STAGE_BLOCKS: return P_.norm_type ? (size_t)nb_ * hx * l * sizeof(float2) : 0;
      case GS_STAGE_ROTATED: return P_.norm_type ? (size_t)n_ * hx * l * sizeof(float2) : 0;
      case GS_STAGE_MU:
      case GS_STAGE_SIGMA: return P_.norm_type ? (size_t)nb_ * l * l * sizeof(double) : 0;
      case GS_STAGE_BEST: return P_.norm_type ? (size_t)nb_ * l * l * sizeof(float) : 0;
      case GS_STAGE_ARG: return P_.norm_type ? (size_t)nb_ * l * l * sizeof(int32_t) : 0;
      case GS_STAGE_PLANES: return P_.norm_type ? 0 : planes_host_.size() * sizeof(double);
      default: return 0;
    }
  }

  void fetch(int stage, void* out, size_t bytes) {
    GS_REQUIRE(ran_, kState, "no completed run");
    const size_t need = stage_bytes(stage);
    GS_REQUIRE(need > 0 && bytes == need, kInvalid, "stage unavailable or size mismatch");
    bind();
    const void* src = nullptr;
    switch (stage) {
      case GS_STAGE_TEMPLATES: src = P_.norm_type ? (const void*)d_tprep_.p : (const void*)d_that_.p; break;
      case GS_STAGE_IMAGE: src = d_imgpre_.p; break;
      case GS_STAGE_BLOCKS: src = d_ihat_.p; break;
      case GS_STAGE_ROTATED: src = d_that_.p; break;
      case GS_STAGE_MU: src = d_mu_.p; break;
      case GS_STAGE_SIGMA: src = d_sigma_.p; break;
      case GS_STAGE_BEST: src = d_best_.p; break;
      case GS_STAGE_ARG: src = d_arg_.p; break;
      case GS_STAGE_PLANES: std::memcpy(out, planes_host_.data(), bytes); return;
    }
    GS_HIP(hipMemcpy(out, src, bytes, hipMemcpyDeviceToHost));
  }

  void set_limits(int na, int nt) {
    lim_angles_ = na;
    lim_templates_ = nt;
  }
  void set_profiling(bool on) {
    profile_ = on;
    clk_col_.reset();
    clk_row_.reset();
  }
  void timing(double* total, double* corr) const {
    if (total) *total = total_ms_;
    if (corr) *corr = corr_ms_;
  }
  void kernel_timing(int which, double* avg_ms, long long* count) const {
    const KernelClock& k = which == 0 ? clk_col_ : clk_row_;
    if (avg_ms) *avg_ms = k.launches ? k.total_ms / k.launches : 0.0;
    if (count) *count = k.launches;
  }
  void set_chunk(int tc) { chunk_ = tc; }
  void set_variant(int kernel, int variant) {
    GS_REQUIRE((kernel == 0 && variant >= 0 && variant <= 13) || (kernel == 1 && variant >= 0 && variant <= 9) ||
                   (kernel == 2 && (variant == 0 || variant == 1)),
               kInvalid, "unknown kernel variant");
    if (kernel == 0)
      col_variant_ = variant;
    else if (kernel == 1)
      row_variant_ = variant;
    else
      overlap_ = variant == 1;
  }

  void set_graph(bool on) {
    use_graph_ = on;
    if (!on) drop_graphs();
  }
  void graph_stats(long long* captures, long long* launches) const {
    if (captures) *captures = graph_captures_;
    if (launches) *launches = graph_launches_;
  }

  // Times each result-invariant option of the norm_type=1 correlation kernels on a one-angle
  // sample of the current micrograph and templates and keeps the fastest: the column-pass
  // variant, then the row-pass variant, then the templates per launch.  Overwrites the results
  // of the last run.
  void autotune(int* col_out, int* row_out, int* chunk_out) {
    GS_REQUIRE(P_.norm_type == 1, kUnsupported, "autotune covers the norm_type=1 correlation kernels");
    GS_REQUIRE(n_ > 0, kState, "no templates loaded");
    GS_REQUIRE(have_image_, kState, "no image set");
    if (!corr_supported(P_.window_size)) {  // one generic correlation: nothing to choose
      if (col_out) *col_out = col_variant_;
      if (row_out) *row_out = row_variant_;
      if (chunk_out) *chunk_out = chunk_;
      return;
    }
    const int la = lim_angles_, lt = lim_templates_;
    const bool prof = profile_, graph = use_graph_;
    lim_angles_ = 1;
    lim_templates_ = 0;
    use_graph_ = false;
    // kernel time of one run: 0 column pass only, 1 row pass only, 2 both
    auto measure = [&](int what) {
      profile_ = false;
      run();  // warm (allocations, caches)
      set_profiling(true);
      run();
      const double c = clk_col_.total_ms, r = clk_row_.total_ms;
      return what == 0 ? c : what == 1 ? r : c + r;
    };
    try {
      // candidates, as (column pass, row pass) pairs timed together: the row-packed layout (column
      // pass 12 packs the rows, the column group the fastest index of its order; row pass 8 reads
      // them and keeps the running maximum across launches), the unpacked layout with the compact recurrence-twiddle column pass and the
      // bank-conflict-free row pass (7, 6), and the natural layouts (3, 1); the other variants
      // measured slower on MI355X at every size tried.  A later candidate replaces the current
      // pick only when more than 3 % faster, so that timing noise between near-equal options
      // does not change the configuration from run to run.
      constexpr double kMargin = 0.97;
      double best = 0;
      bool first = true;
      int pick_col = col_variant_, pick_row = row_variant_;
      for (const auto& c : {std::pair<int, int>{12, 8}, std::pair<int, int>{7, 6}, std::pair<int, int>{3, 1}}) {
        col_variant_ = c.first;
        row_variant_ = c.second;
        const double ms = measure(2);
        if (first || ms < kMargin * best) best = ms, pick_col = c.first, pick_row = c.second;
        first = false;
      }
      col_variant_ = pick_col;
      row_variant_ = pick_row;
      int pick = 0;
      size_t free_b = 0, total_b = 0;
      GS_HIP(hipMemGetInfo(&free_b, &total_b));
      const int keep = chunk_;
      const int nb = nb_ > 0 ? nb_ : 1;
      const double plane = (double)P_.window_size * (P_.window_size / 2) * sizeof(float2);
      first = true;
      // largest first: more templates per launch cost less record traffic, and a smaller count
      // must be more than 3 % faster to replace a larger one
      for (int c : {kMaxChunk, 256, 128, 64}) {
        if (c > 64 && c / 2 >= n_) continue;  // a smaller count already takes every template at once
        // the inter-pass buffer of c templates must fit in half the free memory (plus what it holds now)
        if (plane * nb * c > 0.5 * (double)free_b + (double)d_Z_.n * sizeof(float2)) continue;
        chunk_ = c;
        const double ms = measure(2);
        if (first || ms < kMargin * best) best = ms, pick = c;
        first = false;
      }
      chunk_ = pick ? pick : keep;
    } catch (...) {
      lim_angles_ = la, lim_templates_ = lt, use_graph_ = graph;
      set_profiling(prof);
      throw;
    }
    lim_angles_ = la, lim_templates_ = lt, use_graph_ = graph;
    set_profiling(prof);
    if (col_out) *col_out = col_variant_;
    if (row_out) *row_out = row_variant_;
    if (chunk_out) *chunk_out = chunk_;
  }

  // Records of the last norm_type=1 run selected again at another threshold and NMS radius from
  // the per-pixel records kept on the device (update_CCG + the host scan, src/kernels.cu:709-721
  // and src/norm.cu:583-597, without a new search).  The run's own records are unchanged.
  std::vector<gs_record> select(float threshold, float nms_radius) {
    GS_REQUIRE(P_.norm_type == 1, kUnsupported, "select applies to norm_type=1 runs");
    GS_REQUIRE(ran_, kState, "no completed run");
    bind();
    const int max_cand = 1 << 20;
    GS_HIP(hipMemsetAsync(d_count_.p, 0, sizeof(int), stream_));
    launch_finalize_norm(d_S_.p, d_SS_.p, d_best_.p, d_arg_.p, nb_, P_.window_size, total_n_, threshold,
                         d_offx_.p, d_offy_.p, bx_, W_, H_, P_.diameter, d_mu_.p, d_sigma_.p, d_cand_.p,
                         d_count_.p, max_cand, stream_);
    GS_HIP(hipStreamSynchronize(stream_));
    std::vector<gs_record> keep;
    keep.swap(records_);
    try {
      collect_norm();
    } catch (...) {
      records_.swap(keep);
      throw;
    }
    std::vector<gs_record> out;
    out.swap(records_);
    records_.swap(keep);
    nms(out, nms_radius);
    return out;
  }

  // A norm_type=1 record as a pick to refine: its block, pixel and template, zero shift and
  // defocus offset, and theta = the angle-list value of its angle index (psi = 360 - theta).
  gs_peak peak_from_record(const gs_record& r) const {
    GS_REQUIRE(P_.norm_type == 1, kUnsupported, "picks are defined for norm_type=1 records");
    const auto angles = angle_list(P_.phistep);
    GS_REQUIRE(r.angle_index >= 0 && r.angle_index < (int)angles.size(), kInvalid, "angle index out of range");
    gs_peak p{};
    p.block = r.block;
    p.pixel = r.pixel;
    p.template_id = r.template_id;
    p.theta = angles[r.angle_index];
    return p;
  }

  // Adjoint refinement of k picks on the micrograph of the last norm_type=1 run: `iters` ascent
  // steps p += step * scale^2 * dz/dp per parameter, then a final evaluation.  iters = 0 only
  // evaluates z and its gradient.  mu and sigma of each pick's block pixel stay as the run
  // computed them.  Outputs (any may be null): refined picks, final z, starting z, final
  // gradient (k x 4: per px, px, deg, Angstrom).
  void refine(const gs_peak* in, size_t k, int iters, double step, const double* scale, gs_peak* out_pk,
              double* z, double* z0, double* grad) {
    Range range("refine");
    GS_REQUIRE(P_.norm_type == 1, kUnsupported, "refinement is defined for norm_type=1");
    GS_REQUIRE(ran_ && have_image_ && n_ > 0 && nb_ > 0, kState, "refinement needs a completed norm_type=1 run");
    GS_REQUIRE(in && k > 0, kInvalid, "no picks");
    GS_REQUIRE(iters >= 0 && iters <= 100000, kInvalid, "iterations out of range");
    GS_REQUIRE(std::isfinite(step) && scale && std::isfinite(scale[0]) && std::isfinite(scale[1]) &&
                   std::isfinite(scale[2]) && std::isfinite(scale[3]),
               kInvalid, "step and scales must be finite");
    const int l = P_.window_size;
    const size_t L2 = (size_t)l * l;
    for (size_t i = 0; i < k; ++i) {
      const gs_peak& p = in[i];
      GS_REQUIRE(p.block >= 0 && p.block < nb_ && p.pixel >= 0 && (size_t)p.pixel < L2 && p.template_id >= 0 &&
                     p.template_id < n_,
                 kInvalid, "pick block, pixel or template out of range");
      GS_REQUIRE(std::isfinite(p.dx) && std::isfinite(p.dy) && std::isfinite(p.theta) && std::isfinite(p.ddf),
                 kInvalid, "pick parameters must be finite");
    }
    bind();
    set_params();
    GS_HIP(hipEventRecord(ev_start_, stream_));
    if (!white_ok_) {
      whiten_templates(d_twhite_.ensure((size_t)n_ * L2), d_scratch_.ensure((size_t)n_ * L2));
      white_ok_ = true;
    }
    const double count = ref_region_count(l);
    // weight table over the band-pass support and its slot map; block spectra as [b][ky][kx]
    const int hx = l / 2 + 1;
    const std::vector<int32_t> idx = ref_band_indices(l, dp_.hr_t, dp_.lr_t);
    const int ntab = (int)idx.size();
    // rows ky of the weighted spectra that can be non-zero: min(ky, l - ky) <= band
    int band = 0;
    for (int32_t e : idx) band = std::max(band, std::min(e / l, l - e / l));
    int32_t* didx = d_ridx_.ensure(std::max(ntab, 1));
    RefWeight* tab = d_rtab_.ensure(std::max(ntab, 1));
    if (ntab) GS_HIP(hipMemcpyAsync(didx, idx.data(), sizeof(int32_t) * ntab, hipMemcpyHostToDevice, stream_));
    launch_ref_table(didx, ntab, l, dp_, tab, stream_);
    float2* ihat_t = d_rIT_.ensure((size_t)nb_ * hx * l);
    transpose_batched(d_ihat_.p, ihat_t, nb_, hx, l, l, stream_);
    // the rotation's mask table and the rows it leaves non-zero (the fused rotation kernels)
    int ylo = 0, yhi = l - 1;
    float* mtab = d_rmask_.ensure((size_t)L2);
    launch_ref_mask(l, P_.diameter, 4.0f, mtab, &ylo, &yhi, stream_);
    // Picks are independent but for the B planes, which hold the defocus fields of picks 2m and
    // 2m + 1: batches start at even indices, so a pick's partner (and its result) does not
    // depend on the batching.  Where the plane transforms are the stateless LDS kernels
    // (l = 320, 720), two batches run side by side on two streams, each with its own buffers, so
    // that one batch's transforms overlap the other's rotation and dot and each fills the
    // other's launch gaps and kernel tails.
    const int nl = corr_supported(l) && k >= 2 * kRefineMinLane ? 2 : 1;
    size_t per = (k + nl - 1) / nl;
    per += per & 1;
    const int Kc = (int)std::min<size_t>(per, kRefineChunk);
    hipStream_t ls[2] = {stream_, stream_};
    if (nl == 2) {
      ensure_pipeline();
      ls[1] = stream2_;
      GS_HIP(hipEventRecord(ev_rfork_, stream_));  // tables, mask and block spectra are ready
      GS_HIP(hipStreamWaitEvent(stream2_, ev_rfork_, 0));
    }
    for (int ln = 0; ln < nl; ++ln) {
      RefLane& R = rlane_[ln];
      // weighted spectra (written on the band-pass support only, zero elsewhere for the whole
      // call) and their inverse transforms
      launch_ref_clear(R.H.ensure((size_t)Kc * L2), (long long)Kc * L2, ls[ln]);
      R.T.ensure((size_t)Kc * L2);
      // K planes A (r + i dr/dtheta) and (K + 1) / 2 planes holding the B fields of two picks each
      R.AB.ensure((size_t)(Kc + (Kc + 1) / 2) * L2);
      R.pk.ensure(Kc);
      R.part.ensure((size_t)Kc * 16);
      R.part6.ensure((size_t)Kc * rotation_groups(l) * 6);
      R.stats.ensure((size_t)Kc * 8);
      R.cg.ensure((size_t)Kc * std::max(kRefDotSplit, ref_cols_dot_parts(l)) * 5);
      R.ph.ensure((size_t)Kc * (2 * l + 2 * hx));
      R.z.ensure(Kc);
      R.z0.ensure(Kc);
      R.grad.ensure((size_t)Kc * 4);
      R.hp.resize(Kc);
      R.hz.resize(Kc);
      R.hz0.resize(Kc);
      R.hg.resize((size_t)Kc * 4);
    }
    // scratch of the runtime-planned transform, needed only where no LDS plane transform exists
    float2* S = corr_supported(l) ? nullptr : d_rS_.ensure((size_t)(Kc + (Kc + 1) / 2) * L2);
    // one batch of K picks from c0 on lane ln: upload, iters + 1 evaluations, download queued
    auto enqueue = [&](int ln, size_t c0, int K) {
      RefLane& R = rlane_[ln];
      const hipStream_t st = ls[ln];
      RefPeak* rp = R.pk.p;
      float2 *H = R.H.p, *Tt = R.T.p, *AB = R.AB.p;
      for (int j = 0; j < K; ++j) {
        const gs_peak& p = in[c0 + j];
        R.hp[j] = RefPeak{p.block, p.pixel, p.template_id, 0, p.dx, p.dy, p.theta, p.ddf};
      }
      GS_HIP(hipMemcpyAsync(rp, R.hp.data(), sizeof(RefPeak) * K, hipMemcpyHostToDevice, st));
      // LDS plane transforms where built (l = 320, 720), the runtime-planned FFT otherwise
      auto fft2 = [&](float2* d, int batch, int sign, float scale) {
        if (!launch_plane_fft2(d, batch, l, sign, scale, st)) fft2_batched(fft_, d, S, batch, l, l, sign, scale, st);
      };
      for (int it = 0; it <= iters; ++it) {
        launch_ref_template(d_twhite_.p, tab, ntab, rp, K, l, dp_, R.part.p, count, H, st);
        // the rotation's bilinear taps read only columns (and rows) [ylo - 1, yhi + 1] of the
        // processed templates: rotation keeps the distance to the centre, which the mask bounds
        if (!launch_plane_fft2_from(H, Tt, K, l, +1, 1.f, st, band, ylo - 2, yhi + 2)) {
          GS_HIP(hipMemcpyAsync(Tt, H, sizeof(float2) * (size_t)K * L2, hipMemcpyDeviceToDevice, st));
          fft2_batched(fft_, Tt, S, K, l, l, +1, 1.f, st);
        }
        // l = 320, 720: the rotation writes rows already transformed and the column pass of the
        // forward transform is fused with the dot, so the spectra of AB are never stored
        int parts = kRefDotSplit;
        if (launch_ref_rotate_rows(Tt, rp, K, l, mtab, ylo, yhi, R.part6.p, R.stats.p, AB, st)) {
          GS_REQUIRE(launch_ref_cols_dot(d_ihat_.p, AB, rp, R.stats.p, K, l, ylo, yhi, (float)(1.0 / (double)L2),
                                         R.ph.p, R.cg.p, st),
                     kUnsupported, "window_size not supported");
          parts = ref_cols_dot_parts(l);
        } else {
          launch_ref_rotate(Tt, rp, K, l, P_.diameter, 4.0f, R.part6.p, R.stats.p, AB, st);
          fft2(AB, K + (K + 1) / 2, -1, (float)(1.0 / (double)L2));
          GS_REQUIRE(launch_ref_dot(ihat_t, AB, rp, R.stats.p, K, l, R.cg.p, st), kUnsupported,
                     "window_size not supported");
        }
        launch_ref_update(rp, K, R.cg.p, parts, d_mu_.p, d_sigma_.p, l, step, scale, it == 0, it < iters, R.z.p,
                          R.z0.p, R.grad.p, st);
      }
    };
    // results of the batch of K picks from c0 on lane ln (waits for the lane)
    auto collect = [&](int ln, size_t c0, int K) {
      RefLane& R = rlane_[ln];
      const hipStream_t st = ls[ln];
      GS_HIP(hipMemcpyAsync(R.hp.data(), R.pk.p, sizeof(RefPeak) * K, hipMemcpyDeviceToHost, st));
      GS_HIP(hipMemcpyAsync(R.hz.data(), R.z.p, sizeof(double) * K, hipMemcpyDeviceToHost, st));
      GS_HIP(hipMemcpyAsync(R.hz0.data(), R.z0.p, sizeof(double) * K, hipMemcpyDeviceToHost, st));
      GS_HIP(hipMemcpyAsync(R.hg.data(), R.grad.p, sizeof(double) * 4 * K, hipMemcpyDeviceToHost, st));
      GS_HIP(hipStreamSynchronize(st));
      for (int j = 0; j < K; ++j) {
        const size_t i = c0 + j;
        if (out_pk) {
          out_pk[i] = in[i];
          out_pk[i].dx = R.hp[j].dx;
          out_pk[i].dy = R.hp[j].dy;
          out_pk[i].theta = R.hp[j].theta;
          out_pk[i].ddf = R.hp[j].ddf;
        }
        if (z) z[i] = R.hz[j];
        if (z0) z0[i] = R.hz0[j];
        if (grad)
          for (int q = 0; q < 4; ++q) grad[4 * i + q] = R.hg[4 * j + q];
      }
    };
    for (size_t c0 = 0; c0 < k; c0 += (size_t)nl * Kc) {
      int Kl[2] = {0, 0};
      for (int ln = 0; ln < nl && c0 + (size_t)ln * Kc < k; ++ln) {
        Kl[ln] = (int)std::min<size_t>(Kc, k - (c0 + (size_t)ln * Kc));
        enqueue(ln, c0 + (size_t)ln * Kc, Kl[ln]);
      }
      for (int ln = 0; ln < nl; ++ln)
        if (Kl[ln]) collect(ln, c0 + (size_t)ln * Kc, Kl[ln]);
    }
    if (nl == 2) {  // the end event on stream_ covers the second lane's work
      GS_HIP(hipEventRecord(ev_rjoin_, stream2_));
      GS_HIP(hipStreamWaitEvent(stream_, ev_rjoin_, 0));
    }
    GS_HIP(hipEventRecord(ev_end_, stream_));
    GS_HIP(hipEventSynchronize(ev_end_));
    float ms = 0;
    GS_HIP(hipEventElapsedTime(&ms, ev_start_, ev_end_));
    refine_ms_ = ms;
  }
  double refine_ms() const { return refine_ms_; }

  const gs_params& params() const { return P_; }
  float alt(int j) const { return (j >= 0 && j < (int)alt_.size()) ? alt_[j] : 0.f; }
  float az(int j) const { return (j >= 0 && j < (int)az_.size()) ? az_[j] : 0.f; }

 private:
  void validate() {
    GS_REQUIRE(P_.window_size > 0 && P_.window_size % 2 == 0, kInvalid, "window_size must be even and positive");
    GS_REQUIRE(P_.phistep > 0, kInvalid, "phistep must be positive");
    GS_REQUIRE(P_.angpix > 0, kInvalid, "angpix must be positive");
    GS_REQUIRE(P_.overlap >= 0 && P_.overlap < P_.window_size, kInvalid, "overlap out of range");
    GS_REQUIRE(P_.norm_type || P_.window_size <= kTile, kInvalid, "window_size must be <= 720 for norm_type=0");
    GS_REQUIRE(P_.reference_records == 0 || P_.reference_records == 1, kInvalid, "reference_records must be 0 or 1");
    GS_REQUIRE(!P_.reference_records || P_.norm_type == 1, kUnsupported,
               "reference_records applies to norm_type=1 (norm_type=0 records are the reference's already)");
    GS_REQUIRE(!P_.reference_records || ((long long)P_.window_size * P_.window_size) % 256 == 0, kUnsupported,
               "reference_records needs window_size * window_size divisible by 256 (the reference's chunks)");
  }

  void bind() { GS_HIP(hipSetDevice(device_)); }

  void drop_graphs() {
    for (auto& g : graphs_) {
      if (g.second.exec) (void)hipGraphExecDestroy(g.second.exec);
      if (g.second.graph) (void)hipGraphDestroy(g.second.graph);
    }
    graphs_.clear();
  }

  void set_params() {  // SetParams, src/norm.cu:183-192
    DevParams& d = dp_;
    d.apix = P_.angpix;
    d.kk = P_.kk;
    d.highres = P_.highres;
    d.lowres = P_.lowres;
    d.d_m = P_.diameter;
    d.thresh = P_.threshold;
    const double def = ctf_.defocus;
    d.dfdiff = ctf_.dfdiff;
    d.dfang = ctf_.dfang;
    d.dfu = -def + ctf_.dfdiff;
    d.dfv = -def - ctf_.dfdiff;
    const double E = P_.energy;
    d.lambda = 12.2639 / std::sqrt(E * 1000.0 + 0.97845 * E * E);
    d.cs = P_.cs;
    d.ampconst = (double)0.07f;
    d.ehw = 4.0f;
    d.a = -10.81f;
    d.b = 1.f;
    d.b2 = 0.32f;
    d.bf = -18.17f;
    d.bf2 = -15.22f;
    d.bf3 = 1.72f;
    const int l = P_.window_size;
    d.hr_t = (float)l * P_.angpix / P_.highres;
    d.lr_t = (float)l * P_.angpix / P_.lowres;
    d.ds = 1.0f / (P_.angpix * (float)l);
  }

  void set_image_band(int nx, int ny) {
    const int l = std::max(nx, ny);
    dp_.hr_i = (float)l * P_.angpix / P_.highres;
    dp_.lr_i = (float)l * P_.angpix / P_.lowres;
    dp_.hi_thr = (float)l * P_.angpix / 6;
  }

  const double* radial_counts_dev(int nx, int ny) {
    const long long key = (long long)nx * 100000 + ny;
    auto it = rb_cache_.find(key);
    if (it != rb_cache_.end()) return it->second->p;
    auto rb = radial_counts(nx, ny);
    auto buf = std::make_unique<DevBuf<double>>();
    buf->ensure(rb.size());
    GS_HIP(hipMemcpy(buf->p, rb.data(), sizeof(double) * rb.size(), hipMemcpyHostToDevice));
    const double* p = buf->p;
    rb_cache_.emplace(key, std::move(buf));
    return p;
  }

  // Whole-frame (or tile) preprocessing: src/norm.cu:356-402 / src/nonorm.cu:346-389.
  // Whole-frame preparation of nf complex frames of nx x ny in place (src/norm.cu:356-402 with
  // dc_zero, src/nonorm.cu:346-389 per tile): forward transform, phase flip, radial power,
  // whitening and filter, inverse transform.  dc_zero needs nf = 1.
  void prep_frames(float2* F, int nf, int nx, int ny, bool dc_zero) {
    const size_t N = (size_t)nx * ny;
    GS_REQUIRE(!dc_zero || nf == 1, kInvalid, "dc_zero applies to one frame");
    // square frames of a built size (the 720 x 720 tiles of norm_type=0) take the LDS transform
    auto fft2 = [&](int sign, float scale) {
      if (nx != ny || !launch_plane_fft2(F, nf, nx, sign, scale, stream_))
        fft2_batched(fft_, F, d_frame2_.ensure(N * nf), nf, ny, nx, sign, scale, stream_);
    };
    fft2(-1, (float)(1.0 / (double)N));
    set_image_band(nx, ny);
    launch_phase_flip(F, nx, ny, nf, dp_, stream_);
    const int nbins = std::max(nx, ny) / 2;
    double* ra = d_ra_img_.ensure((size_t)nf * nbins);
    double* part = d_ichpart_.ensure((size_t)nf * radial_image_groups(nf) * nbins);
    launch_radial_power_image(F, nx, ny, nf, part, ra, stream_);
    launch_whiten_filter_image(F, nx, ny, nf, ra, radial_counts_dev(nx, ny), dp_, stream_);
    if (dc_zero) launch_zero_dc(F, stream_);
    fft2(+1, 1.f);
  }

  void prep_frame(const float* in, int nx, int ny, bool dc_zero, float* out) {
    const size_t N = (size_t)nx * ny;
    float2* F = d_frame_.ensure(N);
    launch_real_to_complex(in, F, (long long)N, 1.f, stream_);
    prep_frames(F, 1, nx, ny, dc_zero);
    launch_extract_real(F, out, (long long)N, stream_);
  }

  // Template preprocessing into full complex planes in work_ (src/norm.cu:194-354 or
  // src/nonorm.cu:178-344).  norm: ends in real space; nonorm: ends in Fourier space.
  // Padded, edge-standardised, transformed and whitened templates (src/norm.cu:132-318 up to
  // the weighting) into T (n planes l x l); S is scratch of the same size.
  void whiten_templates(float2* T, float2* S) {
    const int l = P_.window_size, n = n_;
    const size_t L2 = (size_t)l * l;
    launch_pad_templates(d_raw_.p, n, w_, h_, l, T, stream_);
    double* mom = d_mom_.ensure((size_t)n * 4);
    double* part = d_chpart_.ensure((size_t)n * plane_reduction_chunks(n) * std::max(4, l / 2));
    launch_template_moments(T, n, l, part, mom, stream_);
    launch_template_standardise(T, n, l, mom, stream_);
    fft2_planes(T, S, n, -1, (float)(1.0 / (double)L2));
    double* ra = d_ra_.ensure((size_t)n * (l / 2));
    launch_radial_power_planes(T, n, l, part, ra, stream_);
    launch_whiten_planes(T, n, l, ra, radial_counts_dev(l, l), stream_);
  }

  void prep_templates() {
    const int l = P_.window_size, n = n_;
    const size_t L2 = (size_t)l * l;
    float2* T = d_work_.ensure((size_t)n * L2);
    float2* S = d_scratch_.ensure((size_t)n * L2);
    whiten_templates(T, S);
    if (!P_.norm_type) {  // src/nonorm.cu:291-299
      fft2_planes(T, S, n, +1, 1.f);
      launch_mask_planes(T, n, l, P_.diameter, 4.0f, stream_);
      fft2_planes(T, S, n, -1, (float)(1.0 / (double)L2));
    }
    launch_weight_planes(T, n, l, dp_, stream_);
    double* m = d_m_.ensure(n);
    launch_area_power_full(T, n, l, d_apart_.ensure((size_t)n * plane_reduction_chunks(n)), m, stream_);
    const float sign = ((l / 2) % 2) ? -1.f : 1.f;  // checkerboard sign for odd l/2 (k:669-677)
    launch_scale_planes(T, n, (long long)L2, m, P_.norm_type ? 1.f : sign, stream_);
    if (P_.norm_type) {
      fft2_planes(T, S, n, +1, 1.f);
      launch_extract_real(T, d_tprep_.ensure((size_t)n * L2), (long long)n * L2, stream_);
    } else {
      const int hx = l / 2 + 1;
      transpose_batched(T, d_that_.ensure((size_t)n * hx * l), n, l, l, hx, stream_);
    }
  }

  // Half spectra [p][kx 0..l/2][ky] scaled by `scale` of `np` complex l x l planes in `planes`
  // (the real-input forward transform of the reference's C2C + scale, src/norm.cu:516-519).
  void half_spectra(float2* planes, int np, float2* out, float scale) {
    const int l = P_.window_size, hx = l / 2 + 1;
    float2* S = d_scratch_.ensure((size_t)np * l * l);
    if (launch_plane_fft2(planes, np, l, -1, scale, stream_)) {  // LDS plane transform, then keep kx <= l/2
      transpose_batched(planes, out, np, l, l, hx, stream_);
      return;
    }
    fft_rows(fft_, planes, S, l, (long long)np * l, -1, 1.f, stream_);
    transpose_batched(planes, out, np, l, l, hx, stream_);
    fft_rows(fft_, out, S, l, (long long)np * hx, -1, scale, stream_);
  }

  // In-place 2-D transforms of np l x l planes: the LDS plane transform where built, else the
  // runtime-planned FFT with scratch S.
  void fft2_planes(float2* T, float2* S, int np, int sign, float scale) {
    const int l = P_.window_size;
    if (!launch_plane_fft2(T, np, l, sign, scale, stream_)) fft2_batched(fft_, T, S, np, l, l, sign, scale, stream_);
  }

  void clock_begin(KernelClock& k, hipStream_t s) {
    if (!profile_) return;
    if (k.used == k.ev.size()) {
      std::pair<hipEvent_t, hipEvent_t> p{nullptr, nullptr};
      GS_HIP(hipEventCreate(&p.first));
      GS_HIP(hipEventCreate(&p.second));
      k.ev.push_back(p);
    }
    GS_HIP(hipEventRecord(k.ev[k.used].first, s));
  }
  void clock_end(KernelClock& k, hipStream_t s) {
    if (!profile_) return;
    GS_HIP(hipEventRecord(k.ev[k.used].second, s));
    k.used += 1;
  }

  // Second stream (default priority) for the row pass (and the norm_type=0 template
  // preparation), and the events of the double-buffered inter-pass handoff; created on first use.
  void ensure_pipeline() {
    if (stream2_) return;
    GS_HIP(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking));
    for (hipEvent_t* e : {&ev_zready_[0], &ev_zready_[1], &ev_zfree_[0], &ev_zfree_[1], &ev_sync_, &ev_tprep_,
                          &ev_rfork_, &ev_rjoin_})
      GS_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }

  // Templates per launch when not set: as many as a memory budget holds, at most kMaxChunk.  The
  // row pass reads and writes the per-pixel records once per launch, so more templates per
  // launch means less record traffic per plane (on MI355X at l = 320 and 234 blocks, 64 -> 512
  // templates per launch takes about 8 % off the row pass); at 1000 the column pass slowed
  // down.  Budget: a quarter of the device memory or half of what was free when first asked,
  // whichever is smaller, halved for the double-buffered two-stream pipeline (512 templates x
  // 234 blocks at l = 320 is a 49 GB inter-pass buffer).
  // The generic correlation (window sizes without tiled kernels) needs two buffers of (l/2 + 1)
  // columns per plane instead of one of l/2, and never double-buffers.
  static constexpr int kMaxChunk = 512;
  int chunk_for(int nb) {
    if (chunk_ > 0) return chunk_;
    if (mem_budget_ <= 0) {
      size_t fr = 0, tot = 0;
      GS_HIP(hipMemGetInfo(&fr, &tot));
      mem_budget_ = std::min(0.25 * (double)tot, 0.5 * (double)fr);
    }
    const int l = P_.window_size;
    const double plane = corr_supported(l) ? (double)l * (l / 2) * sizeof(float2) * (overlap_ ? 2 : 1)
                                           : 2.0 * (double)l * (l / 2 + 1) * sizeof(float2);
    const double per_t = plane * std::max(nb, 1);
    return (int)std::max(4.0, std::min((double)kMaxChunk, std::floor(mem_budget_ / per_t)));
  }

  // norm_type = 1 (SearchNorm::work, src/norm.cu:620-660) with one correlation pass.
  void run_norm() {
    const int l = P_.window_size, hx = l / 2 + 1, n = n_;
    const size_t L2 = (size_t)l * l;
    {
      Range r("templates");
      prep_templates();
    }
    std::optional<Range> img_range(std::in_place, "image_and_blocks");
    float* pre = d_imgpre_.ensure((size_t)W_ * H_);
    prep_frame(d_img_.p, W_, H_, true, pre);
    // SplitImage (src/norm.cu:404-477)
    const auto ox = block_offsets(W_, l, P_.overlap), oy = block_offsets(H_, l, P_.overlap);
    bx_ = (int)ox.size();
    nb_ = (int)(ox.size() * oy.size());
    offx_ = ox;
    offy_ = oy;
    int* doffx = d_offx_.ensure(ox.size());
    int* doffy = d_offy_.ensure(oy.size());
    GS_HIP(hipMemcpyAsync(doffx, ox.data(), sizeof(int) * ox.size(), hipMemcpyHostToDevice, stream_));
    GS_HIP(hipMemcpyAsync(doffy, oy.data(), sizeof(int) * oy.size(), hipMemcpyHostToDevice, stream_));
    float2* blk = d_blocks_.ensure((size_t)nb_ * L2);
    launch_split_blocks(pre, W_, H_, l, doffx, doffy, bx_, nb_, blk, stream_);
    float2* ihat = d_ihat_.ensure((size_t)nb_ * hx * l);
    half_spectra(blk, nb_, ihat, (float)(1.0 / (double)L2));
    double* m = d_m_.ensure(std::max(n, nb_));
    launch_area_power_half(ihat, nb_, l, d_apart_.ensure((size_t)nb_ * plane_reduction_chunks(nb_)), m, stream_);
    // fold the 1/l^2 of add_CCG_to_sum (k:691) and the odd-l/2 checkerboard sign into Ihat
    const float sign = ((l / 2) % 2) ? -1.f : 1.f;
    launch_scale_planes(ihat, nb_, (long long)hx * l, m, (float)(sign / (double)L2), stream_);

    const long long npix = (long long)nb_ * L2;
    launch_init_accumulators(d_S_.ensure(npix), d_SS_.ensure(npix), d_best_.ensure(npix),
                             d_arg_.ensure(npix), npix, stream_);
    auto angles = angle_list(P_.phistep);
    int na = (int)angles.size();
    if (lim_angles_ > 0) na = std::min(na, lim_angles_);
    const int nt = lim_templates_ > 0 ? std::min(n, lim_templates_) : n;
    angles_ = angles;
    const int tc_max = chunk_for(nb_);
    // window sizes without tiled kernels: the generic correlation with its two buffers
    const bool gen = !corr_supported(l);
    const size_t zsz = gen ? 2 * corr_generic_buffer(l, tc_max, nb_) : (size_t)tc_max * nb_ * l * (l / 2);
    if (gen) corr_generic_prepare(fft_, l);
    // With overlap, the column pass of chunk c+1 (stream_) runs while the row pass of chunk c
    // (stream2_) consumes the other half of a double-buffered inter-pass buffer.
    const bool overlap = overlap_ && !gen;
    hipStream_t rs = stream_;
    if (overlap) {
      ensure_pipeline();
      rs = stream2_;
    }
    // column variants 8 to 13 write the row-packed layout that only row variants 7 and 8 read
    // (l = 320; other sizes use one unpacked layout)
    const bool packed_col = col_variant_ >= 8 && col_variant_ <= 13;
    GS_REQUIRE(l != 320 || packed_col == (row_variant_ == 7 || row_variant_ == 8), kInvalid,
               "column variants 8 to 13 (row-packed inter-pass layout) and row variants 7, 8 go together");
    float2* Zb = d_Z_.ensure(zsz * (overlap ? 2 : 1));
    double* part = d_part_.ensure((size_t)n * rotation_groups(l) * 2);
    float* stats = d_stats_.ensure((size_t)n * 2);
    float2* that = d_that_.ensure((size_t)n * hx * l);
    // template pairs share one complex plane per angle (real and imaginary parts)
    float2* work = d_work_.ensure((size_t)((n + 1) / 2) * L2);
    float2* tscr = d_scratch_.ensure((size_t)((n + 1) / 2) * L2);
    img_range.reset();
    std::optional<Range> loop_range(std::in_place, "angle_loop");
    GS_HIP(hipEventRecord(ev_corr0_, stream_));
    if (overlap) {  // accumulators initialised on stream_ before the first row pass
      GS_HIP(hipEventRecord(ev_sync_, stream_));
      GS_HIP(hipStreamWaitEvent(stream2_, ev_sync_, 0));
    }
    auto enqueue_angles = [&]() {
      long long chunk = 0;
      for (int a = 0; a < na; ++a) {
        // RotateTemplate (src/norm.cu:479-521)
        const float cose = (float)std::cos((double)angles[a] * kRefPi / 180);
        const float sine = (float)std::sin((double)angles[a] * kRefPi / 180);
        launch_rotate_mask_pairs(d_tprep_.p, nt, l, cose, sine, P_.diameter, 4.0f, work, part, stream_);
        launch_rotation_stats(part, nt, l, stats, stream_);
        // one forward transform per pair; the split, the standardisation (affine: applied to the
        // spectra) and the 1/l^2 scale follow in one pass that writes the half spectra
        if (!launch_plane_fft2_pairs(work, work, (nt + 1) / 2, l, nt, 1, stats, (float)(1.0 / (double)L2), that,
                                     stream_)) {
          fft2_planes(work, tscr, (nt + 1) / 2, -1, 1.f);
          launch_unpack_pair_spectra(work, nt, 1, l, stats, (float)(1.0 / (double)L2), that, stream_);
        }
        // correlation + per-pixel statistics (ComputeCCGSum / PickParticles in one pass)
        for (int t0 = 0; t0 < nt; t0 += tc_max, ++chunk) {
          const int tc = std::min(tc_max, nt - t0);
          const int zb = overlap ? (int)(chunk % 2) : 0;
          float2* Z = Zb + zb * zsz;
          if (overlap && chunk >= 2) GS_HIP(hipStreamWaitEvent(stream_, ev_zfree_[zb], 0));
          CorrArgs ca{that, ihat, Z, t0, tc, nb_};
          if (gen) {
            clock_begin(clk_col_, stream_);
            launch_corr_generic_pixel(fft_, ca, l, Z, Z + zsz / 2, d_S_.p, d_SS_.p, d_best_.p, d_arg_.p, a * n + t0,
                                      stream_);
            clock_end(clk_col_, stream_);
            continue;
          }
          clock_begin(clk_col_, stream_);
          GS_REQUIRE(launch_colpass(l, ca, col_variant_, stream_), kUnsupported, "window_size not supported");
          clock_end(clk_col_, stream_);
          if (overlap) {
            GS_HIP(hipEventRecord(ev_zready_[zb], stream_));
            GS_HIP(hipStreamWaitEvent(stream2_, ev_zready_[zb], 0));
          }
          clock_begin(clk_row_, rs);
          launch_rowpass_pixel(l, Z, tc, nb_, d_S_.p, d_SS_.p, d_best_.p, d_arg_.p, a * n + t0,
                               row_variant_, rs);
          clock_end(clk_row_, rs);
          if (overlap) GS_HIP(hipEventRecord(ev_zfree_[zb], stream2_));
        }
      }
    };
    if (use_graph_ && graph_allowed_ && !overlap && !profile_ && !gen) {
      // The angle loop has no micrograph-dependent arguments: capture it once per problem shape
      // and buffer set, then replay it for every later micrograph of that shape.
      auto ptr = [](const void* p) { return (long long)(uintptr_t)p; };
      const std::vector<long long> key{
          nb_, n, nt, na, tc_max, l, col_variant_, row_variant_, ptr(d_tprep_.p), ptr(tscr),
          ptr(part), ptr(stats), ptr(work), ptr(that), ptr(ihat), ptr(Zb),
          ptr(d_S_.p), ptr(d_SS_.p), ptr(d_best_.p), ptr(d_arg_.p)};
      auto it = graphs_.find(key);
      if (it == graphs_.end()) {
        // nothing may allocate while the stream is captured: the scratch buffer of the template
        // transforms was sized above, the transform plan for l exists (template preparation made
        // it), and the correlation twiddle tables are built here
        (void)fft_.plan(l);
        corr_prepare();
        if (graphs_.size() >= kMaxGraphs) drop_graphs();
        hipGraph_t g = nullptr;
        GS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
        try {
          enqueue_angles();
        } catch (...) {
          (void)hipStreamEndCapture(stream_, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        GS_HIP(hipStreamEndCapture(stream_, &g));
        hipGraphExec_t ex = nullptr;
        GS_HIP(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
        it = graphs_.emplace(key, CapturedGraph{g, ex}).first;
        ++graph_captures_;
      }
      GS_HIP(hipGraphLaunch(it->second.exec, stream_));
      ++graph_launches_;
    } else {
      enqueue_angles();
    }
    if (overlap) {  // finalisation reads the accumulators written on stream2_
      GS_HIP(hipEventRecord(ev_sync_, stream2_));
      GS_HIP(hipStreamWaitEvent(stream_, ev_sync_, 0));
    }
    GS_HIP(hipEventRecord(ev_corr1_, stream_));
    loop_range.reset();
    Range fin_range("finalize");
    // set_CCG_mean: n = N_t * int(360/phistep) (src/norm.cu:549-550)
    double total_n = (double)n * (int)(360 / P_.phistep);
    if (lim_angles_ > 0 || lim_templates_ > 0) total_n = (double)nt * na;
    total_n_ = total_n;
    const int max_cand = 1 << 20;
    Cand* cand = d_cand_.ensure(max_cand);
    int* count = d_count_.ensure(1);
    GS_HIP(hipMemsetAsync(count, 0, sizeof(int), stream_));
    launch_finalize_norm(d_S_.p, d_SS_.p, d_best_.p, d_arg_.p, nb_, l, total_n, P_.threshold, doffx,
                         doffy, bx_, W_, H_, P_.diameter, d_mu_.ensure(npix), d_sigma_.ensure(npix),
                         cand, count, max_cand, stream_);
    ref_parts_ = false;
    if (!P_.reference_records) return;
    // Second pass of the reference (PickParticles, src/norm.cu:560-600): every correlation again,
    // now with mu and sigma known, reduced to the reference's per-256-chunk records.  At l = 320
    // it runs on the tiled correlation (a row-packed column pass, then the row pass's record
    // form, which appends chunk parts that collect_reference merges); at other sizes on the
    // generic correlation, whose two buffers take the inter-pass buffer's memory.
    RefRec* rec = d_refrec_.ensure(kMaxRefRecords);
    int* rcount = d_refcount_.ensure(1);
    GS_HIP(hipMemsetAsync(rcount, 0, sizeof(int), stream_));
    ref_parts_ = !gen && l == 320;
    const int colv = col_variant_ >= 8 && col_variant_ <= 13 ? col_variant_ : 12;  // a row-packed layout
    int tc2 = tc_max;
    float2 *P2 = nullptr, *W2 = nullptr;
    if (!ref_parts_) {
      corr_generic_prepare(fft_, l);
      const size_t per_t = 2 * corr_generic_buffer(l, 1, nb_);
      tc2 = (int)std::max<size_t>(1, std::min<size_t>((size_t)nt, zsz * (overlap ? 2 : 1) / per_t));
      P2 = d_Z_.ensure(per_t * tc2);
      W2 = P2 + per_t / 2 * tc2;
    }
    for (int a = 0; a < na; ++a) {
      const float cose = (float)std::cos((double)angles[a] * kRefPi / 180);
      const float sine = (float)std::sin((double)angles[a] * kRefPi / 180);
      launch_rotate_mask_pairs(d_tprep_.p, nt, l, cose, sine, P_.diameter, 4.0f, work, part, stream_);
      launch_rotation_stats(part, nt, l, stats, stream_);
      if (!launch_plane_fft2_pairs(work, work, (nt + 1) / 2, l, nt, 1, stats, (float)(1.0 / (double)L2), that,
                                   stream_)) {
        fft2_planes(work, tscr, (nt + 1) / 2, -1, 1.f);
        launch_unpack_pair_spectra(work, nt, 1, l, stats, (float)(1.0 / (double)L2), that, stream_);
      }
      for (int t0 = 0; t0 < nt; t0 += tc2) {
        const int tc = std::min(tc2, nt - t0);
        if (ref_parts_) {
          const CorrArgs ca{that, ihat, Zb, t0, tc, nb_};
          GS_REQUIRE(launch_colpass(l, ca, colv, stream_), kUnsupported, "window_size not supported");
          GS_REQUIRE(launch_rowpass_chunk_records(l, Zb, tc, nb_, d_mu_.p, d_sigma_.p, P_.threshold, a, t0, rec,
                                                  rcount, kMaxRefRecords, stream_),
                     kUnsupported, "window_size not supported");
          continue;
        }
        const CorrArgs ca{that, ihat, P2, t0, tc, nb_};
        launch_corr_generic_records(fft_, ca, l, P2, W2, d_mu_.p, d_sigma_.p, a, P_.threshold, doffx, doffy, bx_,
                                    W_, H_, P_.diameter, rec, rcount, kMaxRefRecords, stream_);
      }
    }
  }

  // The reference records of the last run, in the reference's output order: angle, block
  // (block_id = i + j * block_x), template, chunk (src/norm.cu:560-600).  No NMS: the reference
  // prints every record.
  void collect_reference() {
    int count = 0;
    GS_HIP(hipMemcpy(&count, d_refcount_.p, sizeof(int), hipMemcpyDeviceToHost));
    GS_REQUIRE(count <= kMaxRefRecords, kUnsupported,
               "more than " + std::to_string(kMaxRefRecords) + " reference records above threshold");
    std::vector<RefRec> c(count);
    if (count) GS_HIP(hipMemcpy(c.data(), d_refrec_.p, sizeof(RefRec) * count, hipMemcpyDeviceToHost));
    std::sort(c.begin(), c.end(), [](const RefRec& a, const RefRec& b) {
      if (a.angle != b.angle) return a.angle < b.angle;
      if (a.block != b.block) return a.block < b.block;
      if (a.templ != b.templ) return a.templ < b.templ;
      return a.chunk < b.chunk;
    });
    const int l = P_.window_size;
    if (ref_parts_) {
      // Parts of one (angle, block, template, chunk): the best score wins, ties to the smaller
      // bit-reversed offset in the chunk (get_peak_pos's tree, k:766-793); then the centre range
      // of src/norm.cu:583-590 applies to the winner.
      auto key = [](const RefRec& r) { return __builtin_bitreverse32((unsigned)(r.pixel % 256)) >> 24; };
      std::vector<RefRec> m;
      for (size_t i = 0; i < c.size();) {
        size_t j = i;
        RefRec w = c[i];
        for (; j < c.size() && c[j].angle == c[i].angle && c[j].block == c[i].block && c[j].templ == c[i].templ &&
               c[j].chunk == c[i].chunk;
             ++j)
          if (c[j].score > w.score || (c[j].score == w.score && key(c[j]) < key(w))) w = c[j];
        const int cx = offx_[w.block % bx_] + w.pixel % l, cy = offy_[w.block / bx_] + w.pixel / l;
        const float dm = P_.diameter;
        if (cx >= dm && cx < W_ - dm && cy >= dm && cy < H_ - dm) m.push_back(w);
        i = j;
      }
      c.swap(m);
    }
    records_.reserve(count);
    for (const RefRec& k : c) {
      gs_record r{};
      r.x = (float)(offx_[k.block % bx_] + k.pixel % l);
      r.y = (float)(offy_[k.block / bx_] + k.pixel / l);
      r.score = k.score;
      r.template_id = k.templ;
      r.angle_index = k.angle;
      float e = 360.0f - angles_[k.angle];  // src/norm.cu:654-655
      if (e >= 360.0f) e -= 360.0f;
      r.psi = e;
      r.block = k.block;
      r.pixel = k.pixel;
      r.local_x = (float)(k.pixel % l);
      r.local_y = (float)(k.pixel / l);
      records_.push_back(r);
    }
  }

  void collect_norm() {
    int count = 0;
    GS_HIP(hipMemcpy(&count, d_count_.p, sizeof(int), hipMemcpyDeviceToHost));
    GS_REQUIRE(count <= (1 << 20), kUnsupported, "too many candidates above threshold");
    std::vector<Cand> c(count);
    if (count) GS_HIP(hipMemcpy(c.data(), d_cand_.p, sizeof(Cand) * count, hipMemcpyDeviceToHost));
    const int l = P_.window_size;
    records_.reserve(count);
    for (const Cand& k : c) {
      gs_record r{};
      r.x = k.x;
      r.y = k.y;
      r.score = k.score;
      r.template_id = k.arg % n_;
      r.angle_index = k.arg / n_;
      float e = 360.0f - angles_[r.angle_index];  // src/norm.cu:654-655
      if (e >= 360.0f) e -= 360.0f;
      r.psi = e;
      r.block = k.block;
      r.pixel = k.pixel;
      r.local_x = (float)(k.pixel % l);
      r.local_y = (float)(k.pixel / l);
      records_.push_back(r);
    }
  }

  // norm_type = 0 (SearchNoNorm::work, src/nonorm.cu:559-589).
  void run_nonorm() {
    const int l = P_.window_size, hx = l / 2 + 1, n = n_;
    const size_t L2 = (size_t)l * l, T2 = (size_t)kTile * kTile;
    // The template preparation (-> d_that_ [n][hx][l]) runs on the second stream beside the
    // tiles' preparation on the main one: both are chains of small launches that leave most of
    // the GPU idle.  The two chains share no buffer (the tiles use d_scratch2_, d_m2_,
    // d_apart2_ and d_ichpart_), and the main stream waits for the templates before the
    // correlation.  The second stream first waits for the uploads already on the main one.
    ensure_pipeline();
    GS_HIP(hipEventRecord(ev_sync_, stream_));
    GS_HIP(hipStreamWaitEvent(stream2_, ev_sync_, 0));
    {
      const hipStream_t main = stream_;
      stream_ = stream2_;
      try {
        prep_templates();
      } catch (...) {
        stream_ = main;
        throw;
      }
      stream_ = main;
    }
    GS_HIP(hipEventRecord(ev_tprep_, stream2_));
    ulx_.clear();
    uly_.clear();
    tile_geometry(W_, H_, ulx_, uly_);
    const int ntiles = (int)ulx_.size();
    nbk_ = (int)block_offsets(kTile, l, P_.overlap).size();
    const int nbb = nbk_ * nbk_;
    // Every tile is searched at once: the blocks of all tiles (g = tile * nbb + block) are one
    // batch for the block normalisation, the rotation and the correlation launches.
    const int B = ntiles * nbb;
    std::vector<int> geo(nbk_ + 2 * ntiles);
    for (int i = 0; i < nbk_; ++i) geo[i] = i * (l - P_.overlap);  // split_IMG (k:615-620)
    for (int k = 0; k < ntiles; ++k) {
      geo[nbk_ + 2 * k] = ulx_[k];
      geo[nbk_ + 2 * k + 1] = uly_[k];
    }
    int* doff = d_tilegeo_.ensure(geo.size());
    const int* dul = doff + nbk_;
    GS_HIP(hipMemcpyAsync(doff, geo.data(), sizeof(int) * geo.size(), hipMemcpyHostToDevice, stream_));
    auto angles = angle_list(P_.phistep);
    int na = (int)angles.size();
    if (lim_angles_ > 0) na = std::min(na, lim_angles_);
    angles_ = angles;
    na_ = na;
    const bool gen = !corr_supported(l);  // generic correlation: one partial per plane
    const int groups = gen ? 1 : rowpass_groups(l);
    if (gen) corr_generic_prepare(fft_, l);
    // Angles are processed in groups: the rotated blocks of ag angles form the "block" axis of
    // one correlation launch (ag * B planes per template), so small micrographs still fill the
    // GPU.  Groups hold an even number of angles so that the rotation's angle pairs never
    // straddle two groups.
    int ag_max = std::max(1, std::min(na, 2048 / std::max(B, 1)));
    if (ag_max < na && ag_max > 1) ag_max &= ~1;
    const int nbp_max = ag_max * B;
    const int tc_max = std::min(n, chunk_for(nbp_max));
    const size_t zsz = gen ? 2 * corr_generic_buffer(l, tc_max, nbp_max) : (size_t)tc_max * nbp_max * l * (l / 2);
    float2* Z = d_Z_.ensure(zsz);
    float4* part = reinterpret_cast<float4*>(d_plane_part_.ensure((size_t)tc_max * nbp_max * groups * 4));
    // plane records [template][angle][tile * nbb + block][4]
    double* planes = d_planes_.ensure((size_t)n * na * B * 4);
    float2* F = d_frame_.ensure((size_t)ntiles * T2);
    float2* blk = d_blocks_.ensure((size_t)B * L2);
    const size_t npair_max = (size_t)(ag_max + 1) / 2 * B;
    float2* rotb = d_work2_.ensure(npair_max * L2);
    float2* ihat = d_ihat_.ensure((size_t)nbp_max * hx * l);
    float2* S = d_scratch2_.ensure(std::max((size_t)B, npair_max) * L2);
    double* m = d_m2_.ensure(B);
    std::vector<float2> cs(na);
    for (int a = 0; a < na; ++a)
      cs[a] = make_float2((float)std::cos((double)angles[a] * kRefPi / 180),
                          (float)std::sin((double)angles[a] * kRefPi / 180));
    float2* dcs = d_cs_.ensure(na);
    GS_HIP(hipMemcpyAsync(dcs, cs.data(), sizeof(float2) * na, hipMemcpyHostToDevice, stream_));
    const float border = P_.diameter / 4;
    std::optional<Range> loop_range(std::in_place, "angle_loop");
    GS_HIP(hipEventRecord(ev_corr0_, stream_));
    // TileImages (src/utils/tileimages.cpp:66-71), the per-tile preparation (src/nonorm.cu:346-389)
    // and SplitImage (src/nonorm.cu:391-450) of every tile
    launch_extract_tiles(d_img_.p, W_, dul, kTile, ntiles, P_.invert ? -1.f : 1.f, F, stream_);
    if (P_.phase_flip == 1) prep_frames(F, ntiles, kTile, kTile, false);
    launch_split_tile_blocks(F, kTile, l, doff, nbk_, ntiles, blk, stream_);
    if (P_.phase_flip == 1) {  // src/nonorm.cu:405-447
      fft2_planes(blk, S, B, -1, (float)(1.0 / (double)L2));
      launch_area_power_full(blk, B, l, d_apart2_.ensure((size_t)B * plane_reduction_chunks(B)), m, stream_);
      launch_scale_planes(blk, B, (long long)L2, m, 1.f, stream_);
      fft2_planes(blk, S, B, +1, 1.f);
    }
    float* blkr = d_blkreal_.ensure((size_t)B * L2);
    launch_extract_real(blk, blkr, (long long)B * L2, stream_);
    GS_HIP(hipStreamWaitEvent(stream_, ev_tprep_, 0));  // the template spectra d_that_
    for (int a0 = 0; a0 < na; a0 += ag_max) {
      const int ag = std::min(ag_max, na - a0), nbp = ag * B, np2 = (ag + 1) / 2 * B;
      // rotate_subIMG of every block for angles a0..a0+ag (k:534-603), two angles per plane, and
      // the forward transforms: half spectra [angle][block] scaled by 1/l^2 (src/nonorm.cu:457-470)
      launch_rotate_real_pairs(blkr, B, l, dcs + a0, ag, rotb, stream_);
      if (!launch_plane_fft2_pairs(rotb, rotb, np2, l, ag, B, nullptr, (float)(1.0 / (double)L2), ihat, stream_)) {
        fft2_planes(rotb, S, np2, -1, 1.f);
        launch_unpack_pair_spectra(rotb, ag, B, l, nullptr, (float)(1.0 / (double)L2), ihat, stream_);
      }
      for (int t0 = 0; t0 < n; t0 += tc_max) {
        const int tc = std::min(tc_max, n - t0);
        CorrArgs ca{d_that_.p, ihat, Z, t0, tc, nbp};
        clock_begin(clk_col_, stream_);
        if (gen) {
          launch_corr_generic_plane(fft_, ca, l, Z, Z + zsz / 2, part, border, stream_);
          clock_end(clk_col_, stream_);
        } else {
          // the norm_type=0 row pass reads the unpacked layout: variant 8 runs as 7 here
          GS_REQUIRE(launch_colpass_plane(l, ca, col_variant_ >= 8 ? 7 : col_variant_, stream_), kUnsupported,
                     "window_size not supported");
          clock_end(clk_col_, stream_);
          clock_begin(clk_row_, stream_);
          launch_rowpass_plane(l, Z, tc, nbp, part, border, stream_);
          clock_end(clk_row_, stream_);
        }
        // record of (template t0+t, angle a0+q, block g) at ((t0+t)*na + a0 + q)*B + g
        launch_finalize_plane(part, tc, nbp, groups, (long long)na * B,
                              planes + ((size_t)t0 * na + a0) * B * 4, stream_);
      }
    }
    GS_HIP(hipEventRecord(ev_corr1_, stream_));
    loop_range.reset();
    planes_host_.resize((size_t)n * na * B * 4);
  }

  void collect_nonorm() {
    GS_HIP(hipMemcpy(planes_host_.data(), d_planes_.p, sizeof(double) * planes_host_.size(),
                     hipMemcpyDeviceToHost));
    const int l = P_.window_size, n = n_, nbb = nbk_ * nbk_, B = (int)ulx_.size() * nbb;
    const double rc = (double)l * l - 1;
    for (int tI = 0; tI < (int)ulx_.size(); ++tI)
      for (int a = 0; a < na_; ++a) {
        const double e = angles_[a];
        const double ce = std::cos(e * kRefPi / 180), se = std::sin(e * kRefPi / 180);
        const double* P = planes_host_.data();
        for (int t = 0; t < n; ++t) {
          double sc = 0, cxs = 0, cys = 0;
          int pix = 0, blkid = 0;
          for (int b = 0; b < nbb; ++b) {  // PickParticles host scoring, src/nonorm.cu:509-534
            const double* q = P + (((size_t)t * na_ + a) * B + tI * nbb + b) * 4;
            const double peak = q[0], pos = q[1], sum = q[2], sum2 = q[3];
            const double ra = sum - peak, rb = sum2 - peak * peak;
            const double sd = std::sqrt(rb / rc - (ra / rc) * (ra / rc));
            const double score = sd == 0 ? 0.0 : peak / sd;
            const int cx = (int)pos % l, cy = (int)pos / l;
            const int bi = b % nbk_, bj = b / nbk_;
            const float centerx = (float)(bi * (l - P_.overlap) + (cx - l / 2) * ce + (cy - l / 2) * se + l / 2);
            const float centery = (float)(bj * (l - P_.overlap) + (cy - l / 2) * ce - (cx - l / 2) * se + l / 2);
            if (sc < score && centerx >= 0 && centerx < kTile && centery >= 0 && centery < kTile) {
              sc = score;
              cxs = centerx;
              cys = centery;
              pix = (int)pos;
              blkid = b;
            }
          }
          if (sc >= P_.threshold) {  // OutputScore, src/nonorm.cu:539-557
            gs_record r{};
            r.local_x = (float)cxs;
            r.local_y = (float)cys;
            r.x = (float)(ulx_[tI] + cxs);
            r.y = (float)(uly_[tI] + cys);
            r.template_id = t;
            r.angle_index = a;
            r.psi = angles_[a];
            r.score = (float)sc;
            r.block = tI;
            r.pixel = blkid * (int)(l * l) + pix;
            records_.push_back(r);
          }
        }
      }
  }

  gs_params P_;
  int device_;
  hipStream_t stream_ = nullptr, stream2_ = nullptr;
  hipEvent_t ev_zready_[2] = {nullptr, nullptr}, ev_zfree_[2] = {nullptr, nullptr}, ev_sync_ = nullptr;
  hipEvent_t ev_tprep_ = nullptr, ev_rfork_ = nullptr, ev_rjoin_ = nullptr;
  bool overlap_ = false;
  bool ref_parts_ = false;  // the reference records of the last run are chunk parts (tiled pass)
  hipEvent_t ev_start_ = nullptr, ev_corr0_ = nullptr, ev_corr1_ = nullptr, ev_end_ = nullptr;
  KernelClock clk_col_, clk_row_;
  bool profile_ = false;
  FftCache fft_;
  DevParams dp_{};
  int n_ = 0, w_ = 0, h_ = 0;
  std::vector<float> alt_, az_;
  int W_ = 0, H_ = 0;
  gs_ctf ctf_{};
  bool have_image_ = false, ran_ = false;
  int nb_ = 0, bx_ = 0, nbk_ = 0, na_ = 0;
  std::vector<int> offx_, offy_, ulx_, uly_;
  std::vector<float> angles_;
  // kernel variants default to the fastest measured on the production shape (l = 320)
  int lim_angles_ = 0, lim_templates_ = 0, chunk_ = 0, row_variant_ = 8, col_variant_ = 12;
  bool graph_allowed_ = true;  // false while a batch runs an item whose shape occurs only once
  double mem_budget_ = 0;  // bytes the automatic templates-per-launch may give the inter-pass buffer
  double total_ms_ = 0, corr_ms_ = 0;
  std::vector<gs_record> records_;
  std::vector<double> planes_host_;
  double total_n_ = 0, refine_ms_ = 0;
  static constexpr size_t kRefineChunk = 2048;  // picks per refinement batch (about 4 MB each at l=320)
  bool white_ok_ = false;
  DevBuf<float2> d_twhite_, d_rS_, d_rIT_;
  DevBuf<int32_t> d_ridx_;
  DevBuf<RefWeight> d_rtab_;
  DevBuf<double2> d_apart_, d_apart2_;
  // buffers of one refinement batch (one per stream lane) and its host copies
  struct RefLane {
    DevBuf<float2> H, T, AB, ph;
    DevBuf<RefPeak> pk;
    DevBuf<double2> part;
    DevBuf<double> part6, stats, cg, z, z0, grad;
    std::vector<RefPeak> hp;
    std::vector<double> hz, hz0, hg;
  };
  RefLane rlane_[2];
  static constexpr size_t kRefineMinLane = 64;  // picks per lane before a second lane is used
  static constexpr size_t kMaxGraphs = 8;
  bool use_graph_ = false;
  std::map<std::vector<long long>, CapturedGraph> graphs_;
  long long graph_captures_ = 0, graph_launches_ = 0;
  PinnedBuf pinned_[2];
  std::map<long long, std::unique_ptr<DevBuf<double>>> rb_cache_;

  DevBuf<float> d_rmask_;
  DevBuf<float> d_raw_, d_img_, d_imgpre_, d_tprep_, d_stats_, d_best_, d_blkreal_;
  DevBuf<float2> d_work_, d_work2_, d_scratch_, d_scratch2_, d_frame_, d_frame2_, d_blocks_, d_ihat_, d_that_, d_Z_,
      d_cs_;
  DevBuf<double> d_mom_, d_ra_, d_ra_img_, d_m_, d_m2_, d_part_, d_S_, d_SS_, d_mu_, d_sigma_, d_planes_, d_chpart_,
      d_ichpart_;
  DevBuf<float> d_plane_part_;
  DevBuf<int32_t> d_arg_;
  DevBuf<int> d_offx_, d_offy_, d_count_, d_tilegeo_;
  DevBuf<Cand> d_cand_;
  static constexpr int kMaxRefRecords = 1 << 22;
  DevBuf<RefRec> d_refrec_;
  DevBuf<int> d_refcount_;
};

}  // namespace gs

// ------------------------------------------------------------------ C ABI
struct gs_engine {
  std::unique_ptr<gs::Engine> impl;
};

namespace {
thread_local std::string g_last_error;

gs_status fail(int code, const std::string& msg) {
  g_last_error = msg;
  return (gs_status)code;
}

template <typename F>
gs_status guarded(F&& f) {
  try {
    f();
    g_last_error.clear();
    return GS_OK;
  } catch (const gs::Error& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(GS_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(GS_ERR_INVALID, e.what());
  } catch (...) {
    return fail(GS_ERR_INVALID, "unknown error");
  }
}
}  // namespace

namespace gs {
// The status plumbing above for the C entry points of other source files.
gs_status c_abi_guarded(const std::function<void()>& f) { return guarded(f); }
}  // namespace gs

extern "C" {

gs_status gs_create(const gs_params* params, int device, gs_engine** out) {
  if (!params || !out) return fail(GS_ERR_INVALID, "null argument");
  return guarded([&] {
    auto e = std::make_unique<gs_engine>();
    e->impl = std::make_unique<gs::Engine>(*params, device);
    *out = e.release();
  });
}

gs_status gs_load_templates(gs_engine* e, const float* data, int n, int w, int h, const float* alt_az) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  return guarded([&] { e->impl->load_templates(data, n, w, h, alt_az); });
}

gs_status gs_set_image(gs_engine* e, const float* image, int width, int height, const gs_ctf* ctf) {
  if (!e || !ctf) return fail(GS_ERR_INVALID, "null argument");
  return guarded([&] { e->impl->set_image(image, width, height, *ctf); });
}

gs_status gs_run(gs_engine* e) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  return guarded([&] { e->impl->run(); });
}

gs_status gs_get_records(gs_engine* e, gs_record** out, size_t* n_out) {
  if (!e || !out || !n_out) return fail(GS_ERR_INVALID, "null argument");
  return guarded([&] {
    const auto& r = e->impl->records();
    gs_record* p = (gs_record*)std::malloc(sizeof(gs_record) * std::max<size_t>(r.size(), 1));
    if (!p) throw std::bad_alloc();
    if (!r.empty()) std::memcpy(p, r.data(), sizeof(gs_record) * r.size());
    *out = p;
    *n_out = r.size();
  });
}

gs_status gs_search(gs_engine* e, const float* image, int width, int height, const gs_ctf* ctf,
                    gs_record** out, size_t* n_out) {
  gs_status s = gs_set_image(e, image, width, height, ctf);
  if (s != GS_OK) return s;
  s = gs_run(e);
  if (s != GS_OK) return s;
  return gs_get_records(e, out, n_out);
}

int gs_format_record(const gs_engine* e, const gs_record* r, int image_index, const char* path,
                     const gs_ctf* ctf, int tile_local, char* buf, size_t cap) {
  if (!e || !r || !ctf) return -1;
  const float cx = tile_local ? r->local_x : r->x, cy = tile_local ? r->local_y : r->y;
  // float fields as the reference stores them (Parameters is float, src/norm.cu:185-187)
  return std::snprintf(buf, cap,
                       "%d\t%s\tdefocus=%f\tdfdiff=%f\tdfang=%f\teuler=%f,%f,%f\tcenter=%f,%f\tscore=%f\n",
                       image_index, path ? path : "", (float)ctf->defocus, (float)ctf->dfdiff,
                       (float)ctf->dfang, e->impl->alt(r->template_id), e->impl->az(r->template_id),
                       r->psi, cx, cy, r->score);
}

void gs_free(void* p) { std::free(p); }

void gs_destroy(gs_engine* e) {
  if (!e) return;
  try {
    delete e;
  } catch (...) {
  }
}

const char* gs_last_error(void) { return g_last_error.c_str(); }

size_t gs_stage_bytes(const gs_engine* e, int stage) {
  if (!e) return 0;
  try {
    return e->impl->stage_bytes(stage);
  } catch (...) {
    return 0;
  }
}

gs_status gs_fetch_stage(gs_engine* e, int stage, void* out, size_t bytes) {
  if (!e || !out) return fail(GS_ERR_INVALID, "null argument");
  return guarded([&] { e->impl->fetch(stage, out, bytes); });
}

gs_status gs_set_limits(gs_engine* e, int n_angles, int n_templates) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  return guarded([&] { e->impl->set_limits(n_angles, n_templates); });
}

gs_status gs_fft_c2c(int device, float* data, int n, int batch, int sign) {
  if (!data || n < 1 || batch < 1 || (sign != 1 && sign != -1)) return fail(GS_ERR_INVALID, "bad FFT arguments");
  return guarded([&] {
    GS_HIP(hipSetDevice(device));
    gs::FftCache cache;
    const size_t N = (size_t)n * batch;
    gs::DevBuf<float2> a, b;
    a.ensure(N);
    b.ensure(N);
    GS_HIP(hipMemcpy(a.p, data, sizeof(float2) * N, hipMemcpyHostToDevice));
    gs::fft_rows(cache, a.p, b.p, n, batch, sign, 1.f, nullptr);
    GS_HIP(hipMemcpy(data, a.p, sizeof(float2) * N, hipMemcpyDeviceToHost));
  });
}

gs_status gs_fft2_c2c(int device, float* data, int ny, int nx, int batch, int sign) {
  if (!data || nx < 1 || ny < 1 || batch < 1 || (sign != 1 && sign != -1))
    return fail(GS_ERR_INVALID, "bad FFT arguments");
  return guarded([&] {
    GS_HIP(hipSetDevice(device));
    gs::FftCache cache;
    const size_t N = (size_t)nx * ny * batch;
    gs::DevBuf<float2> a, b;
    a.ensure(N);
    b.ensure(N);
    GS_HIP(hipMemcpy(a.p, data, sizeof(float2) * N, hipMemcpyHostToDevice));
    // square planes of the built window sizes take the LDS plane transform
    if (nx != ny || !gs::launch_plane_fft2(a.p, batch, nx, sign, 1.f, nullptr))
      gs::fft2_batched(cache, a.p, b.p, batch, ny, nx, sign, 1.f, nullptr);
    GS_HIP(hipMemcpy(data, a.p, sizeof(float2) * N, hipMemcpyDeviceToHost));
  });
}

namespace {
// Streaming copy: four 16-byte loads in flight per thread, non-temporal loads and stores (the
// fastest of the plain, unrolled and non-temporal forms measured on MI355X at 1 and 4 GiB).
constexpr int kCopyUnroll = 4, kCopyGrid = 16384;
__global__ void k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i0 = (long long)blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += stride * kCopyUnroll) {
    float4 v[kCopyUnroll];
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) v[u] = ld_stream(a + i);
    }
#pragma unroll
    for (int u = 0; u < kCopyUnroll; ++u) {
      const long long i = i0 + u * stride;
      if (i < n) st_stream(b + i, v[u]);
    }
  }
}
}  // namespace

gs_status gs_copy_bandwidth(int device, size_t bytes, int reps, double* gbps) {
  if (!gbps || bytes < 16 || reps < 1) return fail(GS_ERR_INVALID, "bad copy arguments");
  return guarded([&] {
    GS_HIP(hipSetDevice(device));
    const long long n = (long long)(bytes / 16);
    gs::DevBuf<float4> a, b;
    a.ensure(n);
    b.ensure(n);
    GS_HIP(hipMemset(a.p, 0, n * 16));
    hipEvent_t e0, e1;
    GS_HIP(hipEventCreate(&e0));
    GS_HIP(hipEventCreate(&e1));
    k_stream_copy<<<kCopyGrid, 256>>>(a.p, b.p, n);
    GS_CHECK_LAUNCH();
    GS_HIP(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) k_stream_copy<<<kCopyGrid, 256>>>(a.p, b.p, n);
    GS_CHECK_LAUNCH();
    GS_HIP(hipEventRecord(e1, nullptr));
    GS_HIP(hipEventSynchronize(e1));
    float ms = 0;
    GS_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *gbps = 2.0 * (double)n * 16 * reps / (ms * 1e-3) / 1e9;  // read + write
  });
}

gs_status gs_last_timing(const gs_engine* e, double* total_ms, double* corr_ms) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  e->impl->timing(total_ms, corr_ms);
  return GS_OK;
}


gs_status gs_set_profiling(gs_engine* e, int on) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  e->impl->set_profiling(on != 0);
  return GS_OK;
}

gs_status gs_kernel_timing(const gs_engine* e, int which, double* avg_ms, long long* count) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  e->impl->kernel_timing(which, avg_ms, count);
  return GS_OK;
}

gs_status gs_set_chunk(gs_engine* e, int templates_per_launch) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  if (templates_per_launch < 0) return fail(GS_ERR_INVALID, "templates per launch must be 0 (automatic) or positive");
  e->impl->set_chunk(templates_per_launch);
  return GS_OK;
}

gs_status gs_set_variant(gs_engine* e, int kernel, int variant) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  return guarded([&] { e->impl->set_variant(kernel, variant); });
}

gs_status gs_search_batch(gs_engine* e, const gs_batch_item* items, int n_items, gs_record** out,
                          size_t* counts) {
  if (!e || !items || !out || !counts || n_items <= 0) return fail(GS_ERR_INVALID, "null argument or empty batch");
  return guarded([&] {
    std::vector<gs_record> all;
    e->impl->search_batch(items, n_items, all, counts);
    gs_record* p = (gs_record*)std::malloc(sizeof(gs_record) * std::max<size_t>(all.size(), 1));
    if (!p) throw std::bad_alloc();
    if (!all.empty()) std::memcpy(p, all.data(), sizeof(gs_record) * all.size());
    *out = p;
  });
}

gs_status gs_set_graph(gs_engine* e, int on) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  return guarded([&] { e->impl->set_graph(on != 0); });
}

gs_status gs_graph_stats(const gs_engine* e, long long* captures, long long* launches) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  e->impl->graph_stats(captures, launches);
  return GS_OK;
}

gs_status gs_autotune(gs_engine* e, int* col_variant, int* row_variant, int* templates_per_launch) {
  if (!e) return fail(GS_ERR_INVALID, "null engine");
  return guarded([&] { e->impl->autotune(col_variant, row_variant, templates_per_launch); });
}

gs_status gs_select(gs_engine* e, float threshold, float nms_radius, gs_record** out, size_t* n_out) {
  if (!e || !out || !n_out) return fail(GS_ERR_INVALID, "null argument");
  return guarded([&] {
    const auto r = e->impl->select(threshold, nms_radius);
    gs_record* p = (gs_record*)std::malloc(sizeof(gs_record) * std::max<size_t>(r.size(), 1));
    if (!p) throw std::bad_alloc();
    if (!r.empty()) std::memcpy(p, r.data(), sizeof(gs_record) * r.size());
    *out = p;
    *n_out = r.size();
  });
}

gs_refine_opts gs_refine_defaults(void) {
  gs_refine_opts o{};
  o.iterations = 20;
  o.step = 0.1;
  o.scale[0] = 2.0;
  o.scale[1] = 2.0;
  o.scale[2] = 3.0;
  o.scale[3] = 500.0;
  return o;
}

gs_status gs_peak_from_record(const gs_engine* e, const gs_record* r, gs_peak* out) {
  if (!e || !r || !out) return fail(GS_ERR_INVALID, "null argument");
  return guarded([&] { *out = e->impl->peak_from_record(*r); });
}

gs_status gs_score_and_grad(gs_engine* e, const gs_peak* peaks, size_t k, double* score, double* grad) {
  if (!e || !peaks || !score || !grad) return fail(GS_ERR_INVALID, "null argument");
  return guarded([&] {
    const double unit[4] = {1, 1, 1, 1};
    e->impl->refine(peaks, k, 0, 0.0, unit, nullptr, score, nullptr, grad);
  });
}

gs_status gs_refine(gs_engine* e, const gs_peak* peaks, size_t k, const gs_refine_opts* opts,
                    gs_refined* out) {
  if (!e || !peaks || !out) return fail(GS_ERR_INVALID, "null argument");
  return guarded([&] {
    const gs_refine_opts o = opts ? *opts : gs_refine_defaults();
    std::vector<gs_peak> pk(k);
    std::vector<double> z(k), z0(k), g(4 * k);
    e->impl->refine(peaks, k, o.iterations, o.step, o.scale, pk.data(), z.data(), z0.data(), g.data());
    for (size_t i = 0; i < k; ++i) {
      out[i].peak = pk[i];
      out[i].score = z[i];
      out[i].score0 = z0[i];
      for (int q = 0; q < 4; ++q) out[i].grad[q] = g[4 * i + q];
    }
  });
}

gs_status gs_search_batch_refine(gs_engine* e, const gs_batch_item* items, int n_items, int top_k,
                                 const gs_refine_opts* opts, gs_record** out, size_t* counts,
                                 gs_refined** refined, size_t* refined_counts) {
  if (!e || !items || !out || !counts || !refined || !refined_counts || n_items <= 0 || top_k < 0)
    return fail(GS_ERR_INVALID, "null argument, empty batch or negative top_k");
  return guarded([&] {
    const gs_refine_opts o = opts ? *opts : gs_refine_defaults();
    std::vector<gs_record> all;
    std::vector<gs_refined> ref;
    std::fill(refined_counts, refined_counts + n_items, (size_t)0);
    e->impl->search_batch(items, n_items, all, counts, top_k, &o, &ref, refined_counts);
    gs_record* p = (gs_record*)std::malloc(sizeof(gs_record) * std::max<size_t>(all.size(), 1));
    gs_refined* q = (gs_refined*)std::malloc(sizeof(gs_refined) * std::max<size_t>(ref.size(), 1));
    if (!p || !q) {
      std::free(p);
      std::free(q);
      throw std::bad_alloc();
    }
    if (!all.empty()) std::memcpy(p, all.data(), sizeof(gs_record) * all.size());
    if (!ref.empty()) std::memcpy(q, ref.data(), sizeof(gs_refined) * ref.size());
    *out = p;
    *refined = q;
  });
}

gs_status gs_refine_timing(const gs_engine* e, double* ms) {
  if (!e || !ms) return fail(GS_ERR_INVALID, "null argument");
  *ms = e->impl->refine_ms();
  return GS_OK;
}

}  // extern "C"
