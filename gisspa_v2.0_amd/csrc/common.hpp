// Shared helpers for the GisSPA HIP library: error handling and complex arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>

namespace gs {

// Host-side roctx range around a stage of the search (template and image preparation, the
// angle loop, finalisation, record collection, refinement); `rocprofv3 --marker-trace` shows
// them beside the kernels.  Kernel launches are asynchronous, so a range covers the stage's
// enqueue time unless the stage ends in a synchronisation (the whole search does).
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes mirrored by include/gisspa.h.
enum : int { kOk = 0, kInvalid = 1, kHip = 2, kUnsupported = 3, kNoMem = 4, kState = 5 };

[[noreturn]] inline void hip_fail(hipError_t e, const char* what, const char* file, int line) {
  throw Error(e == hipErrorOutOfMemory ? kNoMem : kHip,
              std::string(what) + ": " + hipGetErrorString(e) + " (" + file + ":" +
                  std::to_string(line) + ")");
}

}  // namespace gs

#define GS_HIP(x)                                                   \
  do {                                                              \
    hipError_t e_ = (x);                                            \
    if (e_ != hipSuccess) ::gs::hip_fail(e_, #x, __FILE__, __LINE__); \
  } while (0)

#define GS_CHECK_LAUNCH() GS_HIP(hipGetLastError())

#define GS_REQUIRE(cond, code, msg)                 \
  do {                                              \
    if (!(cond)) throw ::gs::Error((code), (msg));  \
  } while (0)

__device__ __forceinline__ float2 c_add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 c_sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 c_mul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 c_conj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 c_scale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// a * conj(b)
__device__ __forceinline__ float2 c_mulconj(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// Streaming store of a value that no later instruction of the kernel reads back (non-temporal
// hint: the line is not kept in L2 at the expense of lines the kernel is still reading).
__device__ __forceinline__ void st_stream(float2* p, float2 v) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  __builtin_nontemporal_store(f2v{v.x, v.y}, reinterpret_cast<f2v*>(p));
}
__device__ __forceinline__ void st_stream(float4* p, float4 v) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  __builtin_nontemporal_store(f4v{v.x, v.y, v.z, v.w}, reinterpret_cast<f4v*>(p));
}
// Streaming load of a value read once (non-temporal hint).
__device__ __forceinline__ float2 ld_stream(const float2* p) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ float4 ld_stream(const float4* p) {
  typedef float f4v __attribute__((ext_vector_type(4)));
  const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
// i^k * a for the sign convention of the caller (s = +1 multiplies by +i, -1 by -i)
__device__ __forceinline__ float2 c_muli(float2 a, int s) {
  return s > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
