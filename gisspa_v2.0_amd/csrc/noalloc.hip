// Allocation-free entry point: the 2-D transform of gs_fft2_c2c on caller-owned device memory,
// with its work buffer and twiddle tables in a file-scope device array.
#include <functional>
#include <mutex>
#include <string>

#include "../../include/gisspa.h"
#include "common.hpp"
#include "fft_generic.hpp"

namespace gs {
// Defined in engine.hip: runs f, maps exceptions to a status and sets gs_last_error.
gs_status c_abi_guarded(const std::function<void()>& f);
}

namespace {

// Work buffer of the no-allocation transform: the Stockham ping-pong buffer (one element per
// element of the input) followed by the twiddle tables of the row and column lengths.  Every
// device the library runs on holds one copy of it, at its own address, until the process ends.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float2 scratch[kScratchBytes / sizeof(float2)];

std::mutex g_mu;
float2* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current).  The array's name in host code is a
// host-side stand-in, not a device address, so kernels only ever receive this pointer.
float2* scratch_ptr(int device) {
  GS_REQUIRE(device >= 0 && device < kMaxDevices, gs::kInvalid, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    GS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GS_REQUIRE(p != nullptr, gs::kHip, "no device address for the transform work buffer");
    g_ptr[device] = static_cast<float2*>(p);
  }
  return g_ptr[device];
}

}  // namespace

extern "C" gs_status gs_fft2_c2c_noalloc(int device, float* d_data, int ny, int nx, int batch, int sign,
                                         void* stream) {
  return gs::c_abi_guarded([&] {
    GS_REQUIRE(d_data && nx >= 1 && ny >= 1 && batch >= 1 && (sign == 1 || sign == -1), gs::kInvalid,
               "bad FFT arguments");
    const size_t n = (size_t)nx * ny * batch;
    const size_t need = n + (size_t)nx + (size_t)ny;
    const size_t cap = kScratchBytes / sizeof(float2);
    GS_REQUIRE(need <= cap, gs::kInvalid,
               "gs_fft2_c2c_noalloc: batch * ny * nx + nx + ny = " + std::to_string(need) +
                   " complex elements exceed the work buffer of " + std::to_string(cap));
    GS_HIP(hipSetDevice(device));
    float2* work = scratch_ptr(device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    gs::FftTable rx{}, ry{};
    rx.nrad = gs::factor_radices(nx, rx.radices, gs::FftTable::kMax);
    ry.nrad = gs::factor_radices(ny, ry.radices, gs::FftTable::kMax);
    float2* twx = work + n;
    float2* twy = twx + nx;
    // every element of the work buffer that is read is first written on the same stream: the
    // twiddles here, the ping-pong part by the first transform stage
    gs::launch_fill_twiddles(twx, nx, s);
    gs::launch_fill_twiddles(twy, ny, s);
    rx.tw = twx;
    ry.tw = twy;
    gs::fft2_table(rx, ry, reinterpret_cast<float2*>(d_data), work, batch, ny, nx, sign, 1.f, s);
  });
}
