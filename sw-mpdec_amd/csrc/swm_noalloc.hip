// swm_noalloc.hip -- TimeStep of a state the caller already holds on the
// device, with no device allocation: swm_timestep_noalloc.
//
// The engine's TimeStep (k_timestep + k_timestep_final) keeps its per-block
// partial maxima in a buffer the library allocates with each handle.  This
// variant keeps them in `scratch`, a file-scope device array of fixed size, so
// a caller that must not allocate (code under HIP graph capture, a serving
// loop that forbids hipMalloc) can compute dt with only its own input and
// output buffers.  Both variants run the same block routines
// (swm_timestep.hpp), so their results are bitwise equal.
//
// Each device holds its own copy of `scratch` until the process ends.  In
// host code the array's name is not a device address: the address is looked
// up once per device with hipGetSymbolAddress and only that pointer is passed
// to kernels.  Every call on a device shares that device's array, so calls
// that overlap in time on one device (two host threads, or two streams
// without an ordering between them) race; calls on one stream are ordered.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "swm.h"
#include "swm_engine.hpp"
#include "swm_timestep.hpp"

namespace swm {

constexpr int kNoAllocMaxBlocks = SWM_TIMESTEP_NOALLOC_MAX_CELLS / 256;

// per-block (max a, max b, NaN seen) triples of one call
static __device__ __align__(256) double scratch[3 * kNoAllocMaxBlocks];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's `scratch`, looked up on the first call on
// the device (with it current) and kept.
hipError_t scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

// Cell k (0-based, x fastest) of U(1:3, 1-g:nx+g, 1-g:ny+g).
struct AosCells {
  const double* U;
  int nx, g;
  __device__ __forceinline__ size_t at(int k) const {
    const int i = k % nx, j = k / nx;
    return 3 * ((size_t)(j + g) * (size_t)(nx + 2 * g) + (size_t)(i + g));
  }
};

__global__ __launch_bounds__(256) void k_timestep_aos(AosCells C, int n, double g,
                                                      double* part) {
  const int k = blockIdx.x * 256 + threadIdx.x;
  const bool in = k < n;
  const size_t o = in ? C.at(k) : 0;
  timestep_partial(in, g, in ? C.U[o] : 0., in ? C.U[o + 1] : 0., in ? C.U[o + 2] : 0.,
                   part + 3 * (size_t)blockIdx.x);
}

__global__ __launch_bounds__(256) void k_timestep_aos_final(AosCells C, int n, int nb,
                                                            double g, double dx, double dy,
                                                            double cfl, const double* part,
                                                            double* out) {
  double ts, LX, LY;
  timestep_final(
      part, nb, n,
      [&](int k, double& a, double& b, double& c) {
        const size_t o = C.at(k);
        a = C.U[o];
        b = C.U[o + 1];
        c = C.U[o + 2];
      },
      g, dx, dy, cfl, ts, LX, LY);
  if (threadIdx.x != 0) return;
  out[0] = ts;
  out[1] = LX;
  out[2] = LY;
}

}  // namespace
}  // namespace swm

extern "C" int swm_timestep_noalloc(int nx, int ny, double dx, double dy, double gravity,
                                    double cfl, const double* U, int ghosts, double* out,
                                    void* stream) {
  using namespace swm;
  if (nx < 1 || ny < 1 || (long long)nx * ny > SWM_TIMESTEP_NOALLOC_MAX_CELLS) {
    set_thread_error("swm_timestep_noalloc: grid " + std::to_string(nx) + " x " +
                     std::to_string(ny) + " is outside 1 .. " +
                     std::to_string(SWM_TIMESTEP_NOALLOC_MAX_CELLS) +
                     " cells (the size of the library's fixed scratch array)");
    return SWM_E_ARG;
  }
  if (ghosts != 0 && ghosts != 3) {
    set_thread_error("swm_timestep_noalloc: ghosts must be 0 or 3");
    return SWM_E_ARG;
  }
  if (!U || !out) {
    set_thread_error("swm_timestep_noalloc: U and out must be device pointers");
    return SWM_E_ARG;
  }
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_thread_error(std::string("swm_timestep_noalloc: hipGetDevice failed: ") +
                     hipGetErrorString(e));
    return SWM_E_HIP;
  }
  if (dev < 0 || dev >= kMaxDevices) {
    set_thread_error("swm_timestep_noalloc: device " + std::to_string(dev) +
                     " is beyond the " + std::to_string(kMaxDevices) +
                     " devices the library keeps scratch addresses for");
    return SWM_E_HIP;
  }
  double* part = nullptr;
  e = scratch_address(dev, &part);
  if (e != hipSuccess) {
    set_thread_error(std::string("swm_timestep_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                     hipGetErrorString(e));
    return SWM_E_HIP;
  }
  const int n = nx * ny, nb = (n + 255) / 256;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const AosCells C{U, nx, ghosts};
  hipLaunchKernelGGL(k_timestep_aos, dim3(nb), dim3(256), 0, s, C, n, gravity, part);
  hipLaunchKernelGGL(k_timestep_aos_final, dim3(1), dim3(256), 0, s, C, n, nb, gravity, dx, dy,
                     cfl, part, out);
  e = hipGetLastError();
  if (e != hipSuccess) {
    set_thread_error(std::string("swm_timestep_noalloc: kernel launch failed: ") +
                     hipGetErrorString(e));
    return SWM_E_HIP;
  }
  return SWM_OK;
}
