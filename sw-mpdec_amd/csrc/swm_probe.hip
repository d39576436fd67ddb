// swm_probe.hip -- measurement of what a step boundary costs on this GPU:
// a kernel boundary (one launch per round, plain or replayed from a HIP
// graph) against a grid-wide barrier inside one cooperative launch.
//
// This informs the engine's design (one launch per stage of a step rather
// than one persistent kernel with grid barriers between stages); it is not
// used by any step.  Each round, every block reads the value its neighbour
// block wrote in the previous round and writes its own, so a round boundary
// must make the writes of every block visible to every other block: at a
// kernel boundary the runtime's release/acquire does it, inside the
// cooperative kernel the barrier's agent-scope fences do.
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "swm.h"
#include "swm_engine.hpp"

namespace swm {
namespace {

// one round of the chain: block b reads block b+1's value of the previous
// round (ping-pong buffers) and writes its own
__device__ __forceinline__ void round_body(const double* in, double* out, int nblk, int r) {
  if (threadIdx.x == 0) {
    const int b = blockIdx.x;
    out[b] = in[(b + 1) % nblk] + (double)r;
  }
}

__global__ void k_probe_round(const double* in, double* out, int nblk, int r) {
  round_body(in, out, nblk, r);
}

// Grid barrier for a cooperative launch: every block arrives on a counter
// (agent-scope release), the last one bumps the generation, the others wait
// for it (agent-scope acquire).  The wait is bounded: after spin_limit polls
// the block sets *failed and leaves the wait, and from then on no barrier
// waits, so the kernel always ends.
__device__ void grid_barrier(unsigned* count, unsigned* gen, unsigned nblk,
                             unsigned long long spin_limit, int* failed) {
  __syncthreads();
  // once a wait has failed, no later barrier waits (the kernel ends quickly)
  if (threadIdx.x == 0 && !__hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
    const unsigned g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived =
        __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == nblk - 1) {
      __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned long long polls = 0;
      while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++polls > spin_limit) {
          __hip_atomic_store(failed, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
  }
  __syncthreads();
}

__global__ void k_probe_barrier_rounds(double* a, double* b, int nblk, int rounds, unsigned* count,
                                       unsigned* gen, unsigned long long spin_limit,
                                       int* failed) {
  for (int r = 0; r < rounds; ++r) {
    round_body((r & 1) ? b : a, (r & 1) ? a : b, nblk, r);
    grid_barrier(count, gen, (unsigned)nblk, spin_limit, failed);
  }
}

}  // namespace
}  // namespace swm

#define PCK(call)                                                                       \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) {                                                             \
      swm::set_thread_error(std::string("swm_debug_sync_cost: ") + #call + ": " +       \
                            hipGetErrorString(_e));                                     \
      rc = SWM_E_HIP;                                                                   \
      goto done;                                                                        \
    }                                                                                   \
  } while (0)

extern "C" int swm_debug_sync_cost(int nblk, int rounds, double us_per_round[3]) {
  using namespace swm;
  if (nblk < 1 || rounds < 1 || rounds > 100000 || !us_per_round) {
    set_thread_error("swm_debug_sync_cost: need nblk >= 1, 1 <= rounds <= 100000");
    return SWM_E_ARG;
  }
  int rc = SWM_OK;
  int dev = 0, ncu = 0, per_cu = 0, coop = 0, hfailed = 0;
  double *a = nullptr, *b = nullptr;
  unsigned* ctr = nullptr;
  int* failed = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  float ms = 0.f;
  const unsigned long long spin_limit = 1ull << 20;  // about a second of polls
  // every block's value after `rounds` rounds from zeros: sum of 0..rounds-1
  const double expect = 0.5 * (double)rounds * (double)(rounds - 1);
  std::vector<double> host(nblk);
  auto reset = [&]() {
    hipError_t e = hipMemsetAsync(a, 0, nblk * sizeof(double), s);
    if (e == hipSuccess) e = hipMemsetAsync(b, 0, nblk * sizeof(double), s);
    return e;
  };
  // the chain's result must reach every block in every mode
  auto check = [&](const char* mode) {
    hipError_t e = hipMemcpy(host.data(), (rounds & 1) ? b : a, nblk * sizeof(double),
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return (int)SWM_E_HIP;
    for (int k = 0; k < nblk; ++k)
      if (host[k] != expect) {
        set_thread_error(std::string("swm_debug_sync_cost: ") + mode + ": block " +
                         std::to_string(k) + " holds " + std::to_string(host[k]) + ", expected " +
                         std::to_string(expect));
        return (int)SWM_E_STATE;
      }
    return (int)SWM_OK;
  };
  PCK(hipGetDevice(&dev));
  PCK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  PCK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  PCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_probe_barrier_rounds, 64, 0));
  if (!coop || nblk > ncu * per_cu) {
    set_thread_error("swm_debug_sync_cost: " + std::to_string(nblk) +
                     " blocks cannot all be resident (cooperative launch limit " +
                     std::to_string(ncu * per_cu) + ")");
    return SWM_E_ARG;
  }
  PCK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  PCK(hipEventCreate(&e0));
  PCK(hipEventCreate(&e1));
  PCK(hipMalloc((void**)&a, nblk * sizeof(double)));
  PCK(hipMalloc((void**)&b, nblk * sizeof(double)));
  PCK(hipMalloc((void**)&ctr, 2 * sizeof(unsigned)));
  PCK(hipMalloc((void**)&failed, sizeof(int)));
  PCK(hipMemsetAsync(ctr, 0, 2 * sizeof(unsigned), s));
  PCK(hipMemsetAsync(failed, 0, sizeof(int), s));
  // 0: one launch per round
  for (int r = 0; r < 2; ++r)  // warm-up
    hipLaunchKernelGGL(k_probe_round, dim3(nblk), dim3(64), 0, s, a, b, nblk, r);
  PCK(reset());
  PCK(hipEventRecord(e0, s));
  for (int r = 0; r < rounds; ++r)
    hipLaunchKernelGGL(k_probe_round, dim3(nblk), dim3(64), 0, s, (r & 1) ? b : a,
                       (r & 1) ? a : b, nblk, r);
  PCK(hipEventRecord(e1, s));
  PCK(hipGetLastError());
  PCK(hipEventSynchronize(e1));
  PCK(hipEventElapsedTime(&ms, e0, e1));
  us_per_round[0] = 1e3 * ms / rounds;
  if ((rc = check("launches")) != SWM_OK) goto done;
  // 1: the same launches replayed from a HIP graph
  PCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  for (int r = 0; r < rounds; ++r)
    hipLaunchKernelGGL(k_probe_round, dim3(nblk), dim3(64), 0, s, (r & 1) ? b : a,
                       (r & 1) ? a : b, nblk, r);
  PCK(hipStreamEndCapture(s, &graph));
  PCK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  PCK(hipGraphLaunch(exec, s));  // warm-up
  PCK(reset());
  PCK(hipEventRecord(e0, s));
  PCK(hipGraphLaunch(exec, s));
  PCK(hipEventRecord(e1, s));
  PCK(hipEventSynchronize(e1));
  PCK(hipEventElapsedTime(&ms, e0, e1));
  us_per_round[1] = 1e3 * ms / rounds;
  if ((rc = check("graph")) != SWM_OK) goto done;
  // 2: one cooperative launch with a grid barrier per round
  {
    unsigned* count = ctr;
    unsigned* gen = ctr + 1;
    int rr = rounds;
    int nb = nblk;
    unsigned long long lim = spin_limit;
    void* args[] = {&a, &b, &nb, &rr, &count, &gen, &lim, &failed};
    PCK(reset());
    PCK(hipEventRecord(e0, s));
    PCK(hipLaunchCooperativeKernel((const void*)k_probe_barrier_rounds, dim3(nblk), dim3(64),
                                   args, 0, s));
    PCK(hipEventRecord(e1, s));
    PCK(hipEventSynchronize(e1));
    PCK(hipEventElapsedTime(&ms, e0, e1));
    PCK(hipMemcpy(&hfailed, failed, sizeof(int), hipMemcpyDeviceToHost));
    us_per_round[2] = 1e3 * ms / rounds;
    if (hfailed) {
      set_thread_error("swm_debug_sync_cost: a grid barrier wait hit its poll limit");
      rc = SWM_E_STATE;
    } else {
      rc = check("grid barrier");
    }
  }
done:
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (s) (void)hipStreamSynchronize(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (ctr) (void)hipFree(ctr);
  if (failed) (void)hipFree(failed);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}
