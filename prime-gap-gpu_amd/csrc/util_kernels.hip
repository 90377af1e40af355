// Small device kernels of the runtime: word fill (counter resets, usable as a
// graph node), verdict compaction (survivor lists for the extra Miller-Rabin
// bases, computed on the device with one atomic per wavefront), the verdicts
// of the survivors from their per-base pass masks, and gathers.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "runtime.hpp"

namespace pgx {
namespace {

__global__ __launch_bounds__(256) void k_fill(uint32_t* p, uint32_t n, uint32_t v) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = v;
}

// Every thread of the wavefront reaches the ballot (no early return), so the
// wavefront's survivors get consecutive slots from one atomicAdd.
__global__ __launch_bounds__(256) void k_compact(const uint32_t* __restrict__ verdict, uint32_t n,
                                                 uint32_t mask, uint32_t* __restrict__ index,
                                                 uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  bool keep = false;
  if (i < n) {
    const uint32_t v = verdict[i];
    keep = (v != PGX_VERDICT_INVALID) && (v & mask);
  }
  const uint64_t ballot = __ballot(keep);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t rank = __popcll(ballot & ((1ull << lane) - 1ull));
  uint32_t base = 0;
  if (lane == 0 && ballot) base = atomicAdd(count, (uint32_t)__popcll(ballot));
  base = __builtin_amdgcn_readfirstlane(base);
  if (keep) index[base + rank] = i;
}

__global__ __launch_bounds__(256) void k_gather(const uint32_t* __restrict__ src,
                                                const uint32_t* __restrict__ index, uint32_t n,
                                                uint32_t* __restrict__ dst) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[index[i]];
}

// Survivor k passed the first base; bit b of mask[k] says whether it passed
// base b + 1: the verdict counts the leading bases passed.
__global__ __launch_bounds__(256) void k_settle(uint32_t* __restrict__ verdict,
                                                const uint32_t* __restrict__ index,
                                                const uint32_t* __restrict__ count, uint32_t n,
                                                const uint32_t* __restrict__ mask) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < min(*count, n)) verdict[index[i]] = 1u + (uint32_t)__builtin_ctz(~mask[i]);
}

}  // namespace

hipError_t launch_settle(uint32_t* verdict, const uint32_t* index, const uint32_t* count,
                         uint32_t n, const uint32_t* mask, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_settle, dim3((n + 255) / 256), dim3(256), 0, s, verdict, index, count, n,
                     mask);
  return hipGetLastError();
}

hipError_t launch_gather(const uint32_t* src, const uint32_t* index, uint32_t n, uint32_t* dst,
                         hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((n + 255) / 256), dim3(256), 0, s, src, index, n, dst);
  return hipGetLastError();
}

hipError_t launch_fill(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s) {
  const uint32_t blocks = n == 0 ? 1 : (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
  hipLaunchKernelGGL(k_fill, dim3(blocks), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_compact(const uint32_t* verdict, uint32_t n, uint32_t mask, uint32_t* index,
                          uint32_t* count, hipStream_t s) {
  hipError_t e = launch_fill(count, 1, 0u, s);
  if (e != hipSuccess || n == 0) return e;
  hipLaunchKernelGGL(k_compact, dim3((n + 255) / 256), dim3(256), 0, s, verdict, n, mask, index,
                     count);
  return hipGetLastError();
}

}  // namespace pgx
