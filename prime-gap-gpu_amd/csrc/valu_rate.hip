// VALU issue-rate microbenchmark: the denominator of the INT32 multiply-add
// roofline.  Each lane runs 8 independent dependency chains of one
// instruction kind; the kernel's duration (HIP events) over the lane-op count
// gives lane-ops per second for that instruction on the whole device.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/pgx.h"
#include "mont.hpp"

namespace {

constexpr int kChains = 8;

template <int OP>
__global__ __launch_bounds__(256) void k_rate(uint32_t iters, uint32_t seed, uint32_t* out) {
  const uint32_t a = threadIdx.x * 2654435761u + seed;
  const uint32_t b = blockIdx.x * 40503u + 17u;
  uint64_t acc[kChains];
  uint32_t hi[kChains];
  double f[kChains];
#pragma unroll
  for (int k = 0; k < kChains; ++k) {
    acc[k] = a + k;
    hi[k] = b + k;
    f[k] = (double)(a + k);
  }
  const double fa = 1.0000001, fb = 0.5;
  for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < kChains; ++k) {
      if constexpr (OP == 0) {  // v_mad_u64_u32 alone (carry-out discarded)
        uint64_t c;
        asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(acc[k]), "=s"(c) : "v"(a), "v"(b));
      } else if constexpr (OP == 1) {  // one dependent multiply-accumulate chain per lane
        acc[0] += pgx::mul64(a + (uint32_t)acc[0], b);
      } else if constexpr (OP == 2) {  // full-rate integer add
        asm volatile("v_add_u32 %0, %1, %0" : "+v"(hi[k]) : "v"(a));
      } else if constexpr (OP == 3) {  // v_mul_lo_u32
        asm volatile("v_mul_lo_u32 %0, %1, %0" : "+v"(hi[k]) : "v"(a));
      } else {  // v_fma_f64
        asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(f[k]) : "v"(fa), "v"(fb));
      }
    }
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kChains; ++k) s ^= (uint32_t)acc[k] ^ (uint32_t)(acc[k] >> 32) ^ hi[k] ^
                                        (uint32_t)(int64_t)f[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <int OP>
hipError_t launch(uint32_t blocks, uint32_t iters, uint32_t* out, hipStream_t s) {
  hipLaunchKernelGGL(k_rate<OP>, dim3(blocks), dim3(256), 0, s, iters, 7u, out);
  return hipGetLastError();
}

}  // namespace

extern "C" {

/* Lane-instructions per second of instruction kind `op` on `device`:
 * 0 v_mad_u64_u32 (8 independent chains per lane: issue rate), 1 v_mad_u64_u32
 * in one dependent chain per lane (latency bound at low occupancy),
 * 2 v_add_u32, 3 v_mul_lo_u32, 4 v_fma_f64.  `blocks` of 256 threads run
 * `iters` x 8 instructions per lane.  Result in *per_sec (for op 1: MACs/s). */
int pgx_valu_rate(int device, int op, uint32_t blocks, uint32_t iters, double* per_sec) {
  if (!per_sec || op < 0 || op > 4 || blocks == 0 || iters == 0) return PGX_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return PGX_E_NODEV;
  uint32_t* out = nullptr;
  hipEvent_t e0, e1;
  if (hipMalloc((void**)&out, (size_t)blocks * 256 * 4) != hipSuccess) return PGX_E_HIP;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  hipError_t err = hipSuccess;
  auto go = [&](uint32_t it) {
    switch (op) {
      case 0: return launch<0>(blocks, it, out, nullptr);
      case 1: return launch<1>(blocks, it, out, nullptr);
      case 2: return launch<2>(blocks, it, out, nullptr);
      case 3: return launch<3>(blocks, it, out, nullptr);
      default: return launch<4>(blocks, it, out, nullptr);
    }
  };
  err = go(iters / 8 + 1);  // warm-up
  if (err == hipSuccess) err = hipDeviceSynchronize();
  if (err == hipSuccess) err = hipEventRecord(e0, nullptr);
  if (err == hipSuccess) err = go(iters);
  if (err == hipSuccess) err = hipEventRecord(e1, nullptr);
  if (err == hipSuccess) err = hipEventSynchronize(e1);
  float ms = 0;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  if (err != hipSuccess || ms <= 0) return PGX_E_HIP;
  *per_sec = (double)blocks * 256.0 * iters * kChains / (ms * 1e-3);
  return PGX_OK;
}

}  // extern "C"
