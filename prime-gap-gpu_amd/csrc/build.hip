// Device-side candidate construction: column i = m[i] * K + sign * x[i],
// written limb-interleaved (the layout every PRP kernel reads).
//
// The reference builds each candidate on the host with mpz_mul_ui +
// mpz_add_ui (fill_batch, gpu_testing.cu:171-172) and packs it with
// mpz_export (miller_rabin.h:373-378) before copying it to the GPU.  Here only
// K (once) and the multipliers (8 bytes per candidate) cross PCIe; one lane
// per candidate forms the product with 32x32->64 multiplies.  A candidate that
// does not fit `out_words` words, or would be negative, sets *status.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "runtime.hpp"

namespace {

// m_i = m64[i], or m_base + m_off[i] (a list of offsets into an m interval).
__global__ __launch_bounds__(256) void k_build(const uint32_t* __restrict__ K, uint32_t k_words,
                                               const uint64_t* __restrict__ m64,
                                               const uint32_t* __restrict__ m_off, uint64_t m_base,
                                               const uint32_t* __restrict__ x, uint32_t x0,
                                               int32_t sign, uint32_t n,
                                               uint32_t* __restrict__ out, uint32_t ld,
                                               uint32_t out_words, uint32_t* __restrict__ status) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t mm = m64 ? m64[i] : m_base + m_off[i];
  const uint32_t lo = (uint32_t)mm, hi = (uint32_t)(mm >> 32);
  const uint32_t xx = x ? x[i] : x0;
  uint64_t carry = 0;  // of the product m*K, < 2^34
  uint32_t cb = 0;     // carry (sign > 0) or borrow (sign < 0) of +-x
  uint32_t bad = 0;
  const uint32_t words = k_words + 2;  // m < 2^64 adds at most two words
  for (uint32_t j = 0; j < words || j < out_words; ++j) {
    // word j of m*K: K_j * lo + K_{j-1} * hi + carry (a 66-bit sum)
    const uint64_t a = (j < k_words) ? (uint64_t)K[j] * lo : 0;
    const uint64_t b = (j >= 1 && j - 1 < k_words) ? (uint64_t)K[j - 1] * hi : 0;
    const uint64_t s1 = a + carry;
    uint32_t ov = (s1 < a) ? 1u : 0u;
    const uint64_t s2 = s1 + b;
    ov += (s2 < s1) ? 1u : 0u;
    const uint32_t pw = (uint32_t)s2;
    carry = (s2 >> 32) | ((uint64_t)ov << 32);
    // +- x
    const uint32_t xw = (j == 0) ? xx : 0u;
    uint32_t w;
    if (sign > 0) {
      const uint64_t t = (uint64_t)pw + xw + cb;
      w = (uint32_t)t;
      cb = (uint32_t)(t >> 32);
    } else {
      const uint64_t t = (uint64_t)pw - xw - cb;
      w = (uint32_t)t;
      cb = (uint32_t)(t >> 63);
    }
    if (j < out_words) out[(size_t)j * ld + i] = w;
    else bad |= (w != 0u);
  }
  bad |= (carry != 0) | (cb != 0);  // too wide, or negative
  if (bad) atomicOr(status, 1u);
}

}  // namespace

hipError_t pgx::launch_build(const uint32_t* K, uint32_t k_words, const uint64_t* m64,
                             const uint32_t* m_off, uint64_t m_base, const uint32_t* x,
                             uint32_t x0, int32_t sign, uint32_t n, uint32_t* out, uint32_t ld,
                             uint32_t out_words, uint32_t* status, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_build, dim3((n + 255) / 256), dim3(256), 0, s, K, k_words, m64, m_off,
                     m_base, x, x0, sign, n, out, ld, out_words, status);
  return hipGetLastError();
}

extern "C" int pgx_build_candidates(pgx_runner_t* h, const uint32_t* d_K, uint32_t k_words,
                                    const uint64_t* d_m, const uint32_t* d_x, uint32_t x0,
                                    int32_t sign, uint32_t n, uint32_t* d_out, uint32_t ld,
                                    uint32_t out_words, uint32_t* d_status, void* stream) {
  if (!h) return PGX_E_ARG;
  if (!d_K || !d_m || !d_out || !d_status || k_words == 0 || out_words == 0 || ld < n ||
      (sign != 1 && sign != -1))
    return pgx::fail(h, PGX_E_ARG, "bad pgx_build_candidates arguments");
  if (n == 0) return PGX_OK;
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  PGX_HIP(h, hipSetDevice(h->device));
  PGX_HIP(h, pgx::launch_build(d_K, k_words, d_m, nullptr, 0, d_x, x0, sign, n, d_out, ld,
                                out_words, d_status, s));
  return PGX_OK;
}
