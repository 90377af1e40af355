// m-interval search on the device: the reference's sieve thread and GPU
// thread for one interval of multipliers m = m_start + m_i (m_i odd,
// 0 < m_i < m_inc), one X at a time, with every per-m set kept as a bitmap in
// HBM (bit t <-> m_i = 2t + 1, as the reference's composites vector).
//
//   reference (gap_search_gpu.cpp, gpu_testing.cu)     here
//   is_coprime_and_valid_m        :263-304             active bitmap A, set up at create
//   run_sieve_thread              :462-699             k_small_starts + k_sieve_tiles +
//                                                      k_sieve_large -> composites C
//   finalize (tests = active \ C) :711-735             k_count + k_scan + k_emit, ascending
//   fill_batch + run_test         gpu_testing.cu:136-195, miller_rabin.h:364-408
//                                                      launch_build + launch_prp per batch
//   process_finished_batch        gpu_testing.cu:104-133  k_consume -> found bitmap F + list
//   remove_vector                 :340-351             k_prune (A &= ~F)
//
// The tile kernel keeps a 4 KB slice of C in LDS (enough workgroups to fill
// the chip at 10^7 multipliers): primes below 64 are laid down as per-word
// patterns, primes below 4096 by one wavefront each walking their hits, the
// rest of the primes below 100000 by one lane each (LDS atomics).  Primes from 100000 up hit a slice at most a few times, so each
// lane takes one prime and marks the whole interval with global atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "runtime.hpp"

namespace {

constexpr uint32_t kSmallPrimeLimit = 100000;  // gap_search_gpu.cpp:495
constexpr uint32_t kTileWords = 1024;
constexpr uint32_t kTileBits = kTileWords * 32;
constexpr uint32_t kTinyPrime = 64;
constexpr uint32_t kCoopPrime = kTileBits / 8;
constexpr uint32_t kBlock = 256;
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kConsumeBlock = 1024;

// Odd primes below `limit` by a segmented sieve of Eratosthenes (the
// reference iterates primesieve, gap_search_gpu.cpp:481-499).
void odd_primes_below(uint32_t limit, std::vector<uint32_t>& out) {
  if (limit <= 3) return;
  const uint32_t root = (uint32_t)std::sqrt((double)limit) + 1;
  std::vector<uint8_t> comp(root + 1, 0);
  std::vector<uint32_t> base;
  for (uint32_t i = 3; i <= root; i += 2)
    if (!comp[i]) {
      base.push_back(i);
      for (uint64_t j = (uint64_t)i * i; j <= root; j += 2 * i) comp[j] = 1;
    }
  const uint64_t kSeg = 1u << 18;  // odd numbers per segment
  std::vector<uint8_t> seg(kSeg);
  out.reserve(out.size() + (size_t)(limit / std::max(1.0, std::log((double)limit)) * 1.2));
  for (uint64_t lo = 3; lo < limit; lo += 2 * kSeg) {
    std::fill(seg.begin(), seg.end(), 0);
    const uint64_t hi = std::min<uint64_t>(lo + 2 * kSeg, limit);
    for (uint32_t q : base) {
      const uint64_t q2 = (uint64_t)q * q;
      if (q2 >= hi) break;
      uint64_t s = std::max<uint64_t>(q2, (lo + q - 1) / q * q);
      if (!(s & 1)) s += q;
      for (uint64_t j = s; j < hi; j += 2ull * q) seg[(j - lo) >> 1] = 1;
    }
    for (uint64_t j = lo; j < hi; j += 2)
      if (!seg[(j - lo) >> 1]) out.push_back((uint32_t)j);
  }
}

// Per prime: -K^-1 mod p and m_start mod p (the reference's
// p_and_neg_inverse_k, gap_search_gpu.cpp:489-497, with _invert :438-458).
__global__ __launch_bounds__(kBlock) void k_prep(const uint32_t* __restrict__ K, uint32_t k_words,
                                                 const uint32_t* __restrict__ p, uint32_t n,
                                                 uint64_t m_start, uint32_t* __restrict__ neg_inv,
                                                 uint32_t* __restrict__ ms_mod) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t q = p[i];
  uint64_t r = 0;
  for (int j = (int)k_words - 1; j >= 0; --j) r = ((r << 32) | K[j]) % q;
  // extended Euclid on (q, r): s1 * r = gcd (mod q)
  int64_t r0 = q, r1 = (int64_t)r, s0 = 0, s1 = 1;
  while (r1) {
    const int64_t t = r0 / r1;
    int64_t u = r0 - t * r1;
    r0 = r1;
    r1 = u;
    u = s0 - t * s1;
    s0 = s1;
    s1 = u;
  }
  const uint32_t inv = (uint32_t)(s0 < 0 ? s0 + q : s0);
  neg_inv[i] = (r == 0) ? 0u : q - inv;  // r == 0 cannot happen for a sieving prime
  ms_mod[i] = (uint32_t)(m_start % q);
}

// v mod q for v < 2^52: the quotient from a double-precision product (off
// by at most one), then one correction; far cheaper than a 64-bit integer
// division on the device.
__device__ __forceinline__ uint32_t mod_u52(uint64_t v, uint32_t q) {
  const uint64_t g = (uint64_t)((double)v * (1.0 / (double)q));
  int64_t r = (int64_t)(v - g * q);
  if (r < 0) r += q;
  else if (r >= (int64_t)q) r -= q;
  return (uint32_t)r;
}

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds,
                                                         uint32_t* total) {
  const uint32_t tid = threadIdx.x, n = blockDim.x;
  lds[tid] = v;
  __syncthreads();
  for (uint32_t off = 1; off < n; off <<= 1) {
    const uint32_t add = (tid >= off) ? lds[tid - off] : 0u;
    __syncthreads();
    lds[tid] += add;
    __syncthreads();
  }
  const uint32_t incl = lds[tid];
  *total = lds[n - 1];
  __syncthreads();
  return incl - v;
}

// Primes in [lo, hi) (lo odd), one segment of kSegOdd odd numbers per
// workgroup: the base primes (odd, up to sqrt(hi)) mark their odd multiples
// in an LDS bitmap (one wavefront per base prime below 4096, one lane per
// larger one), then the unmarked numbers are appended to out[] with one
// atomic per workgroup.  The order of the list is not defined.
constexpr uint32_t kSegWords = 1024;
constexpr uint32_t kSegOdd = kSegWords * 32;

__global__ __launch_bounds__(kBlock) void k_gen_primes(const uint32_t* __restrict__ base,
                                                       uint32_t n_base, uint64_t lo, uint64_t hi,
                                                       uint32_t* __restrict__ out,
                                                       uint32_t* __restrict__ count) {
  __shared__ uint32_t seg[kSegWords];
  __shared__ uint32_t lds[kBlock];
  __shared__ uint32_t block_base;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const uint64_t seg_lo = lo + 2ull * kSegOdd * blockIdx.x;  // odd
  const uint64_t seg_hi = seg_lo + 2ull * kSegOdd;           // exclusive
  for (uint32_t w = tid; w < kSegWords; w += kBlock) seg[w] = 0;
  __syncthreads();
  auto first_index = [&](uint32_t q) -> uint64_t {  // odd-index of the first odd multiple to mark
    const uint64_t q2 = (uint64_t)q * q;
    uint64_t st = (seg_lo + q - 1) / q * q;
    if (st < q2) st = q2;
    if (!(st & 1)) st += q;
    return (st - seg_lo) >> 1;
  };
  for (uint32_t j = wave; j < n_base; j += kBlock / 64) {
    const uint32_t q = base[j];
    if (q >= 4096) break;  // the rest go one lane each
    if ((uint64_t)q * q >= seg_hi) break;
    for (uint64_t t = first_index(q) + (uint64_t)lane * q; t < kSegOdd; t += 64ull * q)
      atomicOr(&seg[t >> 5], 1u << (t & 31u));
  }
  for (uint32_t j = tid; j < n_base; j += kBlock) {
    const uint32_t q = base[j];
    if (q < 4096) continue;
    if ((uint64_t)q * q >= seg_hi) break;
    for (uint64_t t = first_index(q); t < kSegOdd; t += q)
      atomicOr(&seg[t >> 5], 1u << (t & 31u));
  }
  __syncthreads();
  constexpr uint32_t kPer = kSegWords / kBlock;  // words per thread
  uint32_t keep[kPer];
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    const uint32_t w = tid * kPer + k;
    uint32_t v = ~seg[w];
    const uint64_t v0 = seg_lo + 64ull * w;  // value of bit 0 of word w
    if (v0 >= hi) v = 0;
    else if (hi - v0 < 64) v &= (1u << ((hi - v0 + 1) / 2)) - 1u;  // values below hi
    keep[k] = v;
    cnt += __popc(v);
  }
  uint32_t total;
  uint32_t pos = block_exclusive_scan(cnt, lds, &total);
  if (tid == 0) block_base = total ? atomicAdd(count, total) : 0u;
  __syncthreads();
  pos += block_base;
#pragma unroll
  for (uint32_t k = 0; k < kPer; ++k) {
    uint32_t v = keep[k];
    const uint64_t v0 = seg_lo + 64ull * (tid * kPer + k);
    while (v) {
      const uint32_t b = __ffs(v) - 1u;
      v &= v - 1u;
      out[pos++] = (uint32_t)(v0 + 2ull * b);
    }
  }
}

// Smallest odd offset f with q | (m_start + f) K + X, halved: the first bit.
__device__ __forceinline__ uint32_t first_bit(uint32_t q, uint32_t nik, uint32_t msm, uint32_t X) {
  const uint64_t v = (uint64_t)X * nik + (q - msm);  // < 2^52 for X < 2^20
  uint32_t f = (v >> 52) ? (uint32_t)(v % q) : mod_u52(v, q);
  if (!(f & 1u)) f += q;
  return f >> 1;
}

__global__ __launch_bounds__(kBlock) void k_small_starts(const uint32_t* __restrict__ p,
                                                         const uint32_t* __restrict__ nik,
                                                         const uint32_t* __restrict__ msm,
                                                         uint32_t n, uint32_t X,
                                                         uint32_t* __restrict__ start) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) start[i] = first_bit(p[i], nik[i], msm[i], X);
}

// First hit at or above lo of the progression t0 + k q.
__device__ __forceinline__ uint32_t first_at_or_above(uint32_t t0, uint32_t q, uint32_t lo) {
  return (t0 >= lo) ? t0 : lo + (q - (lo - t0) % q) % q;
}

struct SmallPrimes {
  const uint32_t* p;
  const uint32_t* start;
  uint32_t n_tiny, n_coop, n_small;  // [0,n_tiny) tiny, [n_tiny,n_coop) coop, [n_coop,n_small)
};

__global__ __launch_bounds__(kBlock) void k_sieve_tiles(SmallPrimes sp, uint32_t half,
                                                        uint32_t words, uint32_t* __restrict__ C) {
  __shared__ uint32_t tile[kTileWords];
  const uint32_t tid = threadIdx.x;
  const uint32_t w0 = blockIdx.x * kTileWords;
  const uint32_t lo = w0 * 32u;
  const uint32_t hi = min(lo + kTileBits, half);  // lo < 2^30 + 2^17: no overflow
  for (uint32_t w = tid; w < kTileWords; w += kBlock) {
    const uint32_t wl = lo + 32u * w;
    uint32_t bits = 0;
    for (uint32_t j = 0; j < sp.n_tiny; ++j) {
      const uint32_t q = sp.p[j];
      for (uint32_t b = first_at_or_above(sp.start[j], q, wl) - wl; b < 32u; b += q) bits |= 1u << b;
    }
    if (wl >= half) bits = 0;
    else if (half - wl < 32u) bits &= (1u << (half - wl)) - 1u;
    tile[w] = bits;
  }
  __syncthreads();
  if (lo < half) {
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    for (uint32_t j = sp.n_tiny + wave; j < sp.n_coop; j += kBlock / 64) {
      const uint32_t q = sp.p[j];
      for (uint32_t t = first_at_or_above(sp.start[j], q, lo) + lane * q; t < hi; t += 64 * q)
        atomicOr(&tile[(t - lo) >> 5], 1u << (t & 31u));
    }
    for (uint32_t j = sp.n_coop + tid; j < sp.n_small; j += kBlock) {
      const uint32_t q = sp.p[j];
      for (uint32_t t = first_at_or_above(sp.start[j], q, lo); t < hi; t += q)
        atomicOr(&tile[(t - lo) >> 5], 1u << (t & 31u));
    }
  }
  __syncthreads();
  for (uint32_t w = tid; w < kTileWords && w0 + w < words; w += kBlock) C[w0 + w] = tile[w];
}

// Primes >= 100000: one lane per prime, global atomics.  as_written keeps the
// reference's arithmetic (gap_search_gpu.cpp:651-689): a 32-bit X * neg_inv_K
// and, when X * max_prime >= m_start, the full offset used as the bit index.
__global__ __launch_bounds__(kBlock) void k_sieve_large(const uint32_t* __restrict__ p,
                                                        const uint32_t* __restrict__ nik,
                                                        const uint32_t* __restrict__ msm,
                                                        uint32_t n, uint32_t X, uint64_t m_start,
                                                        uint64_t max_prime, uint32_t as_written,
                                                        uint32_t half, uint32_t* __restrict__ C) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t q = p[i];
  uint32_t t0;
  if (!as_written) {
    t0 = first_bit(q, nik[i], msm[i], X);
  } else {
    const uint32_t xk = X * nik[i];  // wraps mod 2^32 as written
    const uint32_t xkm = xk % q, ms = msm[i];
    uint32_t f;
    if ((uint64_t)X * max_prime < m_start) {
      f = (ms + q - xkm) % q;  // (m_start - xk) mod q, m_start > xk here
      f = f ? q - f : 0u;
      if (!(f & 1u)) f += q;
      t0 = f >> 1;
    } else {
      f = (xkm + (q - ms)) % q;
      if (!(f & 1u)) f += q;
      t0 = f;  // not halved
    }
  }
  for (uint32_t t = t0; t < half; t += q) atomicOr(&C[t >> 5], 1u << (t & 31u));
}

// Selection: bit set in A and in neither C nor F (C, F may be NULL).
__device__ __forceinline__ uint32_t select_word(const uint32_t* A, const uint32_t* C,
                                                const uint32_t* F, uint32_t w) {
  uint32_t v = A[w];
  if (C) v &= ~C[w];
  if (F) v &= ~F[w];
  return v;
}

__global__ __launch_bounds__(kBlock) void k_count(const uint32_t* __restrict__ A,
                                                  const uint32_t* __restrict__ C,
                                                  const uint32_t* __restrict__ F, uint32_t words,
                                                  uint32_t* __restrict__ block_count) {
  __shared__ uint32_t lds[kBlock];
  const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t c = (w < words) ? __popc(select_word(A, C, F, w)) : 0u;
  uint32_t total;
  block_exclusive_scan(c, lds, &total);
  if (threadIdx.x == 0) block_count[blockIdx.x] = total;
}

// Exclusive scan of nb block counts in place; *total = their sum.
__global__ __launch_bounds__(kScanThreads) void k_scan(uint32_t* __restrict__ counts, uint32_t nb,
                                                       uint32_t* __restrict__ total) {
  __shared__ uint32_t lds[kScanThreads];
  const uint32_t per = (nb + kScanThreads - 1) / kScanThreads;
  const uint32_t b0 = threadIdx.x * per, b1 = min(b0 + per, nb);
  uint32_t s = 0;
  for (uint32_t b = b0; b < b1; ++b) s += counts[b];
  uint32_t sum;
  uint32_t run = block_exclusive_scan(s, lds, &sum);
  for (uint32_t b = b0; b < b1; ++b) {
    const uint32_t c = counts[b];
    counts[b] = run;
    run += c;
  }
  if (threadIdx.x == 0) *total = sum;
}

__global__ __launch_bounds__(kBlock) void k_emit(const uint32_t* __restrict__ A,
                                                 const uint32_t* __restrict__ C,
                                                 const uint32_t* __restrict__ F, uint32_t words,
                                                 const uint32_t* __restrict__ block_off,
                                                 uint32_t* __restrict__ out) {
  __shared__ uint32_t lds[kBlock];
  const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
  uint32_t v = (w < words) ? select_word(A, C, F, w) : 0u;
  uint32_t total;
  uint32_t pos = block_off[blockIdx.x] + block_exclusive_scan(__popc(v), lds, &total);
  while (v) {
    const uint32_t b = __ffs(v) - 1u;
    v &= v - 1u;
    out[pos++] = 2u * (32u * w + b) + 1u;
  }
}

// Verdicts of one batch: m_i passing every base joins F and the found list.
// One atomic per workgroup reserves the list slots (a single counter taking
// one atomic per wavefront serialises at this rate of hits).
__global__ __launch_bounds__(kConsumeBlock) void k_consume(
    const uint32_t* __restrict__ verdict, const uint32_t* __restrict__ tests, uint32_t n,
    uint32_t rounds, uint32_t X, uint32_t* __restrict__ F, uint32_t* __restrict__ found_mi,
    uint32_t* __restrict__ found_x, uint32_t* __restrict__ found_count) {
  __shared__ uint32_t wave_count[kConsumeBlock / 64];
  __shared__ uint32_t block_base;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool hit = (i < n) && verdict[i] == rounds;
  const uint64_t ballot = __ballot(hit);
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  if (lane == 0) wave_count[wave] = (uint32_t)__popcll(ballot);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t total = 0;
    for (uint32_t w = 0; w < kConsumeBlock / 64; ++w) {
      const uint32_t c = wave_count[w];
      wave_count[w] = total;
      total += c;
    }
    block_base = total ? atomicAdd(found_count, total) : 0u;
  }
  __syncthreads();
  if (hit) {
    const uint32_t mi = tests[i], t = mi >> 1;
    atomicOr(&F[t >> 5], 1u << (t & 31u));
    const uint32_t pos = block_base + wave_count[wave] +
                         (uint32_t)__popcll(ballot & ((1ull << lane) - 1ull));
    found_mi[pos] = mi;
    found_x[pos] = X;
  }
}

__global__ __launch_bounds__(kBlock) void k_prune(uint32_t* __restrict__ A,
                                                  const uint32_t* __restrict__ F, uint32_t words) {
  const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
  if (w < words) A[w] &= ~F[w];
}

// Active multipliers: odd m_i with gcd(m_start + m_i, d) = 1
// (is_coprime_and_valid_m, gap_search_gpu.cpp:263-304): each odd prime of d
// removes the bits t0, t0 + q, ... (t0 = first odd multiple offset, halved).
__global__ __launch_bounds__(kBlock) void k_active(const uint32_t* __restrict__ dp,
                                                   const uint32_t* __restrict__ dp_t0,
                                                   uint32_t n_dp, uint32_t half, uint32_t words,
                                                   uint32_t* __restrict__ A) {
  const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
  if (w >= words) return;
  const uint32_t wl = 32u * w;
  uint32_t bad = 0;
  for (uint32_t j = 0; j < n_dp; ++j) {
    const uint32_t q = dp[j];
    for (uint32_t b = first_at_or_above(dp_t0[j], q, wl) - wl; b < 32u; b += q) bad |= 1u << b;
  }
  uint32_t bits = ~bad;
  if (wl >= half) bits = 0;
  else if (half - wl < 32u) bits &= (1u << (half - wl)) - 1u;
  A[w] = bits;
}

}  // namespace

struct pgx_search {
  pgx_runner* runner = nullptr;
  pgx_search_config_t cfg{};
  uint32_t half = 0, words = 0, n_blocks = 0;
  uint32_t k_words = 0, max_bits = 0;
  uint32_t n_tiny = 0, n_coop = 0, n_small = 0, n_primes = 0;
  uint32_t* d_K = nullptr;
  uint32_t* d_p = nullptr;      // sieving primes, ascending
  uint32_t* d_nik = nullptr;    // -K^-1 mod p
  uint32_t* d_msm = nullptr;    // m_start mod p
  uint32_t* d_start = nullptr;  // first bit of each prime below 100000 for the current X
  uint32_t* d_dp = nullptr;     // odd primes dividing d, then their first bits
  uint32_t n_dp = 0;
  uint32_t* d_A = nullptr;
  uint32_t* d_C = nullptr;
  uint32_t* d_F = nullptr;
  uint32_t* d_block = nullptr;  // per-block counts / offsets of the selection
  uint32_t* d_tests = nullptr;
  uint32_t* d_found_mi = nullptr;
  uint32_t* d_found_x = nullptr;
  uint32_t* d_words = nullptr;  // [0] tests, [1] found, [2] build status, [3] active count
  uint32_t* h_words = nullptr;  // pinned mirror of d_words
  size_t cap = 0;               // active multipliers at create: bound of every list
  uint32_t cur_x = 0;
  uint32_t n_tests = 0;
  bool sieved = false;
  std::string err;
};

namespace {

int sfail(pgx_search* s, int code, const char* msg) {
  s->err = msg;
  return code;
}

#define PGX_SHIP(s, call)                                                            \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      (s)->err = std::string(#call " failed: ") + hipGetErrorString(e_);             \
      return PGX_E_HIP;                                                              \
    }                                                                                \
  } while (0)

void release(pgx_search* s) {
  for (uint32_t* p : {s->d_K, s->d_p, s->d_nik, s->d_msm, s->d_start, s->d_dp, s->d_A, s->d_C,
                      s->d_F, s->d_block, s->d_tests, s->d_found_mi, s->d_found_x, s->d_words})
    if (p) (void)hipFree(p);
  if (s->h_words) (void)hipHostFree(s->h_words);
}

// Ordered list of the bits of A & ~C & ~F into out; count to d_words[slot].
int select_list(pgx_search* s, const uint32_t* C, const uint32_t* F, uint32_t* out,
                uint32_t slot) {
  hipStream_t st = s->runner->stream;
  hipLaunchKernelGGL(k_count, dim3(s->n_blocks), dim3(kBlock), 0, st, s->d_A, C, F, s->words,
                     s->d_block);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, s->d_block, s->n_blocks,
                     s->d_words + slot);
  hipLaunchKernelGGL(k_emit, dim3(s->n_blocks), dim3(kBlock), 0, st, s->d_A, C, F, s->words,
                     s->d_block, out);
  PGX_SHIP(s, hipGetLastError());
  return PGX_OK;
}

int sync_words(pgx_search* s) {
  PGX_SHIP(s, hipMemcpyAsync(s->h_words, s->d_words, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             s->runner->stream));
  PGX_SHIP(s, hipStreamSynchronize(s->runner->stream));
  return PGX_OK;
}

}  // namespace

extern "C" {

int pgx_search_create(pgx_search_t** out, pgx_runner_t* h, const uint32_t* K, uint32_t k_words,
                      const pgx_search_config_t* cfg) {
  if (!out || !h || !K || !k_words || !cfg) return PGX_E_ARG;
  *out = nullptr;
  if (!h->inflight.empty()) return pgx::fail(h, PGX_E_STATE, "the runner has a batch in flight");
  const pgx_search_config_t c = *cfg;
  // the reference's own checks (gap_search_gpu.cpp:502-504, gap_common.cpp:431-446)
  if ((c.d & 1u) || c.d == 0 || (c.m_start & 1u) || (c.m_inc & 1u) || c.m_inc < 2 ||
      c.m_inc >= 0x7FFFFFFFu || c.max_prime < 1000 || c.max_prime > 2000000000u ||
      (c.flags & ~PGX_SEARCH_AS_WRITTEN))
    return pgx::fail(h, PGX_E_ARG, "bad search configuration");
  if (!(K[0] & 1u)) return pgx::fail(h, PGX_E_ARG, "K must be odd");
  if (c.P >= kSmallPrimeLimit) return pgx::fail(h, PGX_E_ARG, "P must be below 100000");
  uint32_t kb = 0;
  for (int j = (int)k_words - 1; j >= 0; --j)
    if (K[j]) {
      kb = 32u * j + 32u - __builtin_clz(K[j]);
      break;
    }
  const uint64_t m_last = c.m_start + c.m_inc;
  const uint32_t mb = 64u - __builtin_clzll(m_last);
  const uint32_t max_bits = kb + mb + 1;
  if (max_bits > 32u * h->limbs)
    return pgx::fail(h, PGX_E_INPUT, "m*K + X reaches %u bits: more than the runner's %u words",
                     max_bits, h->limbs);
  if (max_bits > PGX_MAX_BITS)
    return pgx::fail(h, PGX_E_INPUT, "m*K + X reaches %u bits, more than %d", max_bits,
                     PGX_MAX_BITS);

  pgx_search* s = new (std::nothrow) pgx_search;
  if (!s) return pgx::fail(h, PGX_E_HIP, "out of host memory");
  s->runner = h;
  s->cfg = c;
  s->half = c.m_inc / 2;
  s->words = s->half / 32 + 2;  // with the reference's sentinel words
  s->n_blocks = (s->words + kBlock - 1) / kBlock;
  s->k_words = k_words;
  s->max_bits = max_bits;

  // sieving primes: odd q < max_prime with q > P, or q | d (the d wheel).  The
  // host lists those below 100000 (ascending: the tile kernel's classes);
  // the device generates the rest (k_gen_primes, unordered).
  std::vector<uint32_t> all, primes, dps, base;
  odd_primes_below(std::min<uint32_t>(c.max_prime, kSmallPrimeLimit), all);
  for (uint32_t q : all) {
    if (q <= c.P) {
      if (c.d % q == 0) primes.push_back(q);
      continue;
    }
    primes.push_back(q);
  }
  {
    std::vector<uint32_t> upto_p;
    odd_primes_below(c.P + 1, upto_p);
    for (uint32_t q : upto_p)
      if (c.d % q == 0) dps.push_back(q);
    const size_t nd = dps.size();
    for (size_t j = 0; j < nd; ++j) {  // first bit of the multiples of each prime of d
      const uint32_t q = dps[j];
      uint64_t first = (q - c.m_start % q) % q;
      if (!(first & 1)) first += q;
      dps.push_back((uint32_t)(first >> 1));
    }
  }
  std::vector<uint32_t>().swap(all);
  s->n_small = (uint32_t)primes.size();
  s->n_tiny = (uint32_t)(std::lower_bound(primes.begin(), primes.end(), kTinyPrime) - primes.begin());
  s->n_coop = (uint32_t)(std::lower_bound(primes.begin(), primes.end(), kCoopPrime) - primes.begin());
  s->n_dp = (uint32_t)(dps.size() / 2);
  size_t large_cap = 0;  // bound on the primes in [100001, max_prime): pi(x) < 1.25506 x / ln x
  if (c.max_prime > kSmallPrimeLimit) {
    const double x = (double)c.max_prime;
    large_cap = (size_t)(1.25506 * x / std::log(x)) + 64;
    odd_primes_below((uint32_t)std::sqrt(x) + 2, base);
  }
  int rc = PGX_OK;
  auto dev = [&](uint32_t** p, size_t nwords) -> bool {
    hipError_t e = hipMalloc((void**)p, std::max<size_t>(nwords, 1) * 4);
    if (e != hipSuccess) {
      rc = pgx::fail(h, PGX_E_HIP, "hipMalloc of %zu bytes failed: %s", nwords * 4,
                     hipGetErrorString(e));
      return false;
    }
    return true;
  };
  hipStream_t st = h->stream;
  hipError_t e = hipSetDevice(h->device);
  const size_t p_cap = s->n_small + large_cap;
  uint32_t* d_base = nullptr;
  bool ok = e == hipSuccess && dev(&s->d_K, k_words) && dev(&s->d_p, p_cap) &&
            dev(&s->d_nik, p_cap) && dev(&s->d_msm, p_cap) && dev(&d_base, base.size()) &&
            dev(&s->d_start, s->n_small) && dev(&s->d_dp, 2 * s->n_dp) && dev(&s->d_A, s->words) &&
            dev(&s->d_C, s->words) && dev(&s->d_F, s->words) && dev(&s->d_block, s->n_blocks) &&
            dev(&s->d_words, 4);
  if (ok) {
    e = hipHostMalloc((void**)&s->h_words, 4 * sizeof(uint32_t), hipHostMallocDefault);
    ok = e == hipSuccess;
  }
  if (ok) e = hipMemcpy(s->d_K, K, k_words * 4, hipMemcpyHostToDevice);
  if (ok && e == hipSuccess && s->n_small)
    e = hipMemcpy(s->d_p, primes.data(), primes.size() * 4, hipMemcpyHostToDevice);
  if (ok && e == hipSuccess && !base.empty())
    e = hipMemcpy(d_base, base.data(), base.size() * 4, hipMemcpyHostToDevice);
  if (ok && e == hipSuccess && s->n_dp)
    e = hipMemcpy(s->d_dp, dps.data(), dps.size() * 4, hipMemcpyHostToDevice);
  if (ok && e == hipSuccess) e = hipMemsetAsync(s->d_F, 0, s->words * 4, st);
  if (ok && e == hipSuccess) e = hipMemsetAsync(s->d_C, 0, s->words * 4, st);
  if (ok && e == hipSuccess) e = hipMemsetAsync(s->d_words, 0, 16, st);
  s->n_primes = s->n_small;
  if (ok && e == hipSuccess && large_cap) {  // primes in [100001, max_prime) on the device
    const uint64_t lo = kSmallPrimeLimit + 1, hi = c.max_prime;
    const uint64_t n_seg = (hi - lo + 2ull * kSegOdd - 1) / (2ull * kSegOdd);
    hipLaunchKernelGGL(k_gen_primes, dim3((uint32_t)n_seg), dim3(kBlock), 0, st, d_base,
                       (uint32_t)base.size(), lo, hi, s->d_p + s->n_small, s->d_words + 3);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(s->h_words, s->d_words, 16, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess) {
      if (s->h_words[3] > large_cap) {
        rc = pgx::fail(h, PGX_E_HIP, "prime count %u above its bound %zu", s->h_words[3],
                       large_cap);
        ok = false;
      }
      s->n_primes = s->n_small + s->h_words[3];
    }
    if (e == hipSuccess) e = hipMemsetAsync(s->d_words, 0, 16, st);
  }
  if (d_base) (void)hipFree(d_base);
  if (ok && e == hipSuccess && s->n_primes) {
    hipLaunchKernelGGL(k_prep, dim3((s->n_primes + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       s->d_K, k_words, s->d_p, s->n_primes, c.m_start, s->d_nik, s->d_msm);
    e = hipGetLastError();
  }
  if (ok && e == hipSuccess) {
    hipLaunchKernelGGL(k_active, dim3(s->n_blocks), dim3(kBlock), 0, st, s->d_dp,
                       s->d_dp + s->n_dp, s->n_dp, s->half, s->words, s->d_A);
    e = hipGetLastError();
  }
  if (ok && e != hipSuccess) {
    rc = pgx::fail(h, PGX_E_HIP, "search set-up failed: %s", hipGetErrorString(e));
    ok = false;
  }
  if (ok) {  // list capacity: the active count
    hipLaunchKernelGGL(k_count, dim3(s->n_blocks), dim3(kBlock), 0, st, s->d_A, nullptr, nullptr,
                       s->words, s->d_block);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(kScanThreads), 0, st, s->d_block, s->n_blocks,
                       s->d_words + 3);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(s->h_words, s->d_words, 16, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rc = pgx::fail(h, PGX_E_HIP, "search set-up failed: %s", hipGetErrorString(e));
      ok = false;
    }
  }
  if (ok) {
    s->cap = s->h_words[3];
    ok = dev(&s->d_tests, s->cap) && dev(&s->d_found_mi, s->cap) && dev(&s->d_found_x, s->cap);
  }
  if (!ok) {
    release(s);
    delete s;
    return rc ? rc : pgx::fail(h, PGX_E_HIP, "search set-up failed");
  }
  *out = s;
  return PGX_OK;
}

int pgx_search_destroy(pgx_search_t* s) {
  if (!s) return PGX_OK;
  (void)hipStreamSynchronize(s->runner->stream);
  release(s);
  delete s;
  return PGX_OK;
}

int pgx_search_sieve(pgx_search_t* s, uint32_t X, uint32_t* n_tests) {
  if (!s) return PGX_E_ARG;
  if (X == 0 || (X & 1u)) return sfail(s, PGX_E_ARG, "X must be even and positive");
  if (!s->runner->inflight.empty()) return sfail(s, PGX_E_STATE, "the runner has a batch in flight");
  PGX_SHIP(s, hipSetDevice(s->runner->device));
  hipStream_t st = s->runner->stream;
  if (s->n_small) {
    hipLaunchKernelGGL(k_small_starts, dim3((s->n_small + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       st, s->d_p, s->d_nik, s->d_msm, s->n_small, X, s->d_start);
  }
  SmallPrimes sp{s->d_p, s->d_start, s->n_tiny, s->n_coop, s->n_small};
  hipLaunchKernelGGL(k_sieve_tiles, dim3((s->words + kTileWords - 1) / kTileWords), dim3(kBlock),
                     0, st, sp, s->half, s->words, s->d_C);
  const uint32_t n_large = s->n_primes - s->n_small;
  if (n_large) {
    hipLaunchKernelGGL(k_sieve_large, dim3((n_large + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       s->d_p + s->n_small, s->d_nik + s->n_small, s->d_msm + s->n_small, n_large,
                       X, s->cfg.m_start, (uint64_t)s->cfg.max_prime,
                       (s->cfg.flags & PGX_SEARCH_AS_WRITTEN) ? 1u : 0u, s->half, s->d_C);
  }
  PGX_SHIP(s, hipGetLastError());
  int rc = select_list(s, s->d_C, s->d_F, s->d_tests, 0);
  if (rc != PGX_OK) return rc;
  rc = sync_words(s);
  if (rc != PGX_OK) return rc;
  s->n_tests = s->h_words[0];
  s->cur_x = X;
  s->sieved = true;
  if (n_tests) *n_tests = s->n_tests;
  return PGX_OK;
}

int pgx_search_test(pgx_search_t* s, uint32_t* n_found) {
  if (!s) return PGX_E_ARG;
  if (!s->sieved) return sfail(s, PGX_E_STATE, "no sieved X to test");
  pgx_runner* h = s->runner;
  if (!h->inflight.empty()) return sfail(s, PGX_E_STATE, "the runner has a batch in flight");
  PGX_SHIP(s, hipSetDevice(h->device));
  hipStream_t st = h->stream;
  const uint32_t found_before = s->h_words[1];
  const uint32_t batch = (uint32_t)std::min<size_t>(h->max_n, 0xFFFFFFFFu);
  for (uint32_t off = 0; off < s->n_tests; off += batch) {
    const uint32_t nb = std::min(batch, s->n_tests - off);
    PGX_SHIP(s, pgx::launch_build(s->d_K, s->k_words, nullptr, s->d_tests + off,
                                  s->cfg.m_start, nullptr, s->cur_x, 1, nb, h->b0.d_cand, nb,
                                  h->limbs, s->d_words + 2, st));
    pgx::PrpArgs pa{};
    pa.cand = h->b0.d_cand;
    pa.ld = nb;
    pa.n = nb;
    pa.buf_limbs = h->limbs;
    pa.mode = PGX_MODE_REFERENCE;
    // several bases: the first here, the others side by side on its survivors
    pa.bases = pgx::first_base(h->bases);
    pa.verdict = h->b0.d_verdict;
    int rc = pgx::launch_auto(h, pa, s->max_bits, st);
    if (rc == PGX_OK) rc = pgx::launch_later_bases(h, h->b0, nb, s->max_bits, PGX_MODE_REFERENCE, st);
    if (rc != PGX_OK) return sfail(s, rc, h->err.c_str());
    hipLaunchKernelGGL(k_consume, dim3((nb + kConsumeBlock - 1) / kConsumeBlock),
                       dim3(kConsumeBlock), 0, st,
                       h->b0.d_verdict, s->d_tests + off, nb, (uint32_t)h->bases.count, s->cur_x,
                       s->d_F, s->d_found_mi, s->d_found_x, s->d_words + 1);
    PGX_SHIP(s, hipGetLastError());
  }
  hipLaunchKernelGGL(k_prune, dim3(s->n_blocks), dim3(kBlock), 0, st, s->d_A, s->d_F, s->words);
  PGX_SHIP(s, hipGetLastError());
  int rc = sync_words(s);
  if (rc != PGX_OK) return rc;
  s->sieved = false;
  if (s->h_words[2]) return sfail(s, PGX_E_INPUT, "a candidate did not fit the runner's words");
  if (n_found) *n_found = s->h_words[1] - found_before;
  return PGX_OK;
}

int pgx_search_step(pgx_search_t* s, uint32_t X, uint32_t* n_tests, uint32_t* n_found) {
  int rc = pgx_search_sieve(s, X, n_tests);
  if (rc != PGX_OK) return rc;
  return pgx_search_test(s, n_found);
}

int pgx_search_tests(pgx_search_t* s, uint32_t* out, size_t cap, size_t* n) {
  if (!s || !n) return PGX_E_ARG;
  *n = s->n_tests;
  if (!out) return PGX_OK;
  if (cap < s->n_tests) return sfail(s, PGX_E_CAPACITY, "output smaller than the tests list");
  PGX_SHIP(s, hipStreamSynchronize(s->runner->stream));
  if (s->n_tests)
    PGX_SHIP(s, hipMemcpy(out, s->d_tests, (size_t)s->n_tests * 4, hipMemcpyDeviceToHost));
  return PGX_OK;
}

int pgx_search_found(pgx_search_t* s, uint32_t* m_i, uint32_t* x, size_t cap, size_t* n) {
  if (!s || !n) return PGX_E_ARG;
  int rc = sync_words(s);
  if (rc != PGX_OK) return rc;
  const size_t cnt = s->h_words[1];
  *n = cnt;
  if (!m_i && !x) return PGX_OK;
  if (cap < cnt) return sfail(s, PGX_E_CAPACITY, "output smaller than the found list");
  std::vector<uint32_t> a(cnt), b(cnt);
  if (cnt) {
    PGX_SHIP(s, hipMemcpy(a.data(), s->d_found_mi, cnt * 4, hipMemcpyDeviceToHost));
    PGX_SHIP(s, hipMemcpy(b.data(), s->d_found_x, cnt * 4, hipMemcpyDeviceToHost));
  }
  std::vector<uint32_t> order(cnt);
  for (size_t i = 0; i < cnt; ++i) order[i] = (uint32_t)i;
  std::sort(order.begin(), order.end(), [&](uint32_t u, uint32_t v) { return a[u] < a[v]; });
  for (size_t i = 0; i < cnt; ++i) {
    if (m_i) m_i[i] = a[order[i]];
    if (x) x[i] = b[order[i]];
  }
  return PGX_OK;
}

int pgx_search_bitmap(pgx_search_t* s, uint32_t which, uint32_t* out, size_t cap, size_t* n) {
  if (!s || !n) return PGX_E_ARG;
  const uint32_t* src = which == PGX_SEARCH_ACTIVE       ? s->d_A
                        : which == PGX_SEARCH_COMPOSITES ? s->d_C
                        : which == PGX_SEARCH_FOUND      ? s->d_F
                                                         : nullptr;
  if (!src) return sfail(s, PGX_E_ARG, "unknown bitmap");
  *n = s->words;
  if (!out) return PGX_OK;
  if (cap < s->words) return sfail(s, PGX_E_CAPACITY, "output smaller than the bitmap");
  PGX_SHIP(s, hipStreamSynchronize(s->runner->stream));
  PGX_SHIP(s, hipMemcpy(out, src, (size_t)s->words * 4, hipMemcpyDeviceToHost));
  return PGX_OK;
}

const char* pgx_search_last_error(pgx_search_t* s) {
  return s ? s->err.c_str() : "no search";
}

}  // extern "C"
