// Host-side builder of sieved candidates N = m * K + X, K = P# / d.
//
// Reference roles: init_K (gap_common.cpp:60-65: mpz_primorial_ui then
// mpz_tdiv_q_ui), the X sieve (gap_search_gpu.cpp:586-689: an offset X is
// dropped when some sieving prime p divides m*K + X, found from
// (m mod p)(K mod p)) and fill_batch (gpu_testing.cu:136-195: N = m*K + x
// written for every surviving m at one x).  Here the sieve runs per m over an
// X interval and the survivors are written as little-endian 32-bit limbs, the
// layout from_mpz produces (miller_rabin.h:43-54).  Plain C++, no GMP.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/pgx.h"

namespace {

std::vector<uint32_t> primes_below(uint32_t limit) {
  std::vector<uint32_t> out;
  if (limit < 3) return out;
  std::vector<uint8_t> comp(limit, 0);
  for (uint32_t p = 2; p < limit; ++p) {
    if (comp[p]) continue;
    out.push_back(p);
    for (uint64_t q = (uint64_t)p * p; q < limit; q += p) comp[q] = 1;
  }
  return out;
}

// x <- x * f (+ add), little-endian limbs, growing as needed.
void mul_small(std::vector<uint32_t>& x, uint64_t f, uint32_t add = 0) {
  // f may be up to 64 bits: multiply by the two halves
  const uint32_t lo = (uint32_t)f, hi = (uint32_t)(f >> 32);
  std::vector<uint32_t> r(x.size() + 2, 0);
  for (int half = 0; half < 2; ++half) {
    const uint32_t g = half ? hi : lo;
    if (!g) continue;
    uint64_t carry = 0;
    for (size_t j = 0; j < x.size(); ++j) {
      uint64_t v = (uint64_t)x[j] * g + r[j + half] + carry;
      r[j + half] = (uint32_t)v;
      carry = v >> 32;
    }
    for (size_t j = x.size() + half; carry; ++j) {
      uint64_t v = (uint64_t)r[j] + carry;
      r[j] = (uint32_t)v;
      carry = v >> 32;
    }
  }
  uint64_t carry = add;
  for (size_t j = 0; carry && j < r.size(); ++j) {
    uint64_t v = (uint64_t)r[j] + carry;
    r[j] = (uint32_t)v;
    carry = v >> 32;
  }
  while (r.size() > 1 && r.back() == 0) r.pop_back();
  x.swap(r);
}

// x <- x / d, returns the remainder.
uint32_t div_small(std::vector<uint32_t>& x, uint32_t d) {
  uint64_t rem = 0;
  for (size_t j = x.size(); j-- > 0;) {
    uint64_t cur = (rem << 32) | x[j];
    x[j] = (uint32_t)(cur / d);
    rem = cur % d;
  }
  while (x.size() > 1 && x.back() == 0) x.pop_back();
  return (uint32_t)rem;
}

// x <- x - v; false if that would go negative.
bool sub_small(std::vector<uint32_t>& x, uint32_t v) {
  uint64_t borrow = v;
  for (size_t j = 0; j < x.size() && borrow; ++j) {
    const uint64_t cur = x[j];
    x[j] = (uint32_t)(cur - borrow);
    borrow = cur < borrow ? 1 : 0;
  }
  if (borrow) return false;
  while (x.size() > 1 && x.back() == 0) x.pop_back();
  return true;
}

uint32_t mod_small(const std::vector<uint32_t>& x, uint32_t p) {
  uint64_t rem = 0;
  for (size_t j = x.size(); j-- > 0;) rem = ((rem << 32) | x[j]) % p;
  return (uint32_t)rem;
}

uint32_t bits_of(const std::vector<uint32_t>& x) {
  for (size_t j = x.size(); j-- > 0;)
    if (x[j]) return 32u * (uint32_t)j + 32u - __builtin_clz(x[j]);
  return 0;
}

// K = P# / d; false if d does not divide P#.
bool primorial_over(uint32_t P, uint32_t d, std::vector<uint32_t>& K) {
  K.assign(1, 1);
  for (uint32_t p : primes_below(P + 1)) mul_small(K, p);
  if (d == 0) return false;
  return div_small(K, d) == 0;
}

// splitmix64 / xoshiro256** (Blackman & Vigna), the generator's PRNG.
struct Xoshiro {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
  explicit Xoshiro(uint64_t seed) {
    for (auto& w : s) w = splitmix(seed);
  }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // uniform in [0, r): draws below 2^64 mod r are rejected
  uint64_t below(uint64_t r) {
    const uint64_t reject = (0 - r) % r;
    for (;;) {
      uint64_t v = next();
      if (v >= reject) return v % r;
    }
  }
};

uint64_t gcd64(uint64_t a, uint64_t b) {
  while (b) {
    uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

}  // namespace

extern "C" {

uint32_t pgx_k_bits(uint32_t P, uint32_t d) {
  std::vector<uint32_t> K;
  if (!primorial_over(P, d, K)) return 0;
  return bits_of(K);
}

int pgx_gen_multipliers(uint64_t seed, uint64_t m_lo, uint64_t m_hi, uint32_t d, uint32_t n_m,
                        uint64_t* out_m) {
  if (!out_m || m_hi <= m_lo || d == 0) return PGX_E_ARG;
  // enough coprime values must exist
  uint64_t avail = 0;
  for (uint64_t m = m_lo; m < m_hi && avail < n_m; ++m) avail += gcd64(m, d) == 1;
  if (avail < n_m) return PGX_E_ARG;
  Xoshiro rng(seed);
  std::vector<uint64_t> chosen;
  chosen.reserve(n_m);
  while (chosen.size() < n_m) {
    uint64_t m = m_lo + rng.below(m_hi - m_lo);
    if (gcd64(m, d) != 1) continue;
    if (std::find(chosen.begin(), chosen.end(), m) != chosen.end()) continue;
    chosen.push_back(m);
  }
  std::sort(chosen.begin(), chosen.end());
  std::memcpy(out_m, chosen.data(), n_m * sizeof(uint64_t));
  return PGX_OK;
}

int pgx_gen_sieved(uint32_t P, uint32_t d, const uint64_t* m, uint32_t n_m, uint32_t x_max,
                   uint32_t per_m, uint32_t sieve_limit, uint32_t limbs, uint32_t* out_limbs,
                   uint32_t* out_m, uint32_t* out_x, size_t* written) {
  return pgx_gen_sieved_side(P, d, +1, m, n_m, x_max, per_m, sieve_limit, limbs, out_limbs, out_m,
                             out_x, written);
}

int pgx_gen_sieved_side(uint32_t P, uint32_t d, int32_t sign, const uint64_t* m, uint32_t n_m,
                        uint32_t x_max, uint32_t per_m, uint32_t sieve_limit, uint32_t limbs,
                        uint32_t* out_limbs, uint32_t* out_m, uint32_t* out_x, size_t* written) {
  if (!m || !out_limbs || !written || limbs == 0 || x_max == 0 || (sign != 1 && sign != -1))
    return PGX_E_ARG;
  *written = 0;
  std::vector<uint32_t> K;
  if (!primorial_over(P, d, K)) return PGX_E_INPUT;
  const std::vector<uint32_t> primes = primes_below(sieve_limit);
  std::vector<uint32_t> kmod(primes.size());
  for (size_t k = 0; k < primes.size(); ++k) kmod[k] = mod_small(K, primes[k]);
  std::vector<uint8_t> comp(x_max + 1);
  size_t w = 0;
  for (uint32_t mi = 0; mi < n_m; ++mi) {
    std::fill(comp.begin(), comp.end(), 0);
    for (size_t k = 0; k < primes.size(); ++k) {
      const uint32_t p = primes[k];
      const uint64_t r = (uint64_t)(m[mi] % p) * kmod[k] % p;  // m*K mod p
      // first x >= 1 with m*K + sign*x == 0 (mod p)
      uint64_t x = (sign > 0) ? (p - r) % p : r;
      if (x == 0) x = p;
      for (; x <= x_max; x += p) comp[x] = 1;
    }
    uint32_t taken = 0;
    for (uint32_t x = 1; x <= x_max && taken < per_m; ++x) {
      if (comp[x]) continue;
      std::vector<uint32_t> N = K;
      if (sign > 0) {
        mul_small(N, m[mi], x);
      } else {  // m*K - x (m*K > x_max for every configuration this builds)
        mul_small(N, m[mi], 0);
        if (!sub_small(N, x)) return PGX_E_INPUT;
      }
      if (N.size() > limbs || bits_of(N) > 32u * limbs) return PGX_E_INPUT;
      uint32_t* o = out_limbs + w * limbs;
      std::memset(o, 0, limbs * sizeof(uint32_t));
      std::memcpy(o, N.data(), N.size() * sizeof(uint32_t));
      if (out_m) out_m[w] = mi;
      if (out_x) out_x[w] = x;
      ++w;
      ++taken;
    }
  }
  *written = w;
  return PGX_OK;
}

}  // extern "C"
