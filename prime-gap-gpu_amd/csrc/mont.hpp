// Multi-limb Montgomery arithmetic for one candidate per lane (gfx950).
//
// Numbers are S little-endian 32-bit limbs held in VGPRs.  S is a compile-time
// limb count chosen per bucket so that every candidate N of the bucket satisfies
// N < 2^(32*S - 2), i.e. 4N < R with R = 2^(32*S).  That slack lets every value
// stay in the "lazy" range [0, 2N) through the whole exponentiation: the
// Montgomery product of two lazy values is again lazy, so no conditional
// subtraction is needed inside the hot loop.
//
// The product is formed by finely-integrated product scanning: column k of the
// square x^2 and column k of the reduction m*N are summed into one 96-bit
// accumulator, the reduction digit m_k is derived from column k's low word, and
// the column's low word is retired.  Each 32x32->64 product is one
// v_mad_u64_u32 plus one carry add into the accumulator's third word.
//
// The reference performs the same mathematical operation with CGBN's
// cgbn_mont_sqr / cgbn_mont_mul across a TPI-thread group
// (miller_rabin.h:181, 203, 209); here one lane owns a whole candidate, so no
// cross-lane traffic is needed in the hot loop.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PGX_FN __host__ __device__ __forceinline__

namespace pgx {

// acc(hi:lo) += a * b, with lo 64 bits and hi the third 32-bit word.
// On the device this is v_mad_u64_u32 with its carry-out in an SGPR pair,
// consumed by v_addc_co_u32 (two instructions per product; the portable
// expression below compiles to four).  The host build of the same routine is
// what the CPU tests exercise.
PGX_FN void mac(uint64_t& lo, uint32_t& hi, uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t c;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(lo), "=s"(c) : "v"(a), "v"(b));
  asm("v_addc_co_u32 %0, %1, %0, 0, %1" : "+v"(hi), "+s"(c));
#else
  uint64_t p = (uint64_t)a * b;
  lo += p;
  hi += (lo < p) ? 1u : 0u;
#endif
}

// -N^-1 mod 2^32 for odd n0 (Newton iteration, 3 -> 6 -> 12 -> 24 -> 48 bits).
PGX_FN uint32_t neg_inv32(uint32_t n0) {
  uint32_t inv = n0;
  inv *= 2u - n0 * inv;
  inv *= 2u - n0 * inv;
  inv *= 2u - n0 * inv;
  inv *= 2u - n0 * inv;
  return 0u - inv;
}

// x <- x^2 / R mod N (lazy: x < 2N in, x < 2N out).
template <int S>
PGX_FN void mont_sqr(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t ninv) {
  uint32_t m[S];
  uint64_t t = 0;
  uint32_t t2 = 0;
#pragma unroll
  for (int k = 0; k < 2 * S - 1; ++k) {
    const int i0 = (k < S) ? 0 : k - S + 1;
    // cross products x_i x_j, i < j, i + j = k
    uint64_t c = 0;
    uint32_t c2 = 0;
#pragma unroll
    for (int i = i0; i < k - i; ++i) mac(c, c2, x[i], x[k - i]);
    c2 = (c2 << 1) | (uint32_t)(c >> 63);
    c <<= 1;
    if ((k & 1) == 0) mac(c, c2, x[k >> 1], x[k >> 1]);
    t += c;
    t2 += c2 + ((t < c) ? 1u : 0u);
    // reduction products m_i n_{k-i}
    const int i1 = (k < S) ? k : S;
#pragma unroll
    for (int i = i0; i < i1; ++i) mac(t, t2, m[i], n[k - i]);
    if (k < S) {
      m[k] = (uint32_t)t * ninv;
      mac(t, t2, m[k], n[0]);
    } else {
      x[k - S] = (uint32_t)t;
    }
    t = (t >> 32) | ((uint64_t)t2 << 32);
    t2 = 0;
  }
  x[S - 1] = (uint32_t)t;
}

// x <- x * y / R mod N (lazy).  x and y must not alias.
template <int S>
PGX_FN void mont_mul(uint32_t (&x)[S], const uint32_t (&y)[S], const uint32_t (&n)[S],
                     uint32_t ninv) {
  uint32_t m[S];
  uint64_t t = 0;
  uint32_t t2 = 0;
#pragma unroll
  for (int k = 0; k < 2 * S - 1; ++k) {
    const int i0 = (k < S) ? 0 : k - S + 1;
    const int i1 = (k < S) ? k : S - 1;
#pragma unroll
    for (int i = i0; i <= i1; ++i) mac(t, t2, x[i], y[k - i]);
    const int r1 = (k < S) ? k : S;
#pragma unroll
    for (int i = i0; i < r1; ++i) mac(t, t2, m[i], n[k - i]);
    if (k < S) {
      m[k] = (uint32_t)t * ninv;
      mac(t, t2, m[k], n[0]);
    } else {
      x[k - S] = (uint32_t)t;
    }
    t = (t >> 32) | ((uint64_t)t2 << 32);
    t2 = 0;
  }
  x[S - 1] = (uint32_t)t;
}

// x <- x / R mod N.  For x < 2N the result is < N + 1, hence canonical for any x
// that is not a multiple of N.
template <int S>
PGX_FN void redc(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t ninv) {
#pragma unroll
  for (int i = 0; i < S; ++i) {
    uint32_t m = x[0] * ninv;
    uint64_t c = ((uint64_t)m * n[0] + x[0]) >> 32;
#pragma unroll
    for (int j = 1; j < S; ++j) {
      uint64_t v = (uint64_t)m * n[j] + x[j] + c;
      x[j - 1] = (uint32_t)v;
      c = v >> 32;
    }
    x[S - 1] = (uint32_t)c;
  }
}

// d <- x - n; returns 1 when it borrowed (x < n).
template <int S>
PGX_FN uint32_t sub_borrow(uint32_t (&d)[S], const uint32_t (&x)[S], const uint32_t (&n)[S]) {
  uint32_t br = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    uint64_t v = (uint64_t)x[j] - n[j] - br;
    d[j] = (uint32_t)v;
    br = (uint32_t)(v >> 32) & 1u;
  }
  return br;
}

// Lazy doubling: when bit is set, x <- 2 * (x >= N ? x - N : x), keeping x < 2N.
template <int S>
PGX_FN void dbl_lazy(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t bit) {
  uint32_t d[S];
  const uint32_t lt = sub_borrow<S>(d, x, n);
  uint32_t prev = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    uint32_t y = lt ? x[j] : d[j];
    uint32_t z = (y << 1) | (prev >> 31);
    prev = y;
    x[j] = bit ? z : x[j];
  }
}

// Full doubling for canonical v < N: v <- 2v mod N.
template <int S>
PGX_FN void dbl_full(uint32_t (&v)[S], const uint32_t (&n)[S]) {
  uint32_t y[S];
  uint32_t prev = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    y[j] = (v[j] << 1) | (prev >> 31);
    prev = v[j];
  }
  uint32_t d[S];
  const uint32_t lt = sub_borrow<S>(d, y, n);
#pragma unroll
  for (int j = 0; j < S; ++j) v[j] = lt ? y[j] : d[j];
}

// v <- (v + w) mod N for canonical v, w < N.
template <int S>
PGX_FN void add_full(uint32_t (&v)[S], const uint32_t (&w)[S], const uint32_t (&n)[S]) {
  uint32_t y[S];
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    uint64_t s = (uint64_t)v[j] + w[j] + c;
    y[j] = (uint32_t)s;
    c = (uint32_t)(s >> 32);
  }
  uint32_t d[S];
  const uint32_t lt = sub_borrow<S>(d, y, n);
#pragma unroll
  for (int j = 0; j < S; ++j) v[j] = lt ? y[j] : d[j];
}

// Canonical form of a lazy value: x < 2N -> x mod N.
template <int S>
PGX_FN void canon(uint32_t (&x)[S], const uint32_t (&n)[S]) {
  uint32_t d[S];
  const uint32_t lt = sub_borrow<S>(d, x, n);
#pragma unroll
  for (int j = 0; j < S; ++j) x[j] = lt ? x[j] : d[j];
}

template <int S>
PGX_FN bool equal(const uint32_t (&a)[S], const uint32_t (&b)[S]) {
  uint32_t diff = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) diff |= a[j] ^ b[j];
  return diff == 0;
}

// a == N - b  <=>  a + b == N (both canonical).
template <int S>
PGX_FN bool is_neg(const uint32_t (&a)[S], const uint32_t (&b)[S], const uint32_t (&n)[S]) {
  uint32_t c = 0, diff = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    uint64_t s = (uint64_t)a[j] + b[j] + c;
    diff |= (uint32_t)s ^ n[j];
    c = (uint32_t)(s >> 32);
  }
  return (diff | c) == 0;
}

// Bit length of an S-limb number (0 for zero).
template <int S>
PGX_FN int bit_length(const uint32_t (&n)[S]) {
  int L = 0;
#pragma unroll
  for (int j = 0; j < S; ++j)
    if (n[j]) L = 32 * j + 32 - __builtin_clz(n[j]);
  return L;
}

// Trailing zero count of N - 1 for odd N > 1.
template <int S>
PGX_FN int ctz_nm1(const uint32_t (&n)[S]) {
  int t = -1;
#pragma unroll
  for (int j = S - 1; j >= 0; --j) {
    uint32_t w = (j == 0) ? (n[0] & ~1u) : n[j];
    if (w) t = 32 * j + __builtin_ctz(w);
  }
  return t;
}

// v <- 2^(L-1), the largest power of two below an L-bit odd N.
template <int S>
PGX_FN void set_pow2(uint32_t (&v)[S], int bit) {
#pragma unroll
  for (int j = 0; j < S; ++j) v[j] = (j == (bit >> 5)) ? (1u << (bit & 31)) : 0u;
}

// one <- R mod N = 2^(32S) mod N, by doubling 2^(L-1) (which is < N).
template <int S>
PGX_FN void r_mod_n(uint32_t (&one)[S], const uint32_t (&n)[S], int L) {
  set_pow2<S>(one, L - 1);
  const int steps = 32 * S - (L - 1);
  for (int c = 0; c < steps; ++c) dbl_full<S>(one, n);
}

}  // namespace pgx
