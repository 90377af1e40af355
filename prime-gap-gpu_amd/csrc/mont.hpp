// Multi-limb Montgomery arithmetic for one candidate per lane (gfx950).
//
// Registers hold numbers in radix 2^RB (RB = 29 for up to 31 limbs, 28 above),
// S limbs, each limb < 2^RB.  Candidates arrive and residues leave as 32-bit
// limbs (the reference's layout, miller_rabin.h:39-54, 142-145); the radix is
// internal to the registers.
//
// Why not radix 2^32: v_mad_u64_u32 accumulates into 64 bits, and a 32x32-bit
// product leaves no headroom, so every product would need a second instruction
// to catch the carry (through an SGPR, with a hazard wait state).  With RB-bit
// limbs a product is < 2^(2 RB) and a whole column of up to 2S+1 products
// (cross products counted twice, the square, the reduction products) plus the
// incoming carry stays below 2^64 (RB = 29: 2S+1 <= 63; RB = 28: 2S+1 <= 255).
// Every product is then exactly one v_mad_u64_u32 into a 64-bit accumulator.
//
// Montgomery: R = 2^(RB*S) with R >= 16 N.  Values are kept "lazy": after a
// Montgomery square or product x < 2N; after a doubling x < 4N; and a
// Montgomery square of any x < 4N returns < (16N^2 + RN)/R < 2N.  So the base-2
// doubling is a bare shift by one bit and no conditional subtraction appears
// in the hot loop.  Where the column bound leaves room (RB = 28, or RB = 29
// with S <= 21) the doubling is folded into the squaring itself
// (mont_sqr_dbl: (2^bit x^2 + mN) / R), which removes the shift pass and its
// dependency chain from the loop.
//
// The squaring is finely-integrated product scanning: column k of x^2 (cross
// products accumulated once and doubled) and of m*N share one accumulator, the
// reduction digit m_k comes from column k's low RB bits, and the column's low
// limb is retired.  Two accumulators per product kind keep independent mad
// chains in flight.  The reference performs the same operation with CGBN's
// cgbn_mont_sqr / cgbn_mont_mul across a TPI-thread group (miller_rabin.h:181,
// 203, 209); here one lane owns a whole candidate, so the hot loop has no
// cross-lane traffic.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PGX_FN __host__ __device__ __forceinline__

namespace pgx {

// Radix of a kernel with S limbs.
template <int S>
struct Radix {
  static constexpr int RB = (S <= 31) ? 29 : 28;
  static constexpr uint32_t MASK = (1u << RB) - 1u;
  // largest bit length a candidate may have: R = 2^(RB*S) >= 16 N
  static constexpr int MAX_BITS = RB * S - 4;
  // largest bit length for the squaring with a zero top limb (x < 2N < 2^(RB(S-1)))
  static constexpr int TOP0_MAX_BITS = RB * (S - 1) - 1;
  // 32-bit words spanned by S radix limbs
  static constexpr int W32 = (RB * S + 31) / 32;
};

PGX_FN uint64_t mul64(uint32_t a, uint32_t b) { return (uint64_t)a * b; }

// True when pred holds on some active lane of the wavefront; a host build
// emulates one lane, whose own pred it returns.
PGX_FN bool wave_any(bool pred) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __ballot(pred) != 0;
#else
  return pred;
#endif
}

// Largest v < 2^bits over the active lanes of the wavefront, as a
// wavefront-uniform value (bisection over ballots: exact, and the result lives
// in scalar registers); a host build emulates one lane and returns v.
template <int kBits>
PGX_FN uint32_t wave_max(uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t m = 0;
#pragma unroll
  for (int b = kBits - 1; b >= 0; --b) {
    const uint32_t c = m | (1u << b);
    if (__ballot(v >= c) != 0) m = c;
  }
  return m;
#else
  return v;
#endif
}

// -N^-1 mod 2^32 for odd n0 (Newton iteration, 3 -> 6 -> 12 -> 24 -> 48 bits);
// its low RB bits are -N^-1 mod 2^RB.
PGX_FN uint32_t neg_inv32(uint32_t n0) {
  uint32_t inv = n0;
  inv *= 2u - n0 * inv;
  inv *= 2u - n0 * inv;
  inv *= 2u - n0 * inv;
  inv *= 2u - n0 * inv;
  return 0u - inv;
}

// x <- x^2 / R mod N (lazy: x < 4N in, x < 2N out).
template <int S>
PGX_FN void mont_sqr(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t ninv) {
  constexpr int RB = Radix<S>::RB;
  constexpr uint32_t MASK = Radix<S>::MASK;
  uint32_t m[S];
  uint64_t carry = 0;
#pragma unroll
  for (int k = 0; k < 2 * S - 1; ++k) {
    const int i0 = (k < S) ? 0 : k - S + 1;
    uint64_t c0 = 0, c1 = 0;
#pragma unroll
    for (int i = i0; i < k - i; ++i) {
      if ((i - i0) & 1) c1 += mul64(x[i], x[k - i]);
      else c0 += mul64(x[i], x[k - i]);
    }
    uint64_t t0 = carry, t1 = 0;
    if ((k & 1) == 0) t1 += mul64(x[k >> 1], x[k >> 1]);
    const int i1 = (k < S) ? k : S;
#pragma unroll
    for (int i = i0; i < i1; ++i) {
      if ((i - i0) & 1) t1 += mul64(m[i], n[k - i]);
      else t0 += mul64(m[i], n[k - i]);
    }
    uint64_t t = t0 + t1 + ((c0 + c1) << 1);
    if (k < S) {
      m[k] = ((uint32_t)t * ninv) & MASK;
      t += mul64(m[k], n[0]);  // low RB bits become zero
    } else {
      x[k - S] = (uint32_t)t & MASK;
    }
    carry = t >> RB;
  }
  x[S - 1] = (uint32_t)carry;
}

// x <- 2^bit * x^2 / R mod N with bit in {0, 1}: the base-2 step of a bit of
// the exponent folded into the squaring (cross products shifted by 1 + bit,
// the square term by bit), so no separate doubling pass.  x < 2N in and out:
// (2 * 4N^2 + R N) / R < 2N because R >= 8N.  Column sums stay below 2^64 for
// RB = 28 (any S <= 80) and for RB = 29 up to S = 21 (1.5 S 2^59 < 2^64).
//
// kPreShift: the shifts move onto the operands (2^bit x^2 = sum_{i<j}
// (x_i << (1+bit)) x_j + sum_i (x_i << bit) x_i; limbs < 2^RB, so both fit 32
// bits) and every column is a pure chain of multiply-adds, 2S shifts per
// squaring instead of about three per column.  Otherwise the column sums are
// shifted, which keeps more independent work at the start of a squaring.
template <int S, bool kPreShift = true>
PGX_FN void mont_sqr_dbl(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t ninv, uint32_t bit) {
  constexpr int RB = Radix<S>::RB;
  constexpr uint32_t MASK = Radix<S>::MASK;
  static_assert(RB == 28 || S <= 21, "column bound of the folded doubling");
  if constexpr (kPreShift) {
    uint32_t xs[S], xd[S];
#pragma unroll
    for (int i = 0; i < S; ++i) {
      xs[i] = x[i] << bit;
      xd[i] = xs[i] << 1;
    }
    uint32_t m[S];
    uint64_t carry = 0;
#pragma unroll
    for (int k = 0; k < 2 * S - 1; ++k) {
      const int i0 = (k < S) ? 0 : k - S + 1;
      uint64_t t = carry;
#pragma unroll
      for (int i = i0; i < k - i; ++i) t += mul64(xd[i], x[k - i]);
      if ((k & 1) == 0) t += mul64(xs[k >> 1], x[k >> 1]);
      const int i1 = (k < S) ? k : S;
#pragma unroll
      for (int i = i0; i < i1; ++i) t += mul64(m[i], n[k - i]);
      if (k < S) {
        m[k] = ((uint32_t)t * ninv) & MASK;
        t += mul64(m[k], n[0]);  // low RB bits become zero
      } else {
        x[k - S] = (uint32_t)t & MASK;
      }
      carry = t >> RB;
    }
    x[S - 1] = (uint32_t)carry;
  } else {
    uint32_t m[S];
    uint64_t carry = 0;
#pragma unroll
    for (int k = 0; k < 2 * S - 1; ++k) {
      const int i0 = (k < S) ? 0 : k - S + 1;
      uint64_t c0 = 0, c1 = 0;
#pragma unroll
      for (int i = i0; i < k - i; ++i) {
        if ((i - i0) & 1) c1 += mul64(x[i], x[k - i]);
        else c0 += mul64(x[i], x[k - i]);
      }
      uint64_t t0 = carry, t1 = 0;
      const int i1 = (k < S) ? k : S;
#pragma unroll
      for (int i = i0; i < i1; ++i) {
        if ((i - i0) & 1) t1 += mul64(m[i], n[k - i]);
        else t0 += mul64(m[i], n[k - i]);
      }
      uint64_t t = t0 + t1 + ((c0 + c1) << (1 + bit));
      if ((k & 1) == 0) t += mul64(x[k >> 1], x[k >> 1]) << bit;
      if (k < S) {
        m[k] = ((uint32_t)t * ninv) & MASK;
        t += mul64(m[k], n[0]);  // low RB bits become zero
      } else {
        x[k - S] = (uint32_t)t & MASK;
      }
      carry = t >> RB;
    }
    x[S - 1] = (uint32_t)carry;
  }
}

// The folded squaring (operands shifted) when the top limb of x and of N is
// zero: N < 2^(RB (S-1) - 1), so x < 2N fits in S-1 limbs, while R = 2^(RB S)
// keeps its S digits (the folded doubling needs R > 8N; for N of RB(S-1) - 2
// or RB(S-1) - 1 bits, S-1 digits fall short by a bit or two).  Every product
// with limb S-1 of x or N is zero and is skipped: V(V+1)/2 + S V multiply-adds
// instead of S(S+1)/2 + S^2 (V = S-1), one column and two operand shifts fewer.
// The caller guarantees the bounds (powm_verdict runs it only in wavefronts
// whose candidates all fit).
template <int S>
PGX_FN void mont_sqr_dbl_top0(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t ninv,
                              uint32_t bit) {
  constexpr int RB = Radix<S>::RB;
  constexpr uint32_t MASK = Radix<S>::MASK;
  constexpr int V = S - 1;  // value limbs
  static_assert(S >= 2, "one value limb at least");
  static_assert(RB == 28 || S <= 21, "column bound of the folded doubling");
  uint32_t xs[V], xd[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    xs[i] = x[i] << bit;
    xd[i] = xs[i] << 1;
  }
  uint32_t m[S];
  uint64_t carry = 0;
  // columns 0 .. S+V-2: x^2 fills 0 .. 2V-2, m*N fills 0 .. S+V-2
#pragma unroll
  for (int k = 0; k < S + V - 1; ++k) {
    const int i0 = (k < V) ? 0 : k - V + 1;  // the other index stays below V
    uint64_t t = carry;
#pragma unroll
    for (int i = i0; i < k - i; ++i) t += mul64(xd[i], x[k - i]);
    if ((k & 1) == 0 && (k >> 1) < V) t += mul64(xs[k >> 1], x[k >> 1]);
    const int i1 = (k < S) ? k : S;
#pragma unroll
    for (int i = i0; i < i1; ++i) t += mul64(m[i], n[k - i]);
    if (k < S) {
      m[k] = ((uint32_t)t * ninv) & MASK;
      t += mul64(m[k], n[0]);  // low RB bits become zero
    } else {
      x[k - S] = (uint32_t)t & MASK;
    }
    carry = t >> RB;
  }
  x[V - 1] = (uint32_t)carry;
  x[S - 1] = 0u;
}

// x <- x * y / R mod N (x < 4N, y < 4N with x*y < 16N^2; result < 2N).
// x and y must not alias.
template <int S>
PGX_FN void mont_mul(uint32_t (&x)[S], const uint32_t (&y)[S], const uint32_t (&n)[S],
                     uint32_t ninv) {
  constexpr int RB = Radix<S>::RB;
  constexpr uint32_t MASK = Radix<S>::MASK;
  uint32_t m[S];
  uint64_t carry = 0;
#pragma unroll
  for (int k = 0; k < 2 * S - 1; ++k) {
    const int i0 = (k < S) ? 0 : k - S + 1;
    const int i1 = (k < S) ? k : S - 1;
    uint64_t t0 = carry, t1 = 0;
#pragma unroll
    for (int i = i0; i <= i1; ++i) {
      if ((i - i0) & 1) t1 += mul64(x[i], y[k - i]);
      else t0 += mul64(x[i], y[k - i]);
    }
    const int r1 = (k < S) ? k : S;
#pragma unroll
    for (int i = i0; i < r1; ++i) {
      if ((i - i0) & 1) t0 += mul64(m[i], n[k - i]);
      else t1 += mul64(m[i], n[k - i]);
    }
    uint64_t t = t0 + t1;
    if (k < S) {
      m[k] = ((uint32_t)t * ninv) & MASK;
      t += mul64(m[k], n[0]);
    } else {
      x[k - S] = (uint32_t)t & MASK;
    }
    carry = t >> RB;
  }
  x[S - 1] = (uint32_t)carry;
}

// x <- x / R mod N.  For x < 4N the result is < N + 1, hence canonical for
// every x that is not a multiple of N.
template <int S>
PGX_FN void redc(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t ninv) {
  uint32_t one[S];
#pragma unroll
  for (int j = 0; j < S; ++j) one[j] = (j == 0) ? 1u : 0u;
  mont_mul<S>(x, one, n, ninv);
}

// Lazy doubling: when bit is 1, x <- 2x (a one-bit shift across limbs, no
// reduction: x < 2N becomes < 4N).  bit must be 0 or 1.
template <int S>
PGX_FN void dbl_lazy(uint32_t (&x)[S], uint32_t bit) {
  constexpr int RB = Radix<S>::RB;
  constexpr uint32_t MASK = Radix<S>::MASK;
#pragma unroll
  for (int j = S - 1; j >= 1; --j)
    x[j] = ((x[j] << bit) | (x[j - 1] >> (RB - bit))) & MASK;
  x[0] = (x[0] << bit) & MASK;
}

// d <- x - n over normalized limbs; returns 1 when it borrowed (x < n).
template <int S>
PGX_FN uint32_t sub_borrow(uint32_t (&d)[S], const uint32_t (&x)[S], const uint32_t (&n)[S]) {
  constexpr int RB = Radix<S>::RB;
  constexpr uint32_t MASK = Radix<S>::MASK;
  uint32_t br = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const uint32_t v = x[j] - n[j] - br;  // wraps when negative
    d[j] = v & MASK;
    br = v >> 31;  // |v| < 2^RB + 1 < 2^31, so bit 31 is the sign
  }
  (void)RB;
  return br;
}

// x <- x - N when x >= N.
template <int S>
PGX_FN void sub_if_ge(uint32_t (&x)[S], const uint32_t (&n)[S]) {
  uint32_t d[S];
  const uint32_t lt = sub_borrow<S>(d, x, n);
#pragma unroll
  for (int j = 0; j < S; ++j) x[j] = lt ? x[j] : d[j];
}

// Canonical form of a lazy value x < 4N.
template <int S>
PGX_FN void canon(uint32_t (&x)[S], const uint32_t (&n)[S]) {
  sub_if_ge<S>(x, n);
  sub_if_ge<S>(x, n);
  sub_if_ge<S>(x, n);
}

// Full doubling for canonical v < N: v <- 2v mod N.
template <int S>
PGX_FN void dbl_full(uint32_t (&v)[S], const uint32_t (&n)[S]) {
  dbl_lazy<S>(v, 1u);
  sub_if_ge<S>(v, n);
}

// v <- (v + w) mod N for canonical v, w < N.
template <int S>
PGX_FN void add_full(uint32_t (&v)[S], const uint32_t (&w)[S], const uint32_t (&n)[S]) {
  constexpr int RB = Radix<S>::RB;
  constexpr uint32_t MASK = Radix<S>::MASK;
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const uint32_t s = v[j] + w[j] + c;
    v[j] = s & MASK;
    c = s >> RB;
  }
  sub_if_ge<S>(v, n);
}

// v scaled by 2^-L as a double: sum_j v_j 2^(RB j - L) (each term exact, S
// rounded additions).  With L = bit length of N, N scales into [0.5, 1).
template <int S>
PGX_FN double scaled(const uint32_t (&v)[S], int L) {
  constexpr int RB = Radix<S>::RB;
  double d = 0.0;
#pragma unroll
  for (int j = 0; j < S; ++j) d += ldexp((double)v[j], RB * j - L);
  return d;
}

// x <- a x mod N for an ordinary integer 1 <= a <= 2^16, which keeps the
// Montgomery form (the step of a set exponent bit for a small base, and 16
// doublings at a time in r_mod_n): x < 2N in (limbs < 2^RB), x < 2N out.
// q = floor(a x / N) is estimated in double precision (relative error below
// 2 S 2^-53 <= 2^-45) and biased down by 2^-40, so it is the true quotient or
// one less (a x / N < 2^17, so the bias moves it by less than 2^-22) and
// a x - q N lies in [0, 2N).  One pass forms a x_j - q n_j with a signed carry
// (|a x_j - q n_j| < 2^47).  inv_n = 1 / scaled(n, L).
template <int S>
PGX_FN void mul_small(uint32_t (&x)[S], uint32_t a, const uint32_t (&n)[S], double inv_n, int L) {
  constexpr int RB = Radix<S>::RB;
  constexpr uint32_t MASK = Radix<S>::MASK;
  const double qd = (double)a * scaled<S>(x, L) * inv_n * (1.0 - 0x1p-40);
  const uint32_t q = (qd > 0.0) ? (uint32_t)qd : 0u;
  int64_t c = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int64_t t = (int64_t)((uint64_t)x[j] * a) - (int64_t)((uint64_t)n[j] * q) + c;
    x[j] = (uint32_t)t & MASK;
    c = t >> RB;  // arithmetic: the floor of t / 2^RB
  }
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
  constexpr int RB = Radix<S>::RB;
  constexpr uint32_t MASK = Radix<S>::MASK;
  uint32_t c = 0, diff = 0;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const uint32_t s = a[j] + b[j] + c;
    diff |= (s & MASK) ^ n[j];
    c = s >> RB;
  }
  return (diff | c) == 0;
}

// v <- 2^bit (bit < RB*S).
template <int S>
PGX_FN void set_pow2(uint32_t (&v)[S], int bit) {
  constexpr int RB = Radix<S>::RB;
  const int q = bit / RB, r = bit - q * RB;
#pragma unroll
  for (int j = 0; j < S; ++j) v[j] = (j == q) ? (1u << r) : 0u;
}

// one <- R mod N = 2^(RB*S) mod N: 2^(L-1) (which is < N) doubled
// RB*S - L + 1 times, 16 doublings per mul_small (x < 2N throughout), then one
// conditional subtraction.  (One doubling per step with a full compare cost
// about 8 S instructions a doubling, 4,400 per 462-bit candidate.)
template <int S>
PGX_FN void r_mod_n(uint32_t (&one)[S], const uint32_t (&n)[S], int L) {
  constexpr int RB = Radix<S>::RB;
  set_pow2<S>(one, L - 1);
  const double inv_n = 1.0 / scaled<S>(n, L);
  for (int steps = RB * S - (L - 1); steps > 0; steps -= 16)
    mul_small<S>(one, 1u << (steps < 16 ? steps : 16), n, inv_n, L);
  sub_if_ge<S>(one, n);
}

// ---- conversion between 32-bit words and radix-2^RB limbs ----------------

// Radix-RB limbs of a number given as W32 32-bit words.
template <int S>
PGX_FN void from_words(uint32_t (&x)[S], const uint32_t (&w)[Radix<S>::W32]) {
  constexpr int RB = Radix<S>::RB;
  constexpr int W = Radix<S>::W32;
  constexpr uint32_t MASK = Radix<S>::MASK;
#pragma unroll
  for (int j = 0; j < S; ++j) {
    const int b = RB * j, q = b >> 5, r = b & 31;
    uint32_t v = w[q] >> r;
    if (r + RB > 32 && q + 1 < W) v |= w[q + 1] << (32 - r);
    x[j] = v & MASK;
  }
}

// 32-bit words of a number given as S radix-RB limbs.
template <int S>
PGX_FN void to_words(uint32_t (&w)[Radix<S>::W32], const uint32_t (&x)[S]) {
  constexpr int RB = Radix<S>::RB;
  constexpr int W = Radix<S>::W32;
#pragma unroll
  for (int q = 0; q < W; ++q) {
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const int lo = RB * j;  // limb j covers bits [lo, lo + RB)
      if (lo + RB <= 32 * q || lo >= 32 * q + 32) continue;
      const int sh = lo - 32 * q;
      v |= (sh >= 0) ? (x[j] << sh) : (x[j] >> -sh);
    }
    w[q] = v;
  }
}

// Bit length of a number given as W 32-bit words (0 for zero).
template <int W>
PGX_FN int bit_length_words(const uint32_t (&w)[W]) {
  int L = 0;
#pragma unroll
  for (int j = 0; j < W; ++j)
    if (w[j]) L = 32 * j + 32 - __builtin_clz(w[j]);
  return L;
}

// Trailing zero count of N - 1 for odd N > 1, N given as W 32-bit words.
template <int W>
PGX_FN int ctz_nm1_words(const uint32_t (&w)[W]) {
  int t = -1;
#pragma unroll
  for (int j = W - 1; j >= 0; --j) {
    const uint32_t v = (j == 0) ? (w[0] & ~1u) : w[j];
    if (v) t = 32 * j + __builtin_ctz(v);
  }
  return t;
}

}  // namespace pgx
