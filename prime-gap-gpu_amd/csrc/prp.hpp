// Per-candidate probable-prime test built on mont.hpp (one candidate per lane).
//
// Follows the verdict semantics of the reference's miller_rabin_t::miller_rabin
// (miller_rabin.h:216-248) for each base a, with d = (N-1) >> t, t = ctz(N-1):
//   y_j = a^(d * 2^j) mod N, j = 0..t
//   SPRP  : y_0 in {1, N-1}  or  y_j == N-1 for some 1 <= j <= t-1
//   QUIRK : SPRP  or  (t >= 2 and y_t == N-1)      (the reference squares t times,
//           miller_rabin.h:236-244, so it also accepts a^(N-1) == -1)
//   FERMAT: y_t == 1                                (a^(N-1) mod N == 1)
// The exponentiation is left-to-right binary over the bits of N-1 (the
// reference uses a fixed 2^W window table, miller_rabin.h:160-214; the result
// does not depend on the chunking).  For base 2 the per-bit multiply is a lazy
// doubling, O(S) instead of a second Montgomery product, so no table is needed.
//
// One loop covers all bits L-2 .. 0 of N-1: the bits below t are zero, so the
// strong-test checks of y_0..y_{t-1} ride on the same squaring call site.  Only
// one copy of the S^2 squaring body is emitted per base kind, which keeps the
// loop inside the instruction cache.
#pragma once

#include "mont.hpp"

namespace pgx {

enum : uint32_t {
  FLAG_FERMAT = 1u,  // a^(N-1) == 1
  FLAG_SPRP = 2u,    // textbook strong probable prime to base a
  FLAG_QUIRK = 4u,   // the reference's verdict (SPRP or t >= 2 and a^(N-1) == -1)
};

// Step that multiplies the running power by the base when an exponent bit is
// set.  Base 2: lazy doubling.
template <int S>
struct MulBy2 {
  PGX_FN void operator()(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t, uint32_t bit) const {
    dbl_lazy<S>(x, n, bit);
  }
};

// General base: Montgomery product by am = Mont(a) on set bits.
template <int S>
struct MulByMont {
  const uint32_t (&am)[S];
  PGX_FN void operator()(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t ninv,
                         uint32_t bit) const {
    if (bit) mont_mul<S>(x, am, n, ninv);
  }
};

// Shared exponentiation + strong-test tail.  x enters as Mont(a) (the leading
// bit of N-1).  res receives a^(N-1) mod N (canonical); returns FLAG_* bits.
template <int S, class LimbFn, class Step>
PGX_FN uint32_t powm_verdict(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t ninv, int L,
                             int t, LimbFn limb, const Step& step, uint32_t (&res)[S]) {
  uint32_t one[S];
  bool sprp = false;
  if (t == L - 1) {  // N - 1 = 2^t: y_0 = a
    r_mod_n<S>(one, n, L);
    canon<S>(x, n);
    sprp = equal<S>(x, one) || is_neg<S>(x, one, n);
  }
  uint32_t word = 0;
  for (int p = L - 2; p >= 0; --p) {
    if (p == L - 2 || (p & 31) == 31) word = limb(p >> 5);
    mont_sqr<S>(x, n, ninv);
    // bit p of N-1: bit p of N for p >= 1, zero for p == 0
    step(x, n, ninv, (p > 0) ? ((word >> (p & 31)) & 1u) : 0u);
    if (p <= t && p >= 1) {  // x = Mont(y_{t-p})
      if (p == t) r_mod_n<S>(one, n, L);
      canon<S>(x, n);
      sprp = sprp || is_neg<S>(x, one, n) || (p == t && equal<S>(x, one));
    }
  }
#pragma unroll
  for (int j = 0; j < S; ++j) res[j] = x[j];
  redc<S>(res, n, ninv);  // canonical: res < N
  uint32_t is_one = res[0] ^ 1u, is_m1 = res[0] ^ (n[0] - 1u);
#pragma unroll
  for (int j = 1; j < S; ++j) {
    is_one |= res[j];
    is_m1 |= res[j] ^ n[j];
  }
  uint32_t f = 0;
  if (is_one == 0) f |= FLAG_FERMAT;
  if (sprp) f |= FLAG_SPRP;
  if (sprp || (t >= 2 && is_m1 == 0)) f |= FLAG_QUIRK;
  return f;
}

// Base-2 test.  `limb(j)` returns limb j of N (runtime j); it is read once per
// 32 exponent bits so the exponent never needs dynamically indexed registers.
template <int S, class LimbFn>
PGX_FN uint32_t test_base2(const uint32_t (&n)[S], uint32_t ninv, int L, int t, LimbFn limb,
                           uint32_t (&res)[S]) {
  uint32_t x[S];
  r_mod_n<S>(x, n, L);
  dbl_full<S>(x, n);  // Mont(2)
  return powm_verdict<S>(x, n, ninv, L, t, limb, MulBy2<S>{}, res);
}

// General base a >= 2 (the extra Miller-Rabin bases).
template <int S, class LimbFn>
PGX_FN uint32_t test_base(uint32_t a, const uint32_t (&n)[S], uint32_t ninv, int L, int t,
                          LimbFn limb, uint32_t (&res)[S]) {
  uint32_t am[S];
  {
    uint32_t one[S];
    r_mod_n<S>(one, n, L);
    // am = a * R mod N by double-and-add over the bits of a
#pragma unroll
    for (int j = 0; j < S; ++j) am[j] = 0;
    for (int b = 31 - __builtin_clz(a); b >= 0; --b) {
      dbl_full<S>(am, n);
      if ((a >> b) & 1u) add_full<S>(am, one, n);
    }
  }
  uint32_t x[S];
#pragma unroll
  for (int j = 0; j < S; ++j) x[j] = am[j];
  return powm_verdict<S>(x, n, ninv, L, t, limb, MulByMont<S>{am}, res);
}

// Validity of a candidate for a kernel of S limbs: odd, > 1 and 4N < 2^(32S).
template <int S>
PGX_FN bool fits(const uint32_t (&n)[S], int L) {
  return (n[0] & 1u) && L >= 2 && L <= 32 * S - 2;
}

}  // namespace pgx
