// Per-candidate probable-prime test built on mont.hpp (one candidate per lane).
//
// Follows the verdict semantics of the reference's miller_rabin_t::miller_rabin
// (miller_rabin.h:216-248) for each base a, with d = (N-1) >> t, t = ctz(N-1):
//   y_j = a^(d * 2^j) mod N, j = 0..t
//   SPRP  : y_0 in {1, N-1}  or  y_j == N-1 for some 1 <= j <= t-1
//   QUIRK : SPRP  or  (t >= 2 and y_t == N-1)      (the reference squares t times,
//           miller_rabin.h:236-244, so it also accepts a^(N-1) == -1)
//   FERMAT: y_t == 1                                (a^(N-1) mod N == 1)
// (a^(N-1) == -1 mod N has no odd solution N > 1, so QUIRK == SPRP; both are
// computed anyway and the tests hold them equal.)
// The exponentiation is left-to-right binary over the bits of N-1 (the
// reference uses a fixed 2^W window table, miller_rabin.h:160-214; the result
// does not depend on the chunking).  For base 2 the per-bit multiply by 2 is
// folded into the squaring ((2^bit x^2 + mN) / R, mont_sqr_dbl) where the
// column bound allows it, and is a one-bit shift otherwise, so base 2 needs
// no table and no second Montgomery product.
//
// One loop covers all bits L-2 .. 0 of N-1: the bits below t are zero, so the
// strong-test checks of y_0..y_{t-1} ride on the same squaring call site and
// only one copy of the unrolled squaring is emitted per base kind.  The bit
// index is the same on every lane of a wavefront (it starts below the
// wavefront's longest N; a shorter N starts from Mont(1), whose squarings
// leave it Mont(1) until its own top bit multiplies in the base), so the loop
// control, the exponent word loads and the guard of the strong-test checks are
// scalar: a wavefront issues one instruction at a time, and per-lane control
// in the loop cost as many issue slots as the squaring's carries.
#pragma once

#include "../../include/pgx_defs.h"
#include "mont.hpp"

namespace pgx {

enum : uint32_t {
  FLAG_FERMAT = PGX_FLAG_FERMAT,  // a^(N-1) == 1
  FLAG_SPRP = PGX_FLAG_SPRP,      // textbook strong probable prime to base a
  FLAG_QUIRK = PGX_FLAG_QUIRK,    // the reference's verdict
};

struct Bases {
  uint32_t count;
  uint32_t a[PGX_MAX_BASES];
};

// Base-2 squarings of one exponent word as generated assembly (gen_sqr.py):
// run(x, n, ninv, w, b_hi, b_lo) does, for b = b_hi down to b_lo,
// x <- 2^((w >> b) & 1) x^2 / R mod N, the same operation as mont_sqr_dbl /
// mont_sqr_dbl_top0 (kTop0), or mont_sqr + dbl_lazy where the doubling is not
// folded (radix 2^29 above 21 limbs).  A translation unit built with
// PGX_ASM_SQUARINGS_INC includes the generated specializations; a width
// without one keeps the C++ loop.
template <int S, bool kTop0>
struct AsmSquarings {
  static constexpr bool kAvailable = false;
  __device__ static void run(uint32_t (&)[S], const uint32_t (&)[S], uint32_t, uint32_t, int, int) {}
};
// The same for a small Miller-Rabin base a: run(..., a - 1, (1 - 2^-40) / N)
// does x <- x^2 / R mod N, then x <- a x mod N where bit b is set (mont_sqr
// then mul_small, with the quotient estimated from all limbs).
template <int S>
struct AsmSmallBase {
  static constexpr bool kAvailable = false;
  __device__ static void run(uint32_t (&)[S], const uint32_t (&)[S], uint32_t, uint32_t, int, int,
                             uint32_t, double) {}
};
#if defined(PGX_ASM_SQUARINGS_INC)
#include PGX_ASM_SQUARINGS_INC
#endif

// Base 2: multiply by the base = doubling, folded into the squaring when the
// radix leaves room for it (RB = 28, or RB = 29 with S <= 21).  kPreShift
// selects the folded squaring's form (mont_sqr_dbl); kTop0 the folded squaring
// that skips a zero top limb (mont_sqr_dbl_top0, candidates of at most
// Radix<S>::TOP0_MAX_BITS bits).
template <int S, bool kPreShift = true, bool kTop0_ = false>
struct MulBy2 {
  static constexpr bool kBase2 = true;
  static constexpr bool kFolded = (Radix<S>::RB == 28) || (S <= 21);
  static constexpr bool kShiftOperands = kPreShift;
  static constexpr bool kTop0 = kTop0_ && kFolded && S >= 2;
  PGX_FN void operator()(uint32_t (&x)[S], const uint32_t (&)[S], uint32_t, uint32_t bit) const {
    dbl_lazy<S>(x, bit);
  }
};

// Small base a (2 <= a < 2^8): multiply by the ordinary integer a on set bits
// (mul_small, about 11 S instructions instead of a Montgomery product).
template <int S>
struct MulBySmall {
  static constexpr bool kBase2 = false;
  static constexpr bool kFolded = false;
  static constexpr bool kShiftOperands = true;
  static constexpr bool kTop0 = false;
  uint32_t a;
  double inv_n;
  int L;
  PGX_FN void operator()(uint32_t (&x)[S], const uint32_t (&n)[S], uint32_t, uint32_t bit) const {
    if (bit) mul_small<S>(x, a, n, inv_n, L);
  }
};

// The exponent loop over bits Lw-2 .. 0 of N-1 (x enters as Mont(a), or
// Mont(1) for lanes shorter than the wavefront's longest N, Lw) with the
// strong-test checks at bits t .. 1.  kSkipTop: the squaring that skips a
// zero top limb (every N of the wavefront has at most
// Radix<S>::TOP0_MAX_BITS bits).  Returns 1 when a check passed.
template <bool kSkipTop, int S, class WordFn, class Step>
PGX_FN uint32_t exponent_bits(uint32_t (&x)[S], const uint32_t (&one)[S],
                              const uint32_t (&n)[S], uint32_t ninv, int Lw, int t, int tw,
                              WordFn word, const Step& step) {
  // a word, not a bool: a bool set in a divergent block is kept as a lane
  // mask in scalar registers, merged on every pass of the loop
  uint32_t sprp = 0;
  const int top = (Lw - 2) >> 5;  // word of the first bit
  uint32_t w_next = word(top);
#if defined(__HIP_DEVICE_COMPILE__)
  constexpr bool kAsm2 = AsmSquarings<S, kSkipTop>::kAvailable && Step::kBase2 && Step::kShiftOperands;
  constexpr bool kAsmA = AsmSmallBase<S>::kAvailable && !Step::kBase2 && !kSkipTop;
  if constexpr (kAsm2 || kAsmA) {
    // Generated loop: the bits above tw (no strong-test check) of a word in
    // one call, then the checked bits one at a time; one call site, so one
    // copy of the squaring.
    uint32_t am1 = 0;
    double invn = 0.0;
    if constexpr (kAsmA) {
      am1 = step.a - 1u;
      invn = ldexp(step.inv_n, -step.L) * (1.0 - 0x1p-40);  // exact scaling, one rounding
    }
    for (int wi = top; wi >= 0; --wi) {
      const uint32_t w = (wi == 0) ? (w_next & ~1u) : w_next;
      if (wi > 0) w_next = word(wi - 1);
      int b = (wi == top) ? ((Lw - 2) & 31) : 31;
      while (b >= 0) {
        const int free_lo = max(tw + 1 - 32 * wi, 0);  // lowest bit of the word above tw
        const int lo = (b >= free_lo) ? free_lo : b;
        if constexpr (kAsm2)
          AsmSquarings<S, kSkipTop>::run(x, n, ninv, w, b, lo);
        else
          AsmSmallBase<S>::run(x, n, ninv, w, b, lo, am1, invn);
        const int p = 32 * wi + lo;
        if (p <= tw) {
          if (p <= t && p >= 1) {  // x = Mont(y_{t-p})
            canon<S>(x, n);
            if (is_neg<S>(x, one, n) || (p == t && equal<S>(x, one))) sprp = 1u;
          }
        }
        b = lo - 1;
      }
    }
    return sprp;
  }
#endif
#if defined(__HIP_DEVICE_COMPILE__)
  // Place the exponent loop just after a 256-byte boundary (the padding runs
  // once).  With the loop body identical instruction for instruction, its
  // speed at one wavefront per SIMD measured up to 20% apart between
  // placements (the 25-limb kernel: 1.58 against 1.92 ms per exponentiation);
  // after this boundary every offset tried ran at the fast speed.  Measured
  // plain-launch gains: +9% at 14 limbs, +13% at 23, +14% at 25, within 1% at
  // 11, 17, 20 and 46, -5% at 36; so the radix-2^29 kernels (S <= 31) align.
  if constexpr (S <= 31) asm volatile(".p2align 8\n\ts_nop 0\n\ts_nop 0");
#endif
  for (int wi = top; wi >= 0; --wi) {
    // 32 bits of N-1 (bit 0 of N-1 is zero; the other bits are N's); the next
    // word is requested one block of 32 squarings ahead
    const uint32_t w = (wi == 0) ? (w_next & ~1u) : w_next;
    if (wi > 0) w_next = word(wi - 1);
    int b = (wi == top) ? ((Lw - 2) & 31) : 31;
    do {
      const uint32_t bit = (w >> b) & 1u;
      if constexpr (kSkipTop) {
        mont_sqr_dbl_top0<S>(x, n, ninv, bit);
      } else if constexpr (Step::kFolded) {
        mont_sqr_dbl<S, Step::kShiftOperands>(x, n, ninv, bit);
      } else {
        mont_sqr<S>(x, n, ninv);
        step(x, n, ninv, bit);
      }
      const int p = 32 * wi + b;
      if (p <= tw) {                // uniform: some lane checks at this bit
        if (p <= t && p >= 1) {     // x = Mont(y_{t-p})
          canon<S>(x, n);
          if (is_neg<S>(x, one, n) || (p == t && equal<S>(x, one))) sprp = 1u;
        }
      }
    } while (--b >= 0);
  }
  return sprp;
}

// Exponentiation + strong-test tail.  x enters as Mont(a) (the leading bit of
// N-1); one = Mont(1) = R mod N, computed once by the caller (the tail
// compares against it: recomputing it where p == t would repeat it for every
// distinct t among a wavefront's lanes).  `word(j)` returns 32-bit word j of
// N; it is read once per 32 exponent bits so the exponent never needs
// dynamically indexed registers.  res gets a^(N-1) mod N (canonical); returns
// FLAG_* bits.  A Step with kTop0 runs the top-limb-zero loop in wavefronts
// whose candidates all fit it and the folded loop in the others, so the
// results never depend on which loop ran.
template <int S, class WordFn, class Step>
PGX_FN uint32_t powm_verdict(uint32_t (&x)[S], const uint32_t (&one)[S], const uint32_t (&n)[S],
                             uint32_t ninv, int L, int t, WordFn word, const Step& step,
                             uint32_t (&res)[S]) {
  uint32_t sprp = 0;
  if (t == L - 1) {  // N - 1 = 2^t: y_0 = a
    canon<S>(x, n);
    sprp = (equal<S>(x, one) || is_neg<S>(x, one, n)) ? 1u : 0u;
  }
  // wavefront-uniform bounds: the longest N and the largest t of the wavefront
  const int Lw = (int)wave_max<12>((uint32_t)L);
  const int tw = (int)wave_max<12>((uint32_t)t);
  if (L != Lw) {  // bits Lw-2 .. L of N-1 are zero: start from Mont(1)
#pragma unroll
    for (int j = 0; j < S; ++j) x[j] = one[j];
  }
  if constexpr (Step::kTop0) {
    if (Lw <= Radix<S>::TOP0_MAX_BITS)
      sprp |= exponent_bits<true>(x, one, n, ninv, Lw, t, tw, word, step);
    else
      sprp |= exponent_bits<false>(x, one, n, ninv, Lw, t, tw, word, step);
  } else {
    sprp |= exponent_bits<false>(x, one, n, ninv, Lw, t, tw, word, step);
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

template <int S, bool kPreShift, bool kTop0 = false, class WordFn>
PGX_FN uint32_t test_base2(const uint32_t (&n)[S], uint32_t ninv, int L, int t, WordFn word,
                           uint32_t (&res)[S]) {
  uint32_t one[S], x[S];
  r_mod_n<S>(one, n, L);
#pragma unroll
  for (int j = 0; j < S; ++j) x[j] = one[j];
  dbl_full<S>(x, n);  // Mont(2)
  return powm_verdict<S>(x, one, n, ninv, L, t, word, MulBy2<S, kPreShift, kTop0>{}, res);
}

// Base 2 <= a < 2^8 (the extra Miller-Rabin bases).
template <int S, class WordFn>
PGX_FN uint32_t test_base(uint32_t a, const uint32_t (&n)[S], uint32_t ninv, int L, int t,
                          WordFn word, uint32_t (&res)[S]) {
  uint32_t one[S], x[S];
  r_mod_n<S>(one, n, L);
  // x = a * R mod N by double-and-add over the bits of a
#pragma unroll
  for (int j = 0; j < S; ++j) x[j] = 0;
  for (int b = 31 - __builtin_clz(a); b >= 0; --b) {
    dbl_full<S>(x, n);
    if ((a >> b) & 1u) add_full<S>(x, one, n);
  }
  const double inv_n = 1.0 / scaled<S>(n, L);
  return powm_verdict<S>(x, one, n, ninv, L, t, word, MulBySmall<S>{a, inv_n, L}, res);
}

// Whole test of one candidate given as W32 words (w).  Returns the verdict word
// of `mode` (PGX_VERDICT_INVALID if N is even, <= 1 or wider than the kernel)
// and writes the first base's a^(N-1) mod N as W32 words into res_w.
// kBase2Only: the base list is {2} (or the mode is FERMAT, which tests base 2
// whatever the list).  kTop0 (base 2 only): the squaring that skips a zero top
// limb in wavefronts whose candidates fit it (same results).
template <int S, bool kBase2Only, bool kPreShift = true, bool kTop0 = false, class WordFn>
PGX_FN uint32_t prp_candidate(const uint32_t (&w)[Radix<S>::W32], WordFn word, uint32_t mode,
                              const Bases& bases, uint32_t (&res_w)[Radix<S>::W32]) {
  static_assert(!kTop0 || kBase2Only, "the top-limb-zero squaring serves base 2");
  constexpr int W = Radix<S>::W32;
  const int L = bit_length_words<W>(w);
  uint32_t r[S];
#pragma unroll
  for (int j = 0; j < S; ++j) r[j] = 0;
  uint32_t v = PGX_VERDICT_INVALID;
  if ((w[0] & 1u) && L >= 2 && L <= Radix<S>::MAX_BITS) {
    const int t = ctz_nm1_words<W>(w);
    uint32_t n[S];
    from_words<S>(n, w);
    const uint32_t ninv = neg_inv32(w[0]);
    if (kBase2Only) {
      // base 2 alone: FERMAT mode returns the flags; the reference and strong
      // modes with the base list {2} return the count of bases passed
      const uint32_t f = test_base2<S, kPreShift, kTop0>(n, ninv, L, t, word, r);
      const uint32_t need = (mode == PGX_MODE_STRONG) ? FLAG_SPRP : FLAG_QUIRK;
      v = (mode == PGX_MODE_FERMAT) ? f : ((f & need) ? 1u : 0u);
    } else {
      const uint32_t need = (mode == PGX_MODE_STRONG) ? FLAG_SPRP : FLAG_QUIRK;
      uint32_t passed = 0;
      bool alive = true;  // every base so far passed
      for (uint32_t b = 0; b < bases.count; ++b) {
        uint32_t f;
        if (b == 0) {
          f = (bases.a[0] == 2u) ? test_base2<S, kPreShift>(n, ninv, L, t, word, r)
                                 : test_base<S>(bases.a[0], n, ninv, L, t, word, r);
        } else {
          uint32_t scratch[S];
          f = test_base<S>(bases.a[b], n, ninv, L, t, word, scratch);
        }
        alive = alive && (f & need);
        passed += alive ? 1u : 0u;
        // The loop ends for the whole wavefront at once: lanes that failed a
        // base run the later ones too and drop the results.  A per-lane exit
        // made the later exponentiations divergent regions, measured 3.7x
        // slower each than the same exponentiation with every lane active.
        if (!wave_any(alive)) break;
      }
      v = passed;
    }
  }
  to_words<S>(res_w, r);
  return v;
}

}  // namespace pgx
