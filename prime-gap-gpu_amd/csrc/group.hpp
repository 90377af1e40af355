// Several lanes per candidate: the PRP test of one number spread over a group
// of G = 2, 4, 8, 16 or 32 consecutive lanes, for launches with few
// candidates, where one lane per candidate leaves most SIMDs idle and the
// latency of a single exponentiation bounds the launch (Miller-Rabin bases on
// a few survivors, serving iterations of wide classes).
//
// A number is Sp = Q*G radix-2^28 limbs; lane r of a group owns limbs
// r*Q .. r*Q+Q-1 in registers.  Cross-lane traffic is DPP (no LDS memory):
// broadcasts inside the group (quad permutations for G <= 8, row_newbcast for
// a row of 16, and for 32 lanes row_newbcast plus row_bcast:15 or an LDS
// swizzle), and moves to the next / previous lane (quad permutations for G <=
// 4, row shifts by 1 lane for 8 and 16, wavefront shifts for 32, whose reads
// across a group boundary callers mask or select away).  The Montgomery
// product is coarsely integrated operand scanning over
// the group: for each limb a_i (broadcast from its owner) every lane adds
// a_i*b_j and m_i*n_j into its 64-bit accumulator slots, the retired position
// i leaves lane 0, and the window moves down one limb (the slot entering a
// lane's top comes from the next lane: only its low 28 bits, the carry above
// them stays in the lane).  Accumulator slots are not normalized during the
// product: each position takes at most 2*Sp products below 2^56 and one carry
// below 2^37, so Sp <= 127 keeps them below 2^64.  The result is then
// normalized: every lane settles its own slots, the carries move up one lane,
// and the single-unit carries left over are resolved for the whole group at
// once from two ballots (carry generated, lane all ones), as a binary
// addition of lane masks, instead of G-1 rounds of lane-to-lane passing.
//
// Values stay lazy as in the single-lane kernels (mont.hpp): R = 2^(28*Sp) >=
// 16N, a product of inputs < 4N is < 2N.  The verdict logic is the one of
// prp.hpp (miller_rabin.h:216-248), with group-wide reductions for the
// comparisons.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

#include "prp.hpp"

namespace pgx {
namespace grp {

constexpr int RB = 28;
constexpr uint32_t MASK = (1u << RB) - 1u;

constexpr int qp(int a, int b, int c, int d) { return a | (b << 2) | (c << 4) | (d << 6); }

// Every control used with dpp() reads a valid lane for every lane (quad
// permutations, row broadcasts and rotations), so no old value is needed and
// none is materialized.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
}

// DPP row shifts (lanes of a row of 16; lanes shifted in from outside the
// row read 0), row rotations, row broadcasts and wavefront shifts by one lane.
constexpr int kRowShl1 = 0x101, kRowShl4 = 0x104, kRowShr1 = 0x111, kRowShr4 = 0x114;
constexpr int kRowRor4 = 0x124, kRowRor8 = 0x128, kWaveShl1 = 0x130, kWaveShr1 = 0x138;
constexpr int kRowBcast15 = 0x142, kRowNewBcast0 = 0x150;

// Shifts: lanes with no source lane read 0 (bound_ctrl).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_row(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

// In groups of 8: the value of the lane 4 above (lanes 0-3 of the group) or 4
// below (lanes 4-7), i.e. the other quad of the group.
__device__ __forceinline__ uint32_t other_quad(uint32_t v) {
  const uint32_t up = dpp_row<kRowShl4>(v), down = dpp_row<kRowShr4>(v);
  return (threadIdx.x & 4u) ? down : up;
}

// LDS-crossbar swizzle inside groups of 32 lanes (no LDS memory): source
// lane ((lane & and_mask) | or_mask) ^ xor_mask.
template <int AND, int OR, int XOR>
__device__ __forceinline__ uint32_t swizzle32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, AND | (OR << 5) | (XOR << 10));
}

// Value of group lane O (compile-time) in every lane of the group.
template <int G, int O>
__device__ __forceinline__ uint32_t bcast(uint32_t v) {
  static_assert(G == 2 || G == 4 || G == 8 || G == 16 || G == 32,
                "groups of 2, 4, 8, 16 or 32 lanes");
  if constexpr (G == 4) {
    return dpp<qp(O, O, O, O)>(v);
  } else if constexpr (G == 2) {
    return dpp<qp(O, O, 2 + O, 2 + O)>(v);
  } else if constexpr (G == 8) {
    const uint32_t q = dpp<qp(O % 4, O % 4, O % 4, O % 4)>(v);  // lane O % 4 of each quad
    const uint32_t o = other_quad(q);
    return ((threadIdx.x & 4u) != 0) == (O >= 4) ? q : o;
  } else if constexpr (G == 16) {
    return dpp<kRowNewBcast0 + O>(v);  // the row is the group
  } else if constexpr (O < 16) {
    // lane O of each row, then the even row's copy into the odd row above it
    const uint32_t q = dpp<kRowNewBcast0 + O>(v);
    return (uint32_t)__builtin_amdgcn_update_dpp((int)q, (int)q, kRowBcast15, 0xA, 0xF, false);
  } else {
    return swizzle32<0, O, 0>(v);
  }
}

// Value of lane r+1 (the top lane gets its own value, or the next group's, or
// 0 back: callers mask it).
template <int G>
__device__ __forceinline__ uint32_t from_next(uint32_t v) {
  if constexpr (G == 32) return dpp_row<kWaveShl1>(v);
  else if constexpr (G == 8 || G == 16) return dpp_row<kRowShl1>(v);
  else if constexpr (G == 4) return dpp<qp(1, 2, 3, 3)>(v);
  else return dpp<qp(1, 1, 3, 3)>(v);
}

// Value of lane r-1 (lane 0 gets its own value, or the previous group's, or 0
// back: callers mask it).
template <int G>
__device__ __forceinline__ uint32_t from_prev(uint32_t v) {
  if constexpr (G == 32) return dpp_row<kWaveShr1>(v);
  else if constexpr (G == 8 || G == 16) return dpp_row<kRowShr1>(v);
  else if constexpr (G == 4) return dpp<qp(0, 0, 1, 2)>(v);
  else return dpp<qp(0, 0, 2, 2)>(v);
}

template <int G>
__device__ __forceinline__ uint64_t from_next64(uint64_t v) {
  return (uint64_t)from_next<G>((uint32_t)v) | ((uint64_t)from_next<G>((uint32_t)(v >> 32)) << 32);
}

template <int G>
__device__ __forceinline__ int64_t from_prev64(int64_t v) {
  const uint64_t u = (uint64_t)v;
  return (int64_t)((uint64_t)from_prev<G>((uint32_t)u) |
                   ((uint64_t)from_prev<G>((uint32_t)(u >> 32)) << 32));
}

template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
  return (uint64_t)dpp<CTRL>((uint32_t)v) | ((uint64_t)dpp<CTRL>((uint32_t)(v >> 32)) << 32);
}

template <int G>
__device__ __forceinline__ uint32_t group_or(uint32_t v) {
  v |= dpp<qp(1, 0, 3, 2)>(v);
  if constexpr (G >= 4) v |= dpp<qp(2, 3, 0, 1)>(v);
  if constexpr (G == 8) v |= other_quad(v);
  if constexpr (G >= 16) {  // the four quads of the row
    v |= dpp<kRowRor4>(v);
    v |= dpp<kRowRor8>(v);
  }
  if constexpr (G == 32) v |= swizzle32<31, 0, 16>(v);  // the other row of the half
  return v;
}

template <int G>
__device__ __forceinline__ double group_sum(double v) {
  auto add = [&](uint64_t o) { v += __longlong_as_double((long long)o); };
  auto bits = [&]() { return (uint64_t)__double_as_longlong(v); };
  add(dpp64<qp(1, 0, 3, 2)>(bits()));
  if constexpr (G >= 4) add(dpp64<qp(2, 3, 0, 1)>(bits()));
  if constexpr (G == 8) {
    const uint64_t u = bits();
    add((uint64_t)other_quad((uint32_t)u) | ((uint64_t)other_quad((uint32_t)(u >> 32)) << 32));
  }
  if constexpr (G >= 16) {
    add(dpp64<kRowRor4>(bits()));
    add(dpp64<kRowRor8>(bits()));
  }
  if constexpr (G == 32) {
    const uint64_t u = bits();
    add((uint64_t)swizzle32<31, 0, 16>((uint32_t)u) |
        ((uint64_t)swizzle32<31, 0, 16>((uint32_t)(u >> 32)) << 32));
  }
  return v;
}

// Lanes that are the top lane of their group, as a wavefront lane mask.
template <int G>
constexpr uint64_t top_lanes() {
  uint64_t m = 0;
  for (int i = G - 1; i < 64; i += G) m |= 1ull << i;
  return m;
}

template <int G>
__device__ __forceinline__ uint32_t lane_id() {
  return threadIdx.x & (G - 1);
}

// Signed limbs t (any int64) -> normalized limbs r, carries passed through the
// group; returns the carry out of the whole number (in every lane).
template <int Q, int G>
__device__ __forceinline__ int64_t settle(uint32_t (&r)[Q], const int64_t (&t)[Q]) {
  const uint32_t lane = lane_id<G>();
  int64_t c = 0;
#pragma unroll
  for (int s = 0; s < Q; ++s) {
    const int64_t v = t[s] + c;
    r[s] = (uint32_t)v & MASK;
    c = v >> RB;
  }
  int64_t top = c;  // carries leaving this lane, summed over the rounds
#pragma unroll
  for (int round = 1; round < G; ++round) {
    int64_t in = from_prev64<G>(c);
    if (lane == 0) in = 0;
#pragma unroll
    for (int s = 0; s < Q; ++s) {
      const int64_t v = (int64_t)r[s] + in;
      r[s] = (uint32_t)v & MASK;
      in = v >> RB;
    }
    c = in;
    top += c;
  }
  return (int64_t)((uint64_t)bcast<G, G - 1>((uint32_t)top) |
                   ((uint64_t)bcast<G, G - 1>((uint32_t)((uint64_t)top >> 32)) << 32));
}

// Unsigned accumulator slots -> normalized limbs (the carry out is zero for
// every value this file produces).  Every lane settles its slots, the carry
// of each lane (< 2^38) enters the lane above, which leaves carries of 0 or 1
// (Q >= 2 limbs of 28 bits absorb 2^38).  Those are resolved at once: a carry
// reaches lane r when some lane below it generated one and every lane in
// between is all ones, i.e. the carries into the lanes are ((g << 1) + p) ^ p
// over the ballot masks g (generated) and p (all ones), neither taken from a
// group's top lane so that nothing crosses into the next group.
template <int Q, int G>
__device__ __forceinline__ void settle_u(uint32_t (&r)[Q], const uint64_t (&t)[Q]) {
  const uint32_t lane = lane_id<G>();
  uint64_t c = 0;
#pragma unroll
  for (int s = 0; s < Q; ++s) {
    const uint64_t v = t[s] + c;
    r[s] = (uint32_t)v & MASK;
    c = v >> RB;
  }
  if constexpr (G >= 4) {
    static_assert(Q >= 2, "two limbs per lane absorb a carry below 2^38");
    uint64_t in = (uint64_t)from_prev64<G>((int64_t)c);
    if (lane == 0) in = 0;
    uint32_t ones = 0;  // zero when every limb of the lane is all ones
#pragma unroll
    for (int s = 0; s < Q; ++s) {
      const uint64_t v = (uint64_t)r[s] + in;
      r[s] = (uint32_t)v & MASK;
      in = v >> RB;
      ones |= r[s] ^ MASK;
    }
    constexpr uint64_t kTop = top_lanes<G>();
    const uint64_t g = __ballot(in != 0) & ~kTop;
    const uint64_t p = __ballot(ones == 0) & ~kTop;
    const uint64_t cin = ((g << 1) + p) ^ p;
    uint32_t k = (uint32_t)(cin >> (threadIdx.x & 63u)) & 1u;
#pragma unroll
    for (int s = 0; s < Q; ++s) {
      const uint32_t v = r[s] + k;
      r[s] = v & MASK;
      k = v >> RB;
    }
    return;
  }
#pragma unroll
  for (int round = 1; round < G; ++round) {
    uint64_t in = (uint64_t)from_prev64<G>((int64_t)c);
    if (lane == 0) in = 0;
#pragma unroll
    for (int s = 0; s < Q; ++s) {
      const uint64_t v = (uint64_t)r[s] + in;
      r[s] = (uint32_t)v & MASK;
      in = v >> RB;
    }
    c = in;
  }
}

// One position of the product scan: add a_I * b and m_I * n, retire
// position I from lane 0, move the window down one limb.  I is a template
// argument so the owner lane of a_I and its slot are constants.  Every lane
// splits its lowest slot (position I + lane*Q) at bit RB: the carry stays in
// the lane, added to the next slot, and only the low RB bits move to the top
// slot of the lane below (one 32-bit DPP move of the slot's low word, masked
// by the receiver; lane 0's low bits are zero by the choice of m and leave the
// window).  low_keep is 0 in the top lane of the group, which has nothing
// above it, and MASK elsewhere.
template <int Q, int G, int I>
__device__ __forceinline__ void scan_step(uint64_t (&acc)[Q], const uint32_t (&a)[Q],
                                          const uint32_t (&b)[Q], const uint32_t (&n)[Q],
                                          uint32_t ninv, uint32_t low_keep) {
  static_assert(Q > 1, "the carry of the lowest slot stays in the lane");
  const uint32_t ai = bcast<G, I / Q>(a[I % Q]);
#pragma unroll
  for (int s = 0; s < Q; ++s) acc[s] += (uint64_t)ai * b[s];
  const uint32_t m = bcast<G, 0>(((uint32_t)acc[0] * ninv) & MASK);
#pragma unroll
  for (int s = 0; s < Q; ++s) acc[s] += (uint64_t)m * n[s];
  acc[1] += acc[0] >> RB;
  const uint32_t nx = from_next<G>((uint32_t)acc[0]) & low_keep;
#pragma unroll
  for (int s = 0; s + 1 < Q; ++s) acc[s] = acc[s + 1];
  acc[Q - 1] = nx;
}

template <int Q, int G, int... Is>
__device__ __forceinline__ void scan_all(std::integer_sequence<int, Is...>, uint64_t (&acc)[Q],
                                         const uint32_t (&a)[Q], const uint32_t (&b)[Q],
                                         const uint32_t (&n)[Q], uint32_t ninv,
                                         uint32_t low_keep) {
  (scan_step<Q, G, Is>(acc, a, b, n, ninv, low_keep), ...);
}

// r <- a * b / R mod N (a, b < 4N with normalized limbs; r < 2N).  r may
// alias a or b.
template <int Q, int G>
__device__ __forceinline__ void mont_mul(uint32_t (&r)[Q], const uint32_t (&a)[Q],
                                         const uint32_t (&b)[Q], const uint32_t (&n)[Q],
                                         uint32_t ninv) {
  uint64_t acc[Q];
#pragma unroll
  for (int s = 0; s < Q; ++s) acc[s] = 0;
  const uint32_t low_keep = (lane_id<G>() == G - 1) ? 0u : MASK;
  scan_all<Q, G>(std::make_integer_sequence<int, Q * G>{}, acc, a, b, n, ninv, low_keep);
  settle_u<Q, G>(r, acc);
}

// x <- 2x for bit == 1 (lazy: x < 2N becomes < 4N).
template <int Q, int G>
__device__ __forceinline__ void dbl_lazy(uint32_t (&x)[Q], uint32_t bit) {
  const uint32_t lane = lane_id<G>();
  uint32_t in = from_prev<G>(x[Q - 1] >> (RB - 1));
  if (lane == 0) in = 0;
  in &= bit;
#pragma unroll
  for (int s = Q - 1; s >= 1; --s) x[s] = ((x[s] << bit) | (x[s - 1] >> (RB - bit))) & MASK;
  x[0] = ((x[0] << bit) & MASK) | in;
}

// x <- x - N when x >= N.
template <int Q, int G>
__device__ __forceinline__ void sub_if_ge(uint32_t (&x)[Q], const uint32_t (&n)[Q]) {
  int64_t t[Q];
#pragma unroll
  for (int s = 0; s < Q; ++s) t[s] = (int64_t)x[s] - (int64_t)n[s];
  uint32_t d[Q];
  const bool lt = settle<Q, G>(d, t) < 0;
#pragma unroll
  for (int s = 0; s < Q; ++s) x[s] = lt ? x[s] : d[s];
}

template <int Q, int G>
__device__ __forceinline__ void canon(uint32_t (&x)[Q], const uint32_t (&n)[Q]) {
  sub_if_ge<Q, G>(x, n);
  sub_if_ge<Q, G>(x, n);
  sub_if_ge<Q, G>(x, n);
}

template <int Q, int G>
__device__ __forceinline__ bool equal(const uint32_t (&a)[Q], const uint32_t (&b)[Q]) {
  uint32_t diff = 0;
#pragma unroll
  for (int s = 0; s < Q; ++s) diff |= a[s] ^ b[s];
  return group_or<G>(diff) == 0;
}

// a + b == N (canonical a, b).
template <int Q, int G>
__device__ __forceinline__ bool is_neg(const uint32_t (&a)[Q], const uint32_t (&b)[Q],
                                       const uint32_t (&n)[Q]) {
  int64_t t[Q];
#pragma unroll
  for (int s = 0; s < Q; ++s) t[s] = (int64_t)a[s] + (int64_t)b[s];
  uint32_t sum[Q];
  const int64_t c = settle<Q, G>(sum, t);
  return c == 0 && equal<Q, G>(sum, n);
}

// v <- 2^bit (bit < 28*Sp).
template <int Q, int G>
__device__ __forceinline__ void set_pow2(uint32_t (&v)[Q], int bit) {
  const int q = bit / RB - (int)lane_id<G>() * Q, rr = bit % RB;
#pragma unroll
  for (int s = 0; s < Q; ++s) v[s] = (s == q) ? (1u << rr) : 0u;
}



// Sum of v_j 2^(28 j - L) over the group's limbs, as a double.
template <int Q, int G>
__device__ __forceinline__ double scaled(const uint32_t (&v)[Q], int L) {
  const int base = (int)lane_id<G>() * Q;
  double d = 0.0;
#pragma unroll
  for (int s = 0; s < Q; ++s) d += ldexp((double)v[s], RB * (base + s) - L);
  return group_sum<G>(d);
}

// x <- a x mod N for an ordinary integer 1 <= a <= 2^16 (x < 4N in, < 2N
// out); the quotient estimate of mont.hpp's mul_small, summed over the group
// (a x / N < 2^18: the 2^-40 bias still moves it by less than 2^-21).
template <int Q, int G>
__device__ __forceinline__ void mul_small(uint32_t (&x)[Q], uint32_t a, const uint32_t (&n)[Q],
                                          double inv_n, int L) {
  const double qd = (double)a * scaled<Q, G>(x, L) * inv_n * (1.0 - 0x1p-40);
  const uint32_t q = (qd > 0.0) ? (uint32_t)qd : 0u;
  int64_t t[Q];
#pragma unroll
  for (int s = 0; s < Q; ++s)
    t[s] = (int64_t)((uint64_t)x[s] * a) - (int64_t)((uint64_t)n[s] * q);
  settle<Q, G>(x, t);
}

// one <- R mod N: 2^(L-1) (< N) doubled RB*Sp - L + 1 times, 16 doublings per
// mul_small, then one conditional subtraction.
template <int Q, int G>
__device__ __forceinline__ void r_mod_n(uint32_t (&one)[Q], const uint32_t (&n)[Q], int L) {
  set_pow2<Q, G>(one, L - 1);
  const double inv_n = 1.0 / scaled<Q, G>(n, L);
  for (int steps = RB * Q * G - (L - 1); steps > 0; steps -= 16)
    mul_small<Q, G>(one, 1u << (steps < 16 ? steps : 16), n, inv_n, L);
  sub_if_ge<Q, G>(one, n);
}

// Exponentiation of x = Mont(base) by N-1 with the strong-test checks (the
// group form of powm_verdict in prp.hpp, with the same wavefront-uniform bit
// index); one = R mod N from the caller.  res <- base^(N-1) mod N.
template <int Q, int G, class WordFn>
__device__ __forceinline__ uint32_t powm_verdict(uint32_t (&x)[Q], const uint32_t (&one)[Q],
                                                 const uint32_t (&n)[Q], uint32_t ninv, int L,
                                                 int t, uint32_t base, double inv_n, WordFn word,
                                                 uint32_t (&res)[Q]) {
  uint32_t sprp = 0;
  if (t == L - 1) {
    canon<Q, G>(x, n);
    sprp = (equal<Q, G>(x, one) || is_neg<Q, G>(x, one, n)) ? 1u : 0u;
  }
  // every lane of a group holds the same L and t
  const int Lw = (int)wave_max<12>((uint32_t)L);
  const int tw = (int)wave_max<12>((uint32_t)t);
  if (L != Lw) {
#pragma unroll
    for (int s = 0; s < Q; ++s) x[s] = one[s];
  }
  const int top = (Lw - 2) >> 5;
  uint32_t w_next = word(top);
  for (int wi = top; wi >= 0; --wi) {
    const uint32_t w = (wi == 0) ? (w_next & ~1u) : w_next;
    if (wi > 0) w_next = word(wi - 1);
    int b = (wi == top) ? ((Lw - 2) & 31) : 31;
    do {
      const uint32_t bit = (w >> b) & 1u;
      mont_mul<Q, G>(x, x, x, n, ninv);
      if (base == 2u) dbl_lazy<Q, G>(x, bit);
      else if (bit) mul_small<Q, G>(x, base, n, inv_n, L);
      const int p = 32 * wi + b;
      if (p <= tw) {
        if (p <= t && p >= 1) {
          canon<Q, G>(x, n);
          if (is_neg<Q, G>(x, one, n) || (p == t && equal<Q, G>(x, one))) sprp = 1u;
        }
      }
    } while (--b >= 0);
  }
  uint32_t unit[Q];
#pragma unroll
  for (int s = 0; s < Q; ++s) unit[s] = (s == 0 && lane_id<G>() == 0) ? 1u : 0u;
  mont_mul<Q, G>(res, x, unit, n, ninv);
  sub_if_ge<Q, G>(res, n);
  uint32_t one_diff = 0, m1_diff = 0;
#pragma unroll
  for (int s = 0; s < Q; ++s) {
    const uint32_t nm1 = (s == 0 && lane_id<G>() == 0) ? n[s] - 1u : n[s];
    one_diff |= res[s] ^ unit[s];
    m1_diff |= res[s] ^ nm1;
  }
  const bool is_one = group_or<G>(one_diff) == 0, is_m1 = group_or<G>(m1_diff) == 0;
  uint32_t f = 0;
  if (is_one) f |= FLAG_FERMAT;
  if (sprp) f |= FLAG_SPRP;
  if (sprp || (t >= 2 && is_m1)) f |= FLAG_QUIRK;
  return f;
}

}  // namespace grp
}  // namespace pgx
