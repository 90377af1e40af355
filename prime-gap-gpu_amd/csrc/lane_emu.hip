// Test-only host build of the per-lane routine of the PRP kernels.
//
// Compiles prp.hpp's templates for the CPU so the CPU test suite can check the
// exact arithmetic the kernels run (same source, host instantiation) against
// the oracle on machines without a GPU.  It is a separate library
// (libpgx_lane_emu.so) that only the tests load; libpgx.so has no CPU path.
#include <stdint.h>

#include "../../include/pgx.h"
#include "prp.hpp"

namespace {

// mode bits selecting the folded squaring with shifted column sums, and the
// folded squaring that skips a zero top limb
constexpr uint32_t kColumnShift = 0x100u;
constexpr uint32_t kTop0 = 0x200u;

template <int S>
uint32_t one_candidate(const uint32_t* x, uint32_t limbs, uint32_t mode, const pgx::Bases& bases,
                       uint32_t* res_out) {
  constexpr int W = pgx::Radix<S>::W32;
  uint32_t w[W], rw[W] = {};
  for (int j = 0; j < W; ++j) w[j] = (j < (int)limbs) ? x[j] : 0u;
  uint32_t spill = 0;
  for (uint32_t j = W; j < limbs; ++j) spill |= x[j];
  uint32_t v = PGX_VERDICT_INVALID;
  if (spill == 0) {
    auto word = [&](int j) { return x[j]; };
    if (mode == (PGX_MODE_FERMAT | kColumnShift))  // the squaring with column sums shifted
      v = pgx::prp_candidate<S, true, false>(w, word, PGX_MODE_FERMAT, bases, rw);
    else if (mode & kTop0)  // the squaring that skips a zero top limb (base 2)
      v = pgx::prp_candidate<S, true, true, true>(w, word, mode & ~kTop0, bases, rw);
    else if (mode == PGX_MODE_FERMAT || (bases.count == 1 && bases.a[0] == 2u))  // as launched
      v = pgx::prp_candidate<S, true>(w, word, mode, bases, rw);
    else
      v = pgx::prp_candidate<S, false>(w, word, mode, bases, rw);
  }
  if (res_out)
    for (uint32_t j = 0; j < limbs; ++j) res_out[j] = (j < (uint32_t)W) ? rw[j] : 0u;
  return v;
}

typedef uint32_t (*OneFn)(const uint32_t*, uint32_t, uint32_t, const pgx::Bases&, uint32_t*);

constexpr int kMaxS = 80;

template <int... Ss>
struct Table {
  OneFn fn[kMaxS + 1] = {};
  Table() { ((fn[Ss] = &one_candidate<Ss>), ...); }
};

// A subset of the kernels' limb counts: the host compiler is slow on the
// unrolled S^2 bodies, and the result does not depend on S (any S with
// 4N < 2^(32S) gives the same residue and verdict).
const Table<PGX_EMU_S_VALUES> table;

}  // namespace

extern "C" {

// Runs the kernel's per-candidate routine on the CPU with kernel limb count S
// (radix limbs; 0 = the smallest built one that holds every candidate).
// Candidate-major limbs in, verdict words (and residues, candidate-major, if
// res != NULL) out.  mode PGX_MODE_FERMAT | 0x100 runs the base-2 test with
// the other form of the folded squaring; a base-2 mode | 0x200 the squaring
// that skips a zero top limb (candidates of at most RB(S-1) - 1 bits).
// Returns the S used, or 0 if none is available.
uint32_t pgx_lane_emu(const uint32_t* limbs_le, size_t n, uint32_t limbs, uint32_t S,
                      uint32_t mode, const uint32_t* bases, uint32_t nbases, uint32_t* verdict,
                      uint32_t* res) {
  if (nbases > PGX_MAX_BASES) return 0;
  pgx::Bases b{};
  b.count = nbases;
  for (uint32_t k = 0; k < nbases; ++k) b.a[k] = bases[k];
  if (S == 0) {
    int max_bits = 0;
    for (size_t i = 0; i < n; ++i)
      for (int j = (int)limbs - 1; j >= 0; --j)
        if (limbs_le[i * limbs + j]) {
          int bits = 32 * j + 32 - __builtin_clz(limbs_le[i * limbs + j]);
          if (bits > max_bits) max_bits = bits;
          break;
        }
    for (S = 1; S <= (uint32_t)kMaxS; ++S)
      if (table.fn[S] && (S <= 31 ? 29 : 28) * (int)S - 4 >= max_bits) break;
  }
  if (S == 0 || S > (uint32_t)kMaxS || !table.fn[S]) return 0;
  for (size_t i = 0; i < n; ++i)
    verdict[i] = table.fn[S](limbs_le + i * limbs, limbs, mode, b, res ? res + i * limbs : nullptr);
  return S;
}

}  // extern "C"
