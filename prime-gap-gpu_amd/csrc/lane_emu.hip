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

template <int S>
uint32_t one_candidate(const uint32_t* x, uint32_t limbs, uint32_t mode, const uint32_t* bases,
                       uint32_t nbases, uint32_t* res_out) {
  uint32_t n[S];
  for (int j = 0; j < S; ++j) n[j] = (j < (int)limbs) ? x[j] : 0u;
  uint32_t spill = 0;
  for (uint32_t j = S; j < limbs; ++j) spill |= x[j];
  const int L = pgx::bit_length<S>(n);
  uint32_t r[S] = {};
  uint32_t v = PGX_VERDICT_INVALID;
  if (spill == 0 && pgx::fits<S>(n, L)) {
    const int t = pgx::ctz_nm1<S>(n);
    const uint32_t ninv = pgx::neg_inv32(n[0]);
    auto limb = [&](int j) { return x[j]; };
    if (mode == PGX_MODE_FERMAT) {
      v = pgx::test_base2<S>(n, ninv, L, t, limb, r);
    } else {
      const uint32_t need = (mode == PGX_MODE_STRONG) ? pgx::FLAG_SPRP : pgx::FLAG_QUIRK;
      uint32_t passed = 0;
      for (uint32_t b = 0; b < nbases; ++b) {
        uint32_t scratch[S];
        uint32_t(&dst)[S] = b == 0 ? r : scratch;
        uint32_t f = (bases[b] == 2u) ? pgx::test_base2<S>(n, ninv, L, t, limb, dst)
                                      : pgx::test_base<S>(bases[b], n, ninv, L, t, limb, dst);
        if (!(f & need)) break;
        ++passed;
      }
      v = passed;
    }
  }
  if (res_out)
    for (uint32_t j = 0; j < limbs; ++j) res_out[j] = (j < (uint32_t)S) ? r[j] : 0u;
  return v;
}

typedef uint32_t (*OneFn)(const uint32_t*, uint32_t, uint32_t, const uint32_t*, uint32_t,
                          uint32_t*);

template <int... Ss>
struct Table {
  OneFn fn[PGX_MAX_KERNEL_LIMBS + 1] = {};
  Table() { ((fn[Ss] = &one_candidate<Ss>), ...); }
};

// A subset of the kernels' limb counts: the host compiler is slow on the
// unrolled S^2 bodies, and the result does not depend on S (any S with
// 4N < 2^(32S) gives the same residue and verdict).
const Table<PGX_EMU_S_VALUES> table;

}  // namespace

extern "C" {

// Runs the kernel's per-candidate routine on the CPU with kernel limb count S
// (0 = the smallest that fits max_bits).  Candidate-major limbs in, verdict
// words (and residues, candidate-major, if res != NULL) out.  Returns the S used,
// or 0 if none is available.
uint32_t pgx_lane_emu(const uint32_t* limbs_le, size_t n, uint32_t limbs, uint32_t S,
                      uint32_t mode, const uint32_t* bases, uint32_t nbases, uint32_t* verdict,
                      uint32_t* res) {
  if (S == 0) {
    uint32_t max_bits = 0;
    for (size_t i = 0; i < n; ++i)
      for (int j = (int)limbs - 1; j >= 0; --j)
        if (limbs_le[i * limbs + j]) {
          uint32_t b = 32u * j + 32u - __builtin_clz(limbs_le[i * limbs + j]);
          if (b > max_bits) max_bits = b;
          break;
        }
    S = (max_bits + 2 + 31) / 32;
    while (S <= PGX_MAX_KERNEL_LIMBS && !table.fn[S]) ++S;
  }
  if (S == 0 || S > PGX_MAX_KERNEL_LIMBS || !table.fn[S]) return 0;
  for (size_t i = 0; i < n; ++i)
    verdict[i] = table.fn[S](limbs_le + i * limbs, limbs, mode, bases, nbases,
                             res ? res + i * limbs : nullptr);
  return S;
}

}  // extern "C"
