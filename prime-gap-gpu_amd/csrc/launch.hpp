// Internal launch interface shared by the kernel translation units and the
// runtime.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pgx.h"

namespace pgx {

struct Bases {
  uint32_t count;
  uint32_t a[PGX_MAX_BASES];
};

// Arguments of one plain (one launch per batch) PRP launch.
struct PrpArgs {
  const uint32_t* cand;  // limb-interleaved: limb j of candidate i at cand[j * ld + i]
  uint32_t ld;
  uint32_t n;
  uint32_t buf_limbs;  // rows stored in cand (limbs beyond the kernel's S must be zero)
  uint32_t mode;       // PGX_MODE_*
  Bases bases;
  uint32_t* verdict;  // n words
  uint32_t* residue;  // buf_limbs rows of res_ld words, or null
  uint32_t res_ld;
};

typedef hipError_t (*PrpLaunchFn)(const PrpArgs&, hipStream_t);

// Launchers indexed by kernel limb count S (1..PGX_MAX_KERNEL_LIMBS); an entry
// is null when no translation unit instantiated that S.
PrpLaunchFn* prp_table();

// Threads per block of the plain kernel (4 wavefronts).
constexpr uint32_t kBlock = 256;

// Smallest kernel limb count for candidates of at most max_bits bits.
inline uint32_t kernel_limbs_for_bits(uint32_t max_bits) { return (max_bits + 2 + 31) / 32; }

}  // namespace pgx
