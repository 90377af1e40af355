// Internal launch interface shared by the kernel translation units and the
// runtime.  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pgx_defs.h"
#include "prp.hpp"

namespace pgx {

// Arguments of one plain (one launch per batch) PRP launch.
// Work item i tests column p = index ? index[i] : i of cand and writes column p
// of verdict / residue.
struct PrpArgs {
  const uint32_t* cand;  // limb-interleaved: limb j of column p at cand[j * ld + p]
  uint32_t ld;
  uint32_t n;          // work items (upper bound when count is set)
  uint32_t buf_limbs;  // rows stored in cand (rows beyond the kernel's width must be zero)
  // Rows [kernel width, check_limbs) are read and a column with a nonzero word
  // there gets PGX_VERDICT_INVALID: buf_limbs where max_bits is the caller's
  // promise (device-resident inputs), 0 where the library derived max_bits
  // from the candidates themselves (no extra reads).
  uint32_t check_limbs;
  uint32_t mode;       // PGX_MODE_*
  Bases bases;
  uint32_t* verdict;         // indexed by column
  uint32_t* residue;         // buf_limbs rows of res_ld words, or null
  uint32_t res_ld;
  const uint32_t* index;     // work item -> column, or null (identity)
  const uint32_t* count;     // device-resident work-item count, or null (use n)
};

// Grid of a persistent launch.
struct PersistShape {
  uint32_t blocks;   // resident workgroups
  uint32_t threads;  // threads per workgroup (multiple of 64, <= kBlock)
  uint32_t chunk;    // work items a wavefront takes at a time (multiple of 64)
};

typedef hipError_t (*PrpLaunchFn)(const PrpArgs&, hipStream_t);
// Base-parallel launch over a device-resident work list (index, count): with
// c = min(*count, n), work item i < c * bases.count tests base
// bases.a[i / c] on column index[i % c] and sets bit i / c of mask[i % c]
// when the candidate passes it (mask cleared by the caller).
typedef hipError_t (*BasesLaunchFn)(const PrpArgs&, uint32_t* mask, hipStream_t);
typedef hipError_t (*PersistLaunchFn)(const PrpArgs&, uint32_t* counter, const PersistShape&,
                                      hipStream_t);

// Kernel limb counts: S radix-2^RB limbs hold candidates of up to
// RB*S - 4 bits (RB = 29 up to 31 limbs, 28 above).
constexpr int kMaxKernelLimbs = 80;
inline int kernel_capacity_bits(int S) { return (S <= 31 ? 29 : 28) * S - 4; }

// Launchers indexed by kernel limb count S (1..kMaxKernelLimbs); an entry is
// null when no translation unit instantiated that S.
PrpLaunchFn* prp_table();
PersistLaunchFn* persist_table();
BasesLaunchFn* bases_table();

// Smallest instantiated S whose capacity holds max_bits, or 0.
inline int kernel_for_bits(uint32_t max_bits) {
  for (int S = 1; S <= kMaxKernelLimbs; ++S)
    if (prp_table()[S] && kernel_capacity_bits(S) >= (int)max_bits) return S;
  return 0;
}

// Launchers of the G-lanes-per-candidate kernels (G = 2, 4, 8, 16 or 32),
// indexed by their radix-2^28 limb count Sp = Q * G (capacity 28*Sp - 4 bits).
constexpr int kMaxGroupLimbs = 96;
PrpLaunchFn* group_table(int G);

inline bool valid_group(int G) { return G == 2 || G == 4 || G == 8 || G == 16 || G == 32; }

// Smallest instantiated Sp of the G-lane kernels that holds max_bits, or 0.
inline int group_width_for_bits(uint32_t max_bits, int G) {
  if (!valid_group(G)) return 0;
  for (int Sp = 1; Sp <= kMaxGroupLimbs; ++Sp)
    if (group_table(G)[Sp] && 28 * Sp - 4 >= (int)max_bits) return Sp;
  return 0;
}

// Threads per block of the plain kernel (4 wavefronts).
constexpr uint32_t kBlock = 256;

}  // namespace pgx
