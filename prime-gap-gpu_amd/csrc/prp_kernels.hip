// Plain PRP kernels: one lane per candidate, one launch per batch.
//
// Compiled several times with -DPGX_S_VALUES=a,b,... so that the fully
// unrolled instantiations (code size and compile time grow as S^2) build in
// parallel; each object registers its launchers in prp_table().
//
// Replaces kernel_miller_rabin<params> (miller_rabin.h:252-270): instead of
// cgbn_load into a TPI-thread group and miller_rabin() over CGBN, each lane
// loads its candidate's limbs from a limb-interleaved array (coalesced rows),
// runs the Montgomery exponentiation in registers and stores its verdict word
// (and optionally the residue) to limb-interleaved outputs.
#include "launch.hpp"
#include "prp.hpp"

#ifndef PGX_S_VALUES
#error "PGX_S_VALUES (comma-separated kernel limb counts) must be defined"
#endif

namespace pgx {

// Form of the folded squaring of the plain kernels (mont_sqr_dbl): operands
// shifted at every width.  (Shifting the column sums instead measured faster
// in the plain launches of 12..24 limbs while the strong-test tail recomputed
// R mod N; since that is computed once per candidate, shifting the operands
// is as fast or faster there too: +9% at 14 limbs, +2% at 17.)
template <int S>
constexpr bool plain_pre_shift() {
  return true;
}

// Tests the candidate in column p of a.cand and writes column p of the outputs.
template <int S, bool kBase2Only, bool kPreShift>
__device__ __forceinline__ void test_column(const PrpArgs& a, uint32_t p) {
  constexpr int W = Radix<S>::W32;
  const uint32_t* c = a.cand + p;
  const size_t ld = a.ld;
  uint32_t w[W];
#pragma unroll
  for (int j = 0; j < W; ++j) w[j] = (j < (int)a.buf_limbs) ? c[j * ld] : 0u;
  uint32_t spill = 0;  // words above the kernel's width must be zero
  for (uint32_t j = W; j < a.check_limbs; ++j) spill |= c[j * ld];
  uint32_t rw[W];
  uint32_t v = PGX_VERDICT_INVALID;
  if (spill == 0) {
    // base 2 alone: the top-limb-zero squaring where a wavefront's candidates fit it
    v = prp_candidate<S, kBase2Only, kPreShift, kBase2Only>(
        w, [&](int j) { return c[j * ld]; }, a.mode, a.bases, rw);
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j) rw[j] = 0;
  }
  a.verdict[p] = v;
  if (a.residue) {
    uint32_t* o = a.residue + p;
    const size_t rld = a.res_ld;
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (j < (int)a.buf_limbs) o[j * rld] = rw[j];
    for (uint32_t j = W; j < a.buf_limbs; ++j) o[j * rld] = 0u;
  }
}

// Number of work items: a.n, or *a.count when the count lives on the device
// (then a.n is only the bound the grid was sized for).
__device__ __forceinline__ uint32_t work_items(const PrpArgs& a) {
  return a.count ? min(*a.count, a.n) : a.n;
}

// One launch per batch: thread i tests work item i.
template <int S, bool kBase2Only>
__global__ __launch_bounds__(kBlock) void k_prp_plain(PrpArgs a) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= work_items(a)) return;
  test_column<S, kBase2Only, plain_pre_shift<S>()>(a, a.index ? a.index[i] : i);
}

// Persistent: a fixed grid of resident wavefronts takes `chunk` work items at a
// time (chunk a multiple of 64) from *counter until the batch is exhausted.
// The counter only grows, so every wavefront reaches the exit.
template <int S, bool kBase2Only>
__global__ __launch_bounds__(kBlock) void k_prp_persistent(PrpArgs a, uint32_t* counter,
                                                           uint32_t chunk) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t n = work_items(a);
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(counter, chunk);
    base = __builtin_amdgcn_readfirstlane(base);
    if (base >= n) break;
    for (uint32_t off = lane; off < chunk; off += 64) {
      const uint32_t i = base + off;
      if (i < n) test_column<S, kBase2Only, true>(a, a.index ? a.index[i] : i);
    }
  }
}

// The base-2 kernels serve FERMAT mode and every mode whose base list is {2}
// (the reference's default, ROUNDS = 1): fewer registers than the general
// Miller-Rabin kernels, which carry the small-base step.
inline bool base2_only(const PrpArgs& a) {
  return a.mode == PGX_MODE_FERMAT || (a.bases.count == 1 && a.bases.a[0] == 2u);
}

template <int S>
hipError_t launch_plain(const PrpArgs& a, hipStream_t stream) {
  const uint32_t grid = (a.n + kBlock - 1) / kBlock;
  if (base2_only(a))
    hipLaunchKernelGGL((k_prp_plain<S, true>), dim3(grid), dim3(kBlock), 0, stream, a);
  else
    hipLaunchKernelGGL((k_prp_plain<S, false>), dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

template <int S>
hipError_t launch_persistent(const PrpArgs& a, uint32_t* counter, const PersistShape& sh,
                             hipStream_t stream) {
  if (base2_only(a))
    hipLaunchKernelGGL((k_prp_persistent<S, true>), dim3(sh.blocks), dim3(sh.threads), 0, stream,
                       a, counter, sh.chunk);
  else
    hipLaunchKernelGGL((k_prp_persistent<S, false>), dim3(sh.blocks), dim3(sh.threads), 0,
                       stream, a, counter, sh.chunk);
  return hipGetLastError();
}

template <int... Ss>
struct Register {
  Register() {
    ((prp_table()[Ss] = &launch_plain<Ss>), ...);
    ((persist_table()[Ss] = &launch_persistent<Ss>), ...);
  }
};

static const Register<PGX_S_VALUES> registered;

}  // namespace pgx
