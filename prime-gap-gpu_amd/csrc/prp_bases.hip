// Base-parallel Miller-Rabin kernels over the survivors of a first base: one
// (candidate, base) pair per lane.  A translation unit of its own, compiled in
// the same parts as prp_kernels.hip (-DPGX_S_VALUES=...), so that adding these
// kernels leaves the code objects of the plain and persistent kernels as they
// were: the exponent loop's speed at one wavefront per SIMD depends on where
// the loop lands in the code object.
//
// Replaces the rounds > 1 loop of miller_rabin_t::miller_rabin
// (miller_rabin.h:216-248) for runners with several bases.
#include "launch.hpp"
#include "prp.hpp"

#ifndef PGX_S_VALUES
#error "PGX_S_VALUES (comma-separated kernel limb counts) must be defined"
#endif

namespace pgx {

// One (candidate, base) pair per lane over the survivors of a first base:
// work item i of c * bases.count (c = min(*count, n) columns listed in index)
// tests base bases.a[i / c] on column index[i % c], base-major so that a
// wavefront's lanes share their base, and sets bit i / c of mask[i % c] when
// the candidate passes.  The survivors are nearly all probable primes, which
// pass every base, so testing the bases side by side wastes next to no work
// and takes one exponentiation's latency where running them in turn takes one
// per base.
template <int S>
__global__ __launch_bounds__(kBlock) void k_prp_bases(PrpArgs a, uint32_t* mask) {
  const uint32_t c = min(*a.count, a.n);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if ((uint64_t)i >= (uint64_t)c * a.bases.count) return;
  const uint32_t b = i / c, k = i - b * c;
  constexpr int W = Radix<S>::W32;
  const uint32_t* col = a.cand + a.index[k];
  const size_t ld = a.ld;
  uint32_t w[W];
#pragma unroll
  for (int j = 0; j < W; ++j) w[j] = (j < (int)a.buf_limbs) ? col[j * ld] : 0u;
  uint32_t spill = 0;  // words above the kernel's width must be zero
  for (uint32_t j = W; j < a.check_limbs; ++j) spill |= col[j * ld];
  const int L = bit_length_words<W>(w);
  if (spill != 0 || !(w[0] & 1u) || L < 2 || L > Radix<S>::MAX_BITS) return;
  const int t = ctz_nm1_words<W>(w);
  uint32_t n[S], r[S];
  from_words<S>(n, w);
  const uint32_t f = test_base<S>(a.bases.a[b], n, neg_inv32(w[0]), L, t,
                                  [&](int j) { return col[j * ld]; }, r);
  const uint32_t need = (a.mode == PGX_MODE_STRONG) ? FLAG_SPRP : FLAG_QUIRK;
  if (f & need) atomicOr(&mask[k], 1u << b);
}

template <int S>
hipError_t launch_bases(const PrpArgs& a, uint32_t* mask, hipStream_t stream) {
  const uint64_t items = (uint64_t)a.n * a.bases.count;
  const uint64_t grid = (items + kBlock - 1) / kBlock;
  if (grid == 0) return hipSuccess;
  if (grid > 0x7FFFFFFFull) return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_prp_bases<S>), dim3((uint32_t)grid), dim3(kBlock), 0, stream, a, mask);
  return hipGetLastError();
}

template <int... Ss>
struct RegisterBases {
  RegisterBases() { ((bases_table()[Ss] = &launch_bases<Ss>), ...); }
};

static const RegisterBases<PGX_S_VALUES> registered_bases;

}  // namespace pgx
