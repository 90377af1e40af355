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

template <int S, bool kBase2Only>
__global__ __launch_bounds__(kBlock) void k_prp_plain(PrpArgs a) {
  constexpr int W = Radix<S>::W32;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t* c = a.cand + i;
  const size_t ld = a.ld;
  uint32_t w[W];
#pragma unroll
  for (int j = 0; j < W; ++j) w[j] = (j < (int)a.buf_limbs) ? c[j * ld] : 0u;
  uint32_t spill = 0;  // words above the kernel's width must be zero
  for (uint32_t j = W; j < a.buf_limbs; ++j) spill |= c[j * ld];
  uint32_t rw[W];
  uint32_t v = PGX_VERDICT_INVALID;
  if (spill == 0) {
    v = prp_candidate<S, kBase2Only>(w, [&](int j) { return c[j * ld]; }, a.mode, a.bases, rw);
  } else {
#pragma unroll
    for (int j = 0; j < W; ++j) rw[j] = 0;
  }
  a.verdict[i] = v;
  if (a.residue) {
    uint32_t* o = a.residue + i;
    const size_t rld = a.res_ld;
#pragma unroll
    for (int j = 0; j < W; ++j)
      if (j < (int)a.buf_limbs) o[j * rld] = rw[j];
    for (uint32_t j = W; j < a.buf_limbs; ++j) o[j * rld] = 0u;
  }
}

template <int S>
hipError_t launch_plain(const PrpArgs& a, hipStream_t stream) {
  const uint32_t grid = (a.n + kBlock - 1) / kBlock;
  if (a.mode == PGX_MODE_FERMAT)
    hipLaunchKernelGGL((k_prp_plain<S, true>), dim3(grid), dim3(kBlock), 0, stream, a);
  else
    hipLaunchKernelGGL((k_prp_plain<S, false>), dim3(grid), dim3(kBlock), 0, stream, a);
  return hipGetLastError();
}

template <int... Ss>
struct Register {
  Register() { ((prp_table()[Ss] = &launch_plain<Ss>), ...); }
};

static const Register<PGX_S_VALUES> registered;

}  // namespace pgx
