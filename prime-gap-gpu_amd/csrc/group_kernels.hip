// PRP kernels with G = 2, 4 or 8 lanes per candidate (group.hpp).  Same
// arguments, modes, verdict words and residues as the one-lane kernels of
// prp_kernels.hip; work item i is tested by lanes G*i .. G*i+G-1.
#include "group.hpp"
#include "launch.hpp"

namespace pgx {
namespace {

__device__ __forceinline__ uint32_t items_of(const PrpArgs& a) {
  return a.count ? min(*a.count, a.n) : a.n;
}

template <int Q, int G>
__device__ __forceinline__ uint32_t group_test_base(uint32_t base, const uint32_t (&n)[Q],
                                              const uint32_t (&one)[Q], uint32_t ninv, int L,
                                              int t, double inv_n, const PrpArgs& a,
                                              const uint32_t* c, uint32_t (&res)[Q]) {
  uint32_t x[Q];  // Mont(base) = base * R mod N
#pragma unroll
  for (int s = 0; s < Q; ++s) x[s] = one[s];
  if (base == 2u) grp::dbl_lazy<Q, G>(x, 1u);
  else grp::mul_small<Q, G>(x, base, n, inv_n, L);
  grp::sub_if_ge<Q, G>(x, n);
  const size_t ld = a.ld;
  auto word = [&](int j) -> uint32_t { return (j < (int)a.buf_limbs) ? c[(size_t)j * ld] : 0u; };
  return grp::powm_verdict<Q, G>(x, one, n, ninv, L, t, base, inv_n, word, res);
}

template <int Q, int G>
__global__ __launch_bounds__(kBlock) void k_prp_group(PrpArgs a) {
  constexpr int Sp = Q * G;
  constexpr uint32_t kGroups = kBlock / G;
  __shared__ uint32_t stage[kGroups * Sp];
  const uint32_t slot = threadIdx.x / G;
  const uint32_t item = blockIdx.x * kGroups + slot;
  const uint32_t lane = threadIdx.x & (G - 1);
  if (item >= items_of(a)) return;  // whole groups leave together
  const uint32_t p = a.index ? a.index[item] : item;
  const uint32_t* c = a.cand + p;
  const size_t ld = a.ld;
  const int W = (int)a.buf_limbs;
  auto word = [&](int j) -> uint32_t { return (j < W) ? c[(size_t)j * ld] : 0u; };

  int L = 0;
  for (int j = W - 1; j >= 0; --j) {
    const uint32_t v = word(j);
    if (v) {
      L = 32 * j + 32 - __builtin_clz(v);
      break;
    }
  }
  const uint32_t w0 = word(0);
  uint32_t v = PGX_VERDICT_INVALID;
  uint32_t res[Q];
#pragma unroll
  for (int s = 0; s < Q; ++s) res[s] = 0;
  if ((w0 & 1u) && L >= 2 && L <= grp::RB * Sp - 4) {
    int t = 0;  // ctz(N - 1)
    if (w0 != 1u) {
      t = __builtin_ctz(w0 - 1u);
    } else {
      for (int j = 1; j < W; ++j)
        if (word(j)) {
          t = 32 * j + __builtin_ctz(word(j));
          break;
        }
    }
    uint32_t n[Q];
#pragma unroll
    for (int s = 0; s < Q; ++s) {
      const int g = (int)lane * Q + s, bit = grp::RB * g, wi = bit >> 5;
      const uint64_t pair = ((uint64_t)word(wi + 1) << 32) | word(wi);
      n[s] = (uint32_t)(pair >> (bit & 31)) & grp::MASK;
    }
    const uint32_t ninv = neg_inv32(w0);
    uint32_t one[Q];
    grp::r_mod_n<Q, G>(one, n, L);
    const double inv_n = 1.0 / grp::scaled<Q, G>(n, L);
    if (a.mode == PGX_MODE_FERMAT) {
      v = group_test_base<Q, G>(2u, n, one, ninv, L, t, inv_n, a, c, res);
    } else {
      const uint32_t need = (a.mode == PGX_MODE_STRONG) ? FLAG_SPRP : FLAG_QUIRK;
      uint32_t passed = 0;
      bool alive = true;  // every base so far passed
      for (uint32_t b = 0; b < a.bases.count; ++b) {
        uint32_t f;
        if (b == 0) {
          f = group_test_base<Q, G>(a.bases.a[0], n, one, ninv, L, t, inv_n, a, c, res);
        } else {
          uint32_t scratch[Q];
          f = group_test_base<Q, G>(a.bases.a[b], n, one, ninv, L, t, inv_n, a, c, scratch);
        }
        alive = alive && (f & need);
        passed += alive ? 1u : 0u;
        // wavefront-uniform exit, as in the one-lane kernels
        if (!wave_any(alive)) break;
      }
      v = passed;
    }
  }
  if (lane == 0) a.verdict[p] = v;
  if (a.residue) {
    // limbs -> 32-bit words through the group's slice of LDS
    uint32_t* st = stage + slot * Sp;
#pragma unroll
    for (int s = 0; s < Q; ++s) st[lane * Q + s] = res[s];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t* o = a.residue + p;
    for (int w = (int)lane; w < W; w += G) {
      const int bit = 32 * w, j0 = bit / grp::RB, off = bit % grp::RB;
      uint64_t acc = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int j = j0 + k, sh = grp::RB * k - off;
        if (j < Sp && sh < 64) acc |= (sh >= 0) ? ((uint64_t)st[j] << sh) : ((uint64_t)st[j] >> -sh);
      }
      o[(size_t)w * a.res_ld] = (uint32_t)acc;
    }
  }
}

template <int Q, int G>
hipError_t launch_group(const PrpArgs& a, hipStream_t s) {
  if (a.n == 0) return hipSuccess;
  constexpr uint32_t kGroups = kBlock / G;
  hipLaunchKernelGGL((k_prp_group<Q, G>), dim3((a.n + kGroups - 1) / kGroups), dim3(kBlock), 0, s,
                     a);
  return hipGetLastError();
}

struct RegisterGroups {
  RegisterGroups() {
    PrpLaunchFn* t2 = group_table(2);
    PrpLaunchFn* t4 = group_table(4);
    PrpLaunchFn* t8 = group_table(8);
    PrpLaunchFn* t16 = group_table(16);
    PrpLaunchFn* t32 = group_table(32);
    t2[16] = &launch_group<8, 2>;
    t2[24] = &launch_group<12, 2>;
    t2[32] = &launch_group<16, 2>;
    t2[40] = &launch_group<20, 2>;
    t2[48] = &launch_group<24, 2>;
    t4[16] = &launch_group<4, 4>;
    t4[24] = &launch_group<6, 4>;
    t4[32] = &launch_group<8, 4>;
    t4[40] = &launch_group<10, 4>;
    t4[48] = &launch_group<12, 4>;
    t4[56] = &launch_group<14, 4>;
    t4[64] = &launch_group<16, 4>;
    t4[76] = &launch_group<19, 4>;
    t8[32] = &launch_group<4, 8>;
    t8[48] = &launch_group<6, 8>;
    t8[64] = &launch_group<8, 8>;
    t8[80] = &launch_group<10, 8>;
    t16[32] = &launch_group<2, 16>;
    t16[48] = &launch_group<3, 16>;
    t16[64] = &launch_group<4, 16>;
    t16[80] = &launch_group<5, 16>;
    t32[64] = &launch_group<2, 32>;
    t32[96] = &launch_group<3, 32>;
  }
} register_groups;

}  // namespace
}  // namespace pgx
