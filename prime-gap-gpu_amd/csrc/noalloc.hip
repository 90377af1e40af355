// pgx_test_device_noalloc: the several-base batch of pgx_submit (first base on
// every candidate, survivors compacted on the device, later bases side by side
// on them) on device-resident candidates, with no runner and no allocation.
// For callers that cannot allocate: code under HIP graph capture, serving loops
// and harnesses that forbid hipMalloc.  No counterpart in the reference.
//
// Its scratch (survivor list, per-survivor pass masks, survivor count) lives in
// a file-scope device array, one copy per GPU, held until the process ends.
// In host code the array's name is only a host-side stand-in; its device
// address comes from hipGetSymbolAddress, once per device (with that device
// current), and only that address is ever handed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/pgx.h"
#include "runtime.hpp"

namespace {

constexpr uint32_t kMaxN = PGX_NOALLOC_MAX_N;
// [0, kMaxN): survivor list; [kMaxN, 2 kMaxN): pass masks; 2 kMaxN: count
constexpr size_t kScratchWords = 2 * (size_t)kMaxN + 64;
static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

int noalloc_fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  pgx::create_error() = buf;
  return code;
}

// The device address of this device's scratch (the device must be current).
int scratch_of(int device, uint32_t** out) {
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return noalloc_fail(PGX_E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s",
                          device, hipGetErrorString(e));
    g_scratch[device] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch[device];
  return PGX_OK;
}

}  // namespace

extern "C" int pgx_test_device_noalloc(int device, const uint32_t* d_cand, uint32_t ld,
                                       uint32_t limbs, size_t n, uint32_t max_bits,
                                       uint32_t flags, const uint32_t* bases, uint32_t nbases,
                                       uint32_t* d_verdict, uint32_t* d_residues, uint32_t res_ld,
                                       void* stream) {
  if (n == 0) return PGX_OK;
  if (!d_cand || !d_verdict) return noalloc_fail(PGX_E_ARG, "null device pointer");
  if (n > kMaxN)
    return noalloc_fail(PGX_E_CAPACITY, "n = %zu exceeds the no-allocation maximum of %u "
                        "candidates per call", n, kMaxN);
  if (ld < n || limbs == 0)
    return noalloc_fail(PGX_E_ARG, "bad shape: n=%zu ld=%u limbs=%u", n, ld, limbs);
  if ((flags & PGX_WANT_RESIDUE) && (!d_residues || res_ld < n))
    return noalloc_fail(PGX_E_ARG, "residues requested without a buffer of res_ld >= n");
  const uint32_t mode = flags & PGX_MODE_MASK;
  if (mode > PGX_MODE_FERMAT) return noalloc_fail(PGX_E_ARG, "bad mode");
  if (max_bits < 2) return noalloc_fail(PGX_E_ARG, "max_bits must be >= 2");
  if (nbases == 0 || nbases > PGX_MAX_BASES)
    return noalloc_fail(PGX_E_ARG, "nbases must be in [1, %d]", PGX_MAX_BASES);
  pgx::Bases all{};
  all.count = nbases;
  if (bases) {
    for (uint32_t k = 0; k < nbases; ++k) {
      if (bases[k] < 2 || bases[k] > PGX_MAX_BASE)
        return noalloc_fail(PGX_E_ARG, "base %u must be in [2, %d]", bases[k], PGX_MAX_BASE);
      all.a[k] = bases[k];
    }
  } else {
    pgx::first_primes(nbases, all.a);
  }
  const int S = pgx::kernel_for_bits(max_bits);
  if (S == 0)
    return noalloc_fail(PGX_E_INPUT, "candidates of %u bits exceed the widest kernel (%d bits)",
                        max_bits, PGX_MAX_BITS);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return noalloc_fail(PGX_E_NODEV, "no HIP device available");
  if (device < 0 || device >= ndev || device >= kMaxDevices)
    return noalloc_fail(PGX_E_ARG, "device %d out of range (%d devices)", device, ndev);
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return noalloc_fail(PGX_E_HIP, "hipSetDevice: %s", hipGetErrorString(e));
  uint32_t* s = nullptr;
  int rc = scratch_of(device, &s);
  if (rc != PGX_OK) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);

  // first base (every base in one pass when there is one base, or in FERMAT
  // mode, as pgx_submit does)
  const bool one_pass = nbases == 1 || mode == PGX_MODE_FERMAT;
  pgx::PrpArgs a{};
  a.cand = d_cand;
  a.ld = ld;
  a.n = (uint32_t)n;
  a.buf_limbs = limbs;
  a.check_limbs = limbs;  // max_bits is the caller's bound
  a.mode = mode;
  a.bases = one_pass ? all : pgx::first_base(all);
  a.verdict = d_verdict;
  a.residue = (flags & PGX_WANT_RESIDUE) ? d_residues : nullptr;
  a.res_ld = res_ld;
  e = pgx::prp_table()[S](a, st);
  if (e == hipSuccess && !one_pass) {
    pgx::Bases later{};
    later.count = nbases - 1;
    for (uint32_t k = 1; k < nbases; ++k) later.a[k - 1] = all.a[k];
    e = pgx::launch_survivor_bases(d_cand, ld, limbs, limbs, (uint32_t)n, S, mode, later,
                                   d_verdict, s, s + 2 * (size_t)kMaxN, s + kMaxN, st);
  }
  if (e != hipSuccess) return noalloc_fail(PGX_E_HIP, "launch failed: %s", hipGetErrorString(e));
  return PGX_OK;
}
