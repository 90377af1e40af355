// C-ABI runtime: runners, staging, launches.  See include/pgx.h for the
// contract of each entry point and the reference function it replaces.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "../../include/pgx.h"
#include "launch.hpp"
#include "runtime.hpp"

namespace pgx {

PrpLaunchFn* prp_table() {
  static PrpLaunchFn table[kMaxKernelLimbs + 1] = {};
  return table;
}

PersistLaunchFn* persist_table() {
  static PersistLaunchFn table[kMaxKernelLimbs + 1] = {};
  return table;
}

BasesLaunchFn* bases_table() {
  static BasesLaunchFn table[kMaxKernelLimbs + 1] = {};
  return table;
}

PrpLaunchFn* group_table(int G) {
  static PrpLaunchFn two[kMaxGroupLimbs + 1] = {};
  static PrpLaunchFn four[kMaxGroupLimbs + 1] = {};
  static PrpLaunchFn eight[kMaxGroupLimbs + 1] = {};
  static PrpLaunchFn sixteen[kMaxGroupLimbs + 1] = {};
  static PrpLaunchFn thirtytwo[kMaxGroupLimbs + 1] = {};
  switch (G) {
    case 2: return two;
    case 4: return four;
    case 8: return eight;
    case 16: return sixteen;
    default: return thirtytwo;
  }
}

// Per thread: runners are created from several host threads at once (one per
// GPU thread of the reference's dispatch).
std::string& create_error() {
  static thread_local std::string msg = "no error";
  return msg;
}

}  // namespace pgx

using pgx::fail;

// First `count` primes, the reference's generate_primes (miller_rabin.h:272-289).
void pgx::first_primes(uint32_t count, uint32_t* out) {
  uint32_t found = 0;
  for (uint32_t c = 2; found < count; ++c) {
    bool prime = true;
    for (uint32_t k = 0; k < found && out[k] * out[k] <= c; ++k)
      if (c % out[k] == 0) {
        prime = false;
        break;
      }
    if (prime) out[found++] = c;
  }
}

namespace {

uint32_t bit_length(const uint32_t* x, uint32_t limbs) {
  for (int j = (int)limbs - 1; j >= 0; --j)
    if (x[j]) return 32u * j + 32u - __builtin_clz(x[j]);
  return 0;
}

// Candidate-major rows (n x S words: the layout at the C ABI) to and from the
// column-major layout of the kernels (word j of candidate i at j*ld + i).
// Tiles of 64 candidates x 64 words pass through LDS so that both the reads
// and the writes are coalesced; the grid strides over candidate tiles and
// blockIdx.y picks the 64-word slice.
constexpr uint32_t kTile = 64;

template <bool kToCols>
__global__ __launch_bounds__(256) void k_transpose(const uint32_t* __restrict__ src,
                                                   uint32_t* __restrict__ dst, uint32_t n,
                                                   uint32_t S, uint32_t ld) {
  __shared__ uint32_t tile[kTile][kTile + 1];  // [candidate][word], odd pitch
  const uint32_t j0 = blockIdx.y * kTile;
  const uint32_t wj = min(kTile, S - j0);
  const uint32_t tiles = (n + kTile - 1) / kTile;
  for (uint32_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    const uint32_t i0 = t * kTile, wi = min(kTile, n - i0);
    for (uint32_t e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
      const uint32_t r = e / kTile, c = e % kTile;
      if (kToCols) {  // candidate r, word c
        if (r < wi && c < wj) tile[r][c] = src[(size_t)(i0 + r) * S + j0 + c];
      } else {  // word r, candidate c
        if (r < wj && c < wi) tile[c][r] = src[(size_t)(j0 + r) * ld + i0 + c];
      }
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
      const uint32_t r = e / kTile, c = e % kTile;
      if (kToCols) {  // word r, candidate c
        if (r < wj && c < wi) dst[(size_t)(j0 + r) * ld + i0 + c] = tile[c][r];
      } else {  // candidate r, word c
        if (r < wi && c < wj) dst[(size_t)(i0 + r) * S + j0 + c] = tile[r][c];
      }
    }
    __syncthreads();
  }
}

hipError_t launch_transpose(bool to_cols, const uint32_t* src, uint32_t* dst, size_t n,
                            uint32_t S, uint32_t ld, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t tiles = (n + kTile - 1) / kTile;
  const dim3 grid((unsigned)std::min<size_t>(tiles, 8192), (S + kTile - 1) / kTile);
  if (to_cols)
    hipLaunchKernelGGL(k_transpose<true>, grid, dim3(256), 0, s, src, dst, (uint32_t)n, S, ld);
  else
    hipLaunchKernelGGL(k_transpose<false>, grid, dim3(256), 0, s, src, dst, (uint32_t)n, S, ld);
  return hipGetLastError();
}

// Queue the D2H copies of a finished batch: verdicts, and residues turned
// candidate-major on the device (into d_cand, free once the test has run).
int queue_results(pgx_runner* h, pgx_batch& b, size_t n, uint32_t flags) {
  PGX_HIP(h, hipMemcpyAsync(b.h_verdict, b.d_verdict, n * 4, hipMemcpyDeviceToHost, b.stream));
  if (flags & PGX_WANT_RESIDUE) {
    PGX_HIP(h, launch_transpose(false, b.d_res, b.d_cand, n, h->limbs, (uint32_t)n, b.stream));
    PGX_HIP(h, hipMemcpyAsync(b.h_res, b.d_cand, (size_t)h->limbs * n * 4,
                              hipMemcpyDeviceToHost, b.stream));
  }
  return PGX_OK;
}

// Staging and device buffers of a batch slot (its stream and count are set by
// the caller).
hipError_t alloc_batch(pgx_runner* h, pgx_batch& b) {
  const size_t words = (size_t)h->limbs * h->max_n;
  hipError_t e = hipHostMalloc((void**)&b.h_cand, words * 4, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b.h_verdict, h->max_n * 4, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void**)&b.h_res, words * 4, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc((void**)&b.d_cand, words * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&b.d_verdict, h->max_n * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&b.d_res, words * 4);
  if (e == hipSuccess && h->bases.count > 1)
    e = hipMalloc((void**)&b.d_survivors, 2 * h->max_n * 4);
  return e;
}

// Frees a batch slot's buffers (not its stream or count).
void release_batch(pgx_batch& b) {
  if (b.h_cand) (void)hipHostFree(b.h_cand);
  if (b.h_verdict) (void)hipHostFree(b.h_verdict);
  if (b.h_res) (void)hipHostFree(b.h_res);
  if (b.d_cand) (void)hipFree(b.d_cand);
  if (b.d_verdict) (void)hipFree(b.d_verdict);
  if (b.d_res) (void)hipFree(b.d_res);
  if (b.d_survivors) (void)hipFree(b.d_survivors);
  if (b.d_K) (void)hipFree(b.d_K);
  if (b.d_m) (void)hipFree(b.d_m);
  if (b.d_status) (void)hipFree(b.d_status);
  if (b.h_m) (void)hipHostFree(b.h_m);
  if (b.h_status) (void)hipHostFree(b.h_status);
}

void release(pgx_runner* h) {
  if (h->b1) {
    if (h->b1->stream) (void)hipStreamDestroy(h->b1->stream);
    if (h->b1->d_count) (void)hipFree(h->b1->d_count);
    release_batch(*h->b1);
  }
  release_batch(h->b0);  // b0's stream is h->stream, its count a counter slot
  if (h->stream) (void)hipStreamDestroy(h->stream);
  if (h->d_counters) (void)hipFree(h->d_counters);
}

// A stream for the runner's batches: non-blocking, on the CUs of its mask
// (every CU when the mask is empty).
hipError_t batch_stream(const pgx_runner* h, hipStream_t* s) {
  if (h->cu_mask.empty()) return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
  return hipExtStreamCreateWithCUMask(s, (uint32_t)h->cu_mask.size(), h->cu_mask.data());
}

// A free batch slot, in submit order: slot 0, or the second slot (made here
// the first time slot 0 is busy); NULL when both hold a batch.
pgx_batch* acquire_batch(pgx_runner* h) {
  if (!h->b0.pending) return &h->b0;
  if (!h->b1) {
    auto b = std::make_unique<pgx_batch>();
    hipError_t e = batch_stream(h, &b->stream);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_count, 4);
    if (e == hipSuccess) e = alloc_batch(h, *b);
    if (e != hipSuccess) {
      if (b->stream) (void)hipStreamDestroy(b->stream);
      if (b->d_count) (void)hipFree(b->d_count);
      release_batch(*b);
      fail(h, PGX_E_HIP, "second batch slot: %s", hipGetErrorString(e));
      return nullptr;
    }
    h->b1 = std::move(b);
  }
  if (!h->b1->pending) return h->b1.get();
  fail(h, PGX_E_STATE, "two batches are already in flight");
  return nullptr;
}

}  // namespace

namespace {

// Plain launch with an explicit base list, over columns 0..n-1 or, with index
// and count, over the work list index[0..min(*count, n)).  `checked`: max_bits
// was derived from the candidates by the library, so the kernel skips the
// check of the rows above its width.
int launch_plain_bases(pgx_runner* h, const uint32_t* d_cand, uint32_t ld, uint32_t limbs,
                       size_t n, uint32_t max_bits, uint32_t flags, bool checked,
                       const pgx::Bases& bases,
                       const uint32_t* index, const uint32_t* count, uint32_t* d_verdict,
                       uint32_t* d_res, uint32_t res_ld, hipStream_t stream) {
  const int S = pgx::kernel_for_bits(max_bits);
  if (S == 0)
    return fail(h, PGX_E_INPUT, "candidates of %u bits exceed the widest kernel (%d bits)",
                max_bits, PGX_MAX_BITS);
  pgx::PrpLaunchFn fn = pgx::prp_table()[S];
  pgx::PrpArgs a{};
  a.cand = d_cand;
  a.ld = ld;
  a.n = (uint32_t)n;
  a.buf_limbs = limbs;
  a.check_limbs = checked ? 0u : limbs;
  a.mode = flags & PGX_MODE_MASK;
  a.bases = bases;
  a.index = index;
  a.count = count;
  a.verdict = d_verdict;
  a.residue = (flags & PGX_WANT_RESIDUE) ? d_res : nullptr;
  a.res_ld = res_ld;
  PGX_HIP(h, fn(a, stream));
  return PGX_OK;
}

// A batch in b.d_cand (pgx_submit, pgx_submit_mk): one pass when the runner
// has one base (or in FERMAT mode), else the first base and then
// launch_later_bases.
int launch_batch(pgx_runner* h, pgx_batch& b, size_t n, uint32_t max_bits, uint32_t flags) {
  const uint32_t mode = flags & PGX_MODE_MASK;
  if (h->bases.count == 1 || mode == PGX_MODE_FERMAT)
    return launch_plain_bases(h, b.d_cand, (uint32_t)n, h->limbs, n, max_bits, flags, true,
                              h->bases, nullptr, nullptr, b.d_verdict, b.d_res, (uint32_t)n,
                              b.stream);
  int rc = launch_plain_bases(h, b.d_cand, (uint32_t)n, h->limbs, n, max_bits, flags, true,
                              pgx::first_base(h->bases), nullptr, nullptr, b.d_verdict, b.d_res,
                              (uint32_t)n, b.stream);
  if (rc != PGX_OK) return rc;
  return pgx::launch_later_bases(h, b, (uint32_t)n, max_bits, mode, b.stream);
}

}  // namespace

pgx::Bases pgx::first_base(const Bases& b) {
  Bases f{};
  f.count = 1;
  f.a[0] = b.a[0];
  return f;
}

// Columns 0..n-1 of d_cand (ld = n) hold a batch whose first-base verdicts
// (0 or 1, or INVALID) are in d_verdict.  The columns that passed are listed
// on the device (k_compact); every other base of the runner then runs on every
// survivor at once, one (survivor, base) pair per lane, into a pass mask per
// survivor (k_prp_bases); k_settle writes each survivor's count of leading
// bases passed.  In one pass over all bases every wavefront holding a probable
// prime (sieved candidates put a few in nearly every wavefront) runs the bases
// one after another; here the later bases cost one exponentiation's latency
// together.
hipError_t pgx::launch_survivor_bases(const uint32_t* cand, uint32_t ld, uint32_t limbs,
                                      uint32_t check_limbs, uint32_t n, int S, uint32_t mode,
                                      const Bases& later, uint32_t* verdict, uint32_t* index,
                                      uint32_t* count, uint32_t* mask, hipStream_t stream) {
  if (n == 0 || later.count == 0) return hipSuccess;
  hipError_t e = pgx::launch_compact(verdict, n, 1u, index, count, stream);
  if (e == hipSuccess) e = pgx::launch_fill(mask, n, 0u, stream);
  if (e != hipSuccess) return e;
  pgx::PrpArgs a{};
  a.cand = cand;
  a.ld = ld;
  a.n = n;
  a.buf_limbs = limbs;
  a.check_limbs = check_limbs;
  a.mode = mode;
  a.bases = later;
  a.index = index;
  a.count = count;
  e = pgx::bases_table()[S](a, mask, stream);
  if (e != hipSuccess) return e;
  return pgx::launch_settle(verdict, index, count, n, mask, stream);
}

int pgx::launch_later_bases(pgx_runner* h, pgx_batch& b, uint32_t n, uint32_t max_bits,
                            uint32_t mode, hipStream_t stream) {
  if (h->bases.count < 2 || n == 0) return PGX_OK;
  const int S = pgx::kernel_for_bits(max_bits);
  if (S == 0 || !b.d_survivors)
    return fail(h, PGX_E_ARG, "no survivor pass for %u-bit candidates", max_bits);
  pgx::Bases later{};
  later.count = h->bases.count - 1;
  for (uint32_t k = 1; k < h->bases.count; ++k) later.a[k - 1] = h->bases.a[k];
  // the survivor list and the masks in d_survivors, the count in d_count
  PGX_HIP(h, pgx::launch_survivor_bases(b.d_cand, n, h->limbs, 0u, n, S, mode, later,
                                        b.d_verdict, b.d_survivors, b.d_count,
                                        b.d_survivors + h->max_n, stream));
  return PGX_OK;
}

int pgx::launch_prp(pgx_runner* h, const uint32_t* d_cand, uint32_t ld, uint32_t limbs, size_t n,
                    uint32_t max_bits, uint32_t flags, uint32_t* d_verdict, uint32_t* d_res,
                    uint32_t res_ld, hipStream_t stream) {
  return launch_plain_bases(h, d_cand, ld, limbs, n, max_bits, flags, false, h->bases, nullptr,
                            nullptr, d_verdict, d_res, res_ld, stream);
}

using pgx::launch_prp;

extern "C" {

int pgx_runner_create(pgx_runner_t** out, int device, size_t max_n, uint32_t limbs,
                      const uint32_t* bases, uint32_t nbases) {
  if (!out) return fail(nullptr, PGX_E_ARG, "out is NULL");
  *out = nullptr;
  if (max_n == 0 || max_n > 0xFFFFFFFFu)
    return fail(nullptr, PGX_E_ARG, "max_n must be in [1, 2^32)");
  if (limbs == 0 || limbs > 4096) return fail(nullptr, PGX_E_ARG, "limbs must be in [1, 4096]");
  if (nbases == 0 || nbases > PGX_MAX_BASES)
    return fail(nullptr, PGX_E_ARG, "nbases must be in [1, %d]", PGX_MAX_BASES);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, PGX_E_NODEV, "no HIP device available");
  if (device < 0 || device >= ndev)
    return fail(nullptr, PGX_E_ARG, "device %d out of range (%d devices)", device, ndev);

  pgx_runner* h = new (std::nothrow) pgx_runner();
  if (!h) return fail(nullptr, PGX_E_ARG, "out of host memory");
  h->device = device;
  h->max_n = max_n;
  h->limbs = limbs;
  h->bases.count = nbases;
  if (bases) {
    for (uint32_t k = 0; k < nbases; ++k) {
      if (bases[k] < 2 || bases[k] > PGX_MAX_BASE) {
        delete h;
        return fail(nullptr, PGX_E_ARG, "base %u must be in [2, %d]", bases[k], PGX_MAX_BASE);
      }
      h->bases.a[k] = bases[k];
    }
  } else {
    pgx::first_primes(nbases, h->bases.a);
  }
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = alloc_batch(h, h->b0);
  if (e == hipSuccess) e = hipMalloc((void**)&h->d_counters, pgx::kCounters * 4);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&h->n_cu, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  h->b0.stream = h->stream;
  h->b0.d_count = h->d_counters;  // slot 0: persistent launches take slots 1..
  if (e != hipSuccess) {
    release(h);
    delete h;
    return fail(nullptr, PGX_E_HIP, "runner allocation failed: %s", hipGetErrorString(e));
  }
  h->err = "no error";
  *out = h;
  return PGX_OK;
}

int pgx_runner_confine(pgx_runner_t* h, uint32_t part, uint32_t parts) {
  if (!h) return PGX_E_ARG;
  const uint32_t n_cu = (uint32_t)h->n_cu;
  if (parts == 0 || part >= parts || parts > n_cu)
    return fail(h, PGX_E_ARG, "part %u of %u: need part < parts <= %u compute units", part, parts,
                n_cu);
  if (!h->inflight.empty()) return fail(h, PGX_E_STATE, "a batch is in flight");
  PGX_HIP(h, hipSetDevice(h->device));
  std::vector<uint32_t> mask;
  if (parts > 1) {
    mask.assign((n_cu + 31) / 32, 0u);
    for (uint32_t cu = part * n_cu / parts; cu < (part + 1) * n_cu / parts; ++cu)
      mask[cu / 32] |= 1u << (cu % 32);
  }
  // new streams first (nothing changes if that fails), then the old ones go
  std::swap(h->cu_mask, mask);
  hipStream_t s0 = nullptr, s1 = nullptr;
  hipError_t e = batch_stream(h, &s0);
  if (e == hipSuccess && h->b1) e = batch_stream(h, &s1);
  if (e != hipSuccess) {
    if (s0) (void)hipStreamDestroy(s0);
    std::swap(h->cu_mask, mask);
    return fail(h, PGX_E_HIP, "stream on compute units: %s", hipGetErrorString(e));
  }
  // work the caller queued on the runner's stream (pgx_test_device with no
  // stream of its own) finishes on the old stream
  PGX_HIP(h, hipStreamSynchronize(h->stream));
  (void)hipStreamDestroy(h->stream);
  h->stream = h->b0.stream = s0;
  if (h->b1) {
    PGX_HIP(h, hipStreamSynchronize(h->b1->stream));
    (void)hipStreamDestroy(h->b1->stream);
    h->b1->stream = s1;
  }
  return PGX_OK;
}

int pgx_runner_destroy(pgx_runner_t* h) {
  if (!h) return PGX_E_ARG;
  (void)hipSetDevice(h->device);
  for (pgx_batch* b : h->inflight) (void)hipStreamSynchronize(b->stream);
  release(h);
  delete h;
  return PGX_OK;
}

// Marks b as the newest batch in flight.
void mark_pending(pgx_runner* h, pgx_batch& b, size_t n, uint32_t flags) {
  b.pending = true;
  b.pending_n = n;
  b.pending_flags = flags;
  h->inflight.push_back(&b);
}

int pgx_submit(pgx_runner_t* h, const uint32_t* limbs_le, size_t n, uint32_t flags) {
  if (!h) return PGX_E_ARG;
  PGX_HIP(h, hipSetDevice(h->device));
  pgx_batch* bp = acquire_batch(h);
  if (!bp) return h->err.find("second batch") == 0 ? PGX_E_HIP : PGX_E_STATE;
  pgx_batch& b = *bp;
  if (n == 0) {  // the reference returns early on an empty batch (miller_rabin.h:365-366)
    mark_pending(h, b, 0, flags);
    return PGX_OK;
  }
  if (!limbs_le) return fail(h, PGX_E_ARG, "limbs_le is NULL");
  if (n > h->max_n)
    return fail(h, PGX_E_CAPACITY, "tests(%zu) > n(%zu)", n, h->max_n);
  if ((flags & PGX_MODE_MASK) > PGX_MODE_FERMAT) return fail(h, PGX_E_ARG, "bad mode");
  const uint32_t S = h->limbs;
  // Candidate-major through pinned staging, about 1 MiB at a time: each
  // piece is checked (odd, 2 <= bits <= the widest kernel) and copied while
  // it is in the host's cache, and its DMA runs while the next piece is
  // checked.  d_res holds the rows until the transpose has read them (the
  // test writes its residues there afterwards, in stream order).
  const size_t piece = std::max<size_t>(1, (size_t)(1u << 20) / (4u * S));
  uint32_t max_bits = 0;
  for (size_t c0 = 0; c0 < n; c0 += piece) {
    const size_t cn = std::min(piece, n - c0);
    const uint32_t* src = limbs_le + c0 * S;
    for (size_t i = 0; i < cn; ++i) {
      const uint32_t* x = src + i * S;
      const uint32_t L = bit_length(x, S);
      if (!(x[0] & 1u) || L < 2 || L > PGX_MAX_BITS) {
        // earlier pieces may still be in flight from the staging buffer
        (void)hipStreamSynchronize(b.stream);
        if (!(x[0] & 1u) || L < 2)
          return fail(h, PGX_E_INPUT, "candidate %zu is even or <= 1", c0 + i);
        return fail(h, PGX_E_INPUT, "candidate %zu has %u bits, more than the kernels support",
                    c0 + i, L);
      }
      max_bits = std::max(max_bits, L);
    }
    std::memcpy(b.h_cand + c0 * S, src, cn * S * 4);
    PGX_HIP(h, hipMemcpyAsync(b.d_res + c0 * S, b.h_cand + c0 * S, cn * S * 4,
                              hipMemcpyHostToDevice, b.stream));
  }
  PGX_HIP(h, launch_transpose(true, b.d_res, b.d_cand, n, S, (uint32_t)n, b.stream));
  int rc = launch_batch(h, b, n, max_bits, flags);
  if (rc != PGX_OK) return rc;
  rc = queue_results(h, b, n, flags);
  if (rc != PGX_OK) return rc;
  mark_pending(h, b, n, flags);
  return PGX_OK;
}

int pgx_submit_mk(pgx_runner_t* h, const uint32_t* K, uint32_t k_words, const uint64_t* m,
                  size_t n, uint32_t x, int32_t sign, uint32_t flags) {
  if (!h) return PGX_E_ARG;
  if (!K || k_words == 0 || (n && !m) || (sign != 1 && sign != -1))
    return fail(h, PGX_E_ARG, "bad pgx_submit_mk arguments");
  if (n > h->max_n) return fail(h, PGX_E_CAPACITY, "tests(%zu) > n(%zu)", n, h->max_n);
  if ((flags & PGX_MODE_MASK) > PGX_MODE_FERMAT) return fail(h, PGX_E_ARG, "bad mode");
  PGX_HIP(h, hipSetDevice(h->device));
  pgx_batch* bp = acquire_batch(h);
  if (!bp) return h->err.find("second batch") == 0 ? PGX_E_HIP : PGX_E_STATE;
  pgx_batch& b = *bp;
  if (n == 0) {
    mark_pending(h, b, 0, flags);
    return PGX_OK;
  }
  // Width bound: bits(m*K + x) <= bits(K) + bits(max m) + 1.
  const uint32_t kb = bit_length(K, k_words);
  uint64_t mmax = 0;
  for (size_t i = 0; i < n; ++i) mmax = std::max(mmax, m[i]);
  const uint32_t mb = mmax ? 64u - __builtin_clzll(mmax) : 0u;
  const uint32_t bound = std::max<uint32_t>(kb + mb + 1, 2u);
  const uint32_t max_bits = std::min<uint32_t>(bound, 32u * h->limbs);
  if (!b.d_m) {
    PGX_HIP(h, hipMalloc((void**)&b.d_m, h->max_n * 8));
    PGX_HIP(h, hipHostMalloc((void**)&b.h_m, h->max_n * 8, hipHostMallocDefault));
    PGX_HIP(h, hipMalloc((void**)&b.d_status, 4));
    PGX_HIP(h, hipHostMalloc((void**)&b.h_status, 4, hipHostMallocDefault));
  }
  if (b.h_K.size() != k_words || std::memcmp(b.h_K.data(), K, (size_t)k_words * 4) != 0) {
    // K changed: upload it once (this slot's stream is idle: its last batch
    // was waited for)
    if (b.d_K) (void)hipFree(b.d_K);
    b.d_K = nullptr;
    b.h_K.clear();
    PGX_HIP(h, hipMalloc((void**)&b.d_K, (size_t)k_words * 4));
    PGX_HIP(h, hipMemcpy(b.d_K, K, (size_t)k_words * 4, hipMemcpyHostToDevice));
    b.h_K.assign(K, K + k_words);
  }
  std::memcpy(b.h_m, m, n * 8);
  PGX_HIP(h, hipMemcpyAsync(b.d_m, b.h_m, n * 8, hipMemcpyHostToDevice, b.stream));
  PGX_HIP(h, pgx::launch_fill(b.d_status, 1, 0u, b.stream));
  int rc = pgx_build_candidates(h, b.d_K, k_words, b.d_m, nullptr, x, sign, (uint32_t)n,
                                b.d_cand, (uint32_t)n, h->limbs, b.d_status, b.stream);
  if (rc != PGX_OK) return rc;
  rc = launch_batch(h, b, n, max_bits, flags);
  if (rc != PGX_OK) return rc;
  PGX_HIP(h, hipMemcpyAsync(b.h_status, b.d_status, 4, hipMemcpyDeviceToHost, b.stream));
  rc = queue_results(h, b, n, flags);
  if (rc != PGX_OK) return rc;
  mark_pending(h, b, n, flags);
  b.status_pending = true;
  return PGX_OK;
}

int pgx_wait(pgx_runner_t* h, uint32_t* verdict, uint32_t* residues) {
  if (!h) return PGX_E_ARG;
  if (h->inflight.empty()) return fail(h, PGX_E_STATE, "no batch in flight");
  pgx_batch& b = *h->inflight.front();  // the oldest batch
  h->inflight.pop_front();
  b.pending = false;
  const size_t n = b.pending_n;
  if (n == 0) return PGX_OK;
  PGX_HIP(h, hipStreamSynchronize(b.stream));
  if (b.status_pending) {
    b.status_pending = false;
    if (*b.h_status) return fail(h, PGX_E_INPUT, "a candidate m*K +- x did not fit or was negative");
  }
  if (verdict) std::memcpy(verdict, b.h_verdict, n * 4);
  if ((b.pending_flags & PGX_WANT_RESIDUE) && residues)
    std::memcpy(residues, b.h_res, (size_t)h->limbs * n * 4);
  return PGX_OK;
}

int pgx_run(pgx_runner_t* h, const uint32_t* limbs_le, size_t n, uint32_t flags,
            uint32_t* verdict, uint32_t* residues) {
  int rc = pgx_submit(h, limbs_le, n, flags);
  if (rc != PGX_OK) return rc;
  return pgx_wait(h, verdict, residues);
}

int pgx_test_device(pgx_runner_t* h, const uint32_t* d_cand, uint32_t ld, uint32_t limbs,
                    size_t n, uint32_t max_bits, uint32_t flags, uint32_t* d_verdict,
                    uint32_t* d_residues, uint32_t res_ld, void* stream) {
  if (!h) return PGX_E_ARG;
  if (n == 0) return PGX_OK;
  if (!d_cand || !d_verdict) return fail(h, PGX_E_ARG, "null device pointer");
  if (n > 0xFFFFFFFFu || ld < n || limbs == 0)
    return fail(h, PGX_E_ARG, "bad shape: n=%zu ld=%u limbs=%u", n, ld, limbs);
  if ((flags & PGX_WANT_RESIDUE) && (!d_residues || res_ld < n))
    return fail(h, PGX_E_ARG, "residues requested without a buffer of res_ld >= n");
  if ((flags & PGX_MODE_MASK) > PGX_MODE_FERMAT) return fail(h, PGX_E_ARG, "bad mode");
  if (max_bits < 2) return fail(h, PGX_E_ARG, "max_bits must be >= 2");
  PGX_HIP(h, hipSetDevice(h->device));
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  return launch_prp(h, d_cand, ld, limbs, n, max_bits, flags, d_verdict, d_residues, res_ld, s);
}

const char* pgx_last_error(const pgx_runner_t* h) {
  return h ? h->err.c_str() : pgx::create_error().c_str();
}

}  // extern "C"
