// C ABI of the batched path: persistent launches, indexed (compacted) launches,
// compaction, the launch-shape autotuner and hipGraph plans over buckets of
// mixed candidate sizes.  See include/pgx.h for each entry point's contract.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "launch.hpp"
#include "runtime.hpp"

using pgx::fail;

namespace {

uint32_t tune_key(int S, uint32_t flags) { return ((uint32_t)S << 8) | (flags & PGX_MODE_MASK); }

pgx_tuning_t default_tuning() {
  pgx_tuning_t t{};
  t.blocks_per_cu = 8;  // 32 wavefronts per CU: more than any kernel keeps resident
  t.waves_per_block = 4;
  t.chunk = 64;
  t.lanes_per_number = 1;
  return t;
}

int check_tuning(pgx_runner* h, const pgx_tuning_t& t) {
  if (t.blocks_per_cu == 0 && t.lanes_per_number == 1) return PGX_OK;  // plain grid
  if (t.waves_per_block < 1 || t.waves_per_block > 4 || t.blocks_per_cu < 1 ||
      t.blocks_per_cu > 64 || t.chunk < 64 || t.chunk % 64 || t.chunk > 4096)
    return fail(h, PGX_E_ARG, "bad tuning: blocks_per_cu=%u waves_per_block=%u chunk=%u",
                t.blocks_per_cu, t.waves_per_block, t.chunk);
  if (t.lanes_per_number != 1 && !pgx::valid_group((int)t.lanes_per_number))
    return fail(h, PGX_E_ARG, "lanes_per_number=%u: 1, 2, 4, 8, 16 or 32", t.lanes_per_number);
  return PGX_OK;
}



// Validates a bucket; returns the kernel limb count (> 0) or a negative code.
int check_bucket(pgx_runner* h, const pgx_bucket_t* b, uint32_t flags) {
  if (!b || !b->d_cand || !b->d_verdict) return fail(h, PGX_E_ARG, "null bucket or pointer");
  if (b->ld < b->n || b->limbs == 0)
    return fail(h, PGX_E_ARG, "bad bucket shape: n=%u ld=%u limbs=%u", b->n, b->ld, b->limbs);
  if ((flags & PGX_WANT_RESIDUE) && (!b->d_residues || b->res_ld < b->n))
    return fail(h, PGX_E_ARG, "residues requested without a buffer of res_ld >= n");
  if ((flags & PGX_MODE_MASK) > PGX_MODE_FERMAT) return fail(h, PGX_E_ARG, "bad mode");
  if (b->max_bits < 2) return fail(h, PGX_E_ARG, "max_bits must be >= 2");
  const int S = pgx::kernel_for_bits(b->max_bits);
  if (S == 0 || !pgx::persist_table()[S])
    return fail(h, PGX_E_INPUT, "candidates of %u bits exceed the widest kernel (%d bits)",
                b->max_bits, PGX_MAX_BITS);
  return S;
}

pgx::PrpArgs args_of(pgx_runner* h, const pgx_bucket_t* b, uint32_t flags) {
  pgx::PrpArgs a{};
  a.cand = b->d_cand;
  a.ld = b->ld;
  a.n = b->n;
  a.buf_limbs = b->limbs;
  a.check_limbs = b->limbs;  // max_bits is the caller's bound
  a.mode = flags & PGX_MODE_MASK;
  a.bases = h->bases;
  a.verdict = b->d_verdict;
  a.residue = (flags & PGX_WANT_RESIDUE) ? b->d_residues : nullptr;
  a.res_ld = b->res_ld;
  return a;
}

pgx::PersistShape shape_of(pgx_runner* h, const pgx_tuning_t& t) {
  pgx::PersistShape sh{};
  sh.blocks = (uint32_t)h->n_cu * t.blocks_per_cu;
  sh.threads = 64u * t.waves_per_block;
  sh.chunk = t.chunk;
  return sh;
}

// Kernel of a bucket under a tuning: one lane per candidate (persistent, or a
// plain grid with one work item per thread when blocks_per_cu is 0), or a
// group of lanes per candidate (one work item per group, plain grid).  The
// counter serves the persistent form only.
hipError_t launch_tuned(pgx_runner* h, int S, const pgx_tuning_t& t, const pgx::PrpArgs& a,
                        uint32_t max_bits, uint32_t* counter, hipStream_t s) {
  if (t.lanes_per_number == 1 && t.blocks_per_cu == 0) return pgx::prp_table()[S](a, s);
  if (t.lanes_per_number == 1) {
    hipError_t e = pgx::launch_fill(counter, 1, 0u, s);
    return e != hipSuccess ? e : pgx::persist_table()[S](a, counter, shape_of(h, t), s);
  }
  const int Sp = pgx::group_width_for_bits(max_bits, (int)t.lanes_per_number);
  if (Sp == 0) return hipErrorInvalidValue;
  return pgx::group_table((int)t.lanes_per_number)[Sp](a, s);
}

pgx_tuning_t tuning_for(pgx_runner* h, int S, uint32_t flags, const pgx_tuning_t* given) {
  if (given) return *given;
  auto it = h->tuned.find(tune_key(S, flags));
  return it == h->tuned.end() ? default_tuning() : it->second;
}

// Next counter slot of the runner: launches in flight on different streams
// (up to kCounters - 1 of them, from any host threads) never share a counter.
uint32_t* next_counter(pgx_runner* h) { return h->d_counters + h->slots.next(); }

int launch_persistent(pgx_runner* h, const pgx_bucket_t* b, const pgx_tuning_t* tuning,
                      uint32_t flags, const uint32_t* d_index, const uint32_t* d_count,
                      hipStream_t s) {
  const int S = check_bucket(h, b, flags);
  if (S < 0) return S;
  const pgx_tuning_t t = tuning_for(h, S, flags, tuning);
  int rc = check_tuning(h, t);
  if (rc != PGX_OK) return rc;
  if (b->n == 0) return PGX_OK;
  if (t.lanes_per_number > 1 && !pgx::group_width_for_bits(b->max_bits, (int)t.lanes_per_number))
    return fail(h, PGX_E_INPUT, "no %u-lane kernel holds %u bits", t.lanes_per_number,
                b->max_bits);
  pgx::PrpArgs a = args_of(h, b, flags);
  a.index = d_index;
  a.count = d_count;
  PGX_HIP(h, hipSetDevice(h->device));
  PGX_HIP(h, launch_tuned(h, S, t, a, b->max_bits, next_counter(h), s));
  return PGX_OK;
}

}  // namespace

int pgx::launch_auto(pgx_runner* h, const PrpArgs& a, uint32_t max_bits, hipStream_t s) {
  const int S = pgx::kernel_for_bits(max_bits);
  if (S == 0 || !pgx::persist_table()[S])
    return fail(h, PGX_E_INPUT, "candidates of %u bits exceed the widest kernel (%d bits)",
                max_bits, PGX_MAX_BITS);
  const pgx_tuning_t t = tuning_for(h, S, a.mode, nullptr);
  if (t.lanes_per_number > 1 && !pgx::group_width_for_bits(max_bits, (int)t.lanes_per_number))
    return fail(h, PGX_E_INPUT, "no %u-lane kernel holds %u bits", t.lanes_per_number, max_bits);
  if (a.n == 0) return PGX_OK;
  PGX_HIP(h, launch_tuned(h, S, t, a, max_bits, next_counter(h), s));
  return PGX_OK;
}

struct pgx_plan {
  pgx_runner* h = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint32_t* d_counters = nullptr;
  std::vector<hipStream_t> streams;  // [0] captures, [1..] one per bucket
  std::vector<hipEvent_t> events;
};

static void plan_release(pgx_plan* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (hipEvent_t e : p->events) (void)hipEventDestroy(e);
  for (hipStream_t s : p->streams) (void)hipStreamDestroy(s);
  if (p->d_counters) (void)hipFree(p->d_counters);
}

extern "C" {

int pgx_test_persistent(pgx_runner_t* h, const pgx_bucket_t* b, const pgx_tuning_t* tuning,
                        uint32_t flags, void* stream) {
  if (!h) return PGX_E_ARG;
  return launch_persistent(h, b, tuning, flags, nullptr, nullptr,
                           stream ? (hipStream_t)stream : h->stream);
}

int pgx_test_indexed(pgx_runner_t* h, const pgx_bucket_t* b, const uint32_t* d_index,
                     const uint32_t* d_count, const pgx_tuning_t* tuning, uint32_t flags,
                     void* stream) {
  if (!h) return PGX_E_ARG;
  if (!d_index || !d_count) return fail(h, PGX_E_ARG, "null index list or count");
  return launch_persistent(h, b, tuning, flags, d_index, d_count,
                           stream ? (hipStream_t)stream : h->stream);
}

int pgx_compact(pgx_runner_t* h, const uint32_t* d_verdict, uint32_t n, uint32_t mask,
                uint32_t* d_index, uint32_t* d_count, void* stream) {
  if (!h) return PGX_E_ARG;
  if (!d_verdict || !d_index || !d_count) return fail(h, PGX_E_ARG, "null device pointer");
  PGX_HIP(h, hipSetDevice(h->device));
  PGX_HIP(h, pgx::launch_compact(d_verdict, n, mask, d_index, d_count,
                                 stream ? (hipStream_t)stream : h->stream));
  return PGX_OK;
}

// Repetitions per shape; a shape scores its slowest one.  Workgroups of one
// wavefront can land unevenly on a compute unit's four SIMDs, and the fastest
// of a few runs rewards a lucky placement: on the m-interval search, 8
// one-wavefront workgroups per CU timed level with 2 of 4 wavefronts on the
// tuning sample and ran 15% slower on the search's own batches.
constexpr int kTuneReps = 3;

int pgx_autotune(pgx_runner_t* h, const pgx_bucket_t* b, uint32_t flags, pgx_tuning_t* best,
                 double* per_sec) {
  if (!h) return PGX_E_ARG;
  const int S = check_bucket(h, b, flags);
  if (S < 0) return S;
  if (b->n == 0) return fail(h, PGX_E_ARG, "empty bucket");
  // time on at most 2^18 candidates of the bucket
  pgx_bucket_t sample = *b;
  sample.n = std::min<uint32_t>(b->n, 1u << 18);
  hipEvent_t e0, e1;
  PGX_HIP(h, hipSetDevice(h->device));
  PGX_HIP(h, hipEventCreate(&e0));
  PGX_HIP(h, hipEventCreate(&e1));
  pgx_tuning_t top = default_tuning();
  float top_ms = 0;
  int rc = PGX_OK;
  const uint32_t waves[] = {1, 2, 4};
  const uint32_t blocks[] = {1, 2, 3, 4, 6, 8, 12, 16};
  const uint32_t chunks[] = {64, 256};
  // one untimed launch first (code object load, clocks)
  rc = launch_persistent(h, &sample, &top, flags, nullptr, nullptr, h->stream);
  for (uint32_t w : waves)
    for (uint32_t bl : blocks)
      for (uint32_t ch : chunks) {
        if (rc != PGX_OK) break;
        if (w * bl > 32) continue;  // beyond the 32 wavefronts a CU can hold
        // fewer than two chunks per wavefront: alone on the GPU every
        // wavefront takes its one chunk and the time is level, but next to
        // other kernels (a plan's concurrent buckets) the wavefronts start at
        // different times and a chunk is a large part of the whole, so the
        // last one to start sets the time (2^19 987-bit candidates, 4 blocks of
        // one wavefront per CU: 256-candidate chunks took 48.8 or 58.9 ms in a
        // plan beside three other buckets, 64-candidate chunks 48.9-51.0)
        if ((uint64_t)h->n_cu * bl * w * 2 > ((uint64_t)sample.n + ch - 1) / ch) continue;
        pgx_tuning_t t{bl, w, ch, 1};
        float shape_ms = 0;
        for (int rep = 0; rep < kTuneReps && rc == PGX_OK; ++rep) {
          hipError_t e = hipEventRecord(e0, h->stream);
          if (e == hipSuccess) rc = launch_persistent(h, &sample, &t, flags, nullptr, nullptr, h->stream);
          if (e == hipSuccess && rc == PGX_OK) e = hipEventRecord(e1, h->stream);
          if (e == hipSuccess && rc == PGX_OK) e = hipEventSynchronize(e1);
          float ms = 0;
          if (e == hipSuccess && rc == PGX_OK) e = hipEventElapsedTime(&ms, e0, e1);
          if (e != hipSuccess) rc = fail(h, PGX_E_HIP, "autotune timing: %s", hipGetErrorString(e));
          if (rc == PGX_OK && ms > shape_ms) shape_ms = ms;  // the slowest rep
        }
        if (rc == PGX_OK && (top_ms == 0 || shape_ms < top_ms)) {
          top_ms = shape_ms;
          top = t;
        }
      }
  for (uint32_t lanes : {1u, 2u, 4u, 8u, 16u, 32u}) {  // the plain grid, then lane groups
    if (rc != PGX_OK) break;
    if (lanes > 1 && !pgx::group_width_for_bits(b->max_bits, (int)lanes)) continue;
    pgx_tuning_t t = default_tuning();
    t.lanes_per_number = lanes;
    if (lanes == 1) t.blocks_per_cu = 0;
    float shape_ms = 0;
    for (int rep = 0; rep < kTuneReps && rc == PGX_OK; ++rep) {
      hipError_t e = hipEventRecord(e0, h->stream);
      if (e == hipSuccess) rc = launch_persistent(h, &sample, &t, flags, nullptr, nullptr, h->stream);
      if (e == hipSuccess && rc == PGX_OK) e = hipEventRecord(e1, h->stream);
      if (e == hipSuccess && rc == PGX_OK) e = hipEventSynchronize(e1);
      float ms = 0;
      if (e == hipSuccess && rc == PGX_OK) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) rc = fail(h, PGX_E_HIP, "autotune timing: %s", hipGetErrorString(e));
      if (rc == PGX_OK && ms > shape_ms) shape_ms = ms;  // the slowest rep
    }
    if (rc == PGX_OK && (top_ms == 0 || shape_ms < top_ms)) {  // 0: no persistent shape fit
      top_ms = shape_ms;
      top = t;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PGX_OK) return rc;
  h->tuned[tune_key(S, flags)] = top;
  if (best) *best = top;
  if (per_sec) *per_sec = top_ms > 0 ? sample.n / (top_ms * 1e-3) : 0.0;
  return PGX_OK;
}

int pgx_plan_create(pgx_plan_t** out, pgx_runner_t* h, const pgx_bucket_t* buckets,
                    uint32_t nbuckets, const pgx_tuning_t* tuning, uint32_t flags) {
  if (!out || !h) return PGX_E_ARG;
  *out = nullptr;
  if (!buckets || nbuckets == 0 || nbuckets > 64)
    return fail(h, PGX_E_ARG, "nbuckets must be in [1, 64]");
  std::vector<int> S(nbuckets);
  std::vector<pgx_tuning_t> T(nbuckets);
  for (uint32_t k = 0; k < nbuckets; ++k) {
    S[k] = check_bucket(h, &buckets[k], flags);
    if (S[k] < 0) return S[k];
    T[k] = tuning_for(h, S[k], flags, tuning ? &tuning[k] : nullptr);
    int rc = check_tuning(h, T[k]);
    if (rc != PGX_OK) return rc;
    if (T[k].lanes_per_number > 1 &&
        !pgx::group_width_for_bits(buckets[k].max_bits, (int)T[k].lanes_per_number))
      return fail(h, PGX_E_INPUT, "no %u-lane kernel holds %u bits", T[k].lanes_per_number,
                  buckets[k].max_bits);
  }
  PGX_HIP(h, hipSetDevice(h->device));
  pgx_plan* p = new (std::nothrow) pgx_plan();
  if (!p) return fail(h, PGX_E_ARG, "out of host memory");
  p->h = h;
  hipError_t e = hipMalloc((void**)&p->d_counters, nbuckets * 4);
  for (uint32_t k = 0; k <= nbuckets && e == hipSuccess; ++k) {
    hipStream_t s = nullptr;
    e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) p->streams.push_back(s);
  }
  for (uint32_t k = 0; k <= nbuckets && e == hipSuccess; ++k) {
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) p->events.push_back(ev);
  }
  if (e != hipSuccess) {
    plan_release(p);
    delete p;
    return fail(h, PGX_E_HIP, "plan resources: %s", hipGetErrorString(e));
  }
  // capture: a bucket of at least one wavefront per SIMD fills the GPU by
  // itself and runs on the capture stream after the previous such bucket
  // (side by side, their persistent wavefronts start at staggered times and
  // a tuned shape loses its balance: 2^21 candidates in four buckets took
  // 48.8-58.9 ms as four branches against 48.7-51.3 ms one after another);
  // smaller buckets fork into branches of their own that fill the gaps, and
  // every branch joins back before the end
  hipStream_t cap = p->streams[0];
  const uint64_t fills = (uint64_t)h->n_cu * 4u * 64u;
  e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) e = hipEventRecord(p->events[0], cap);
  for (uint32_t k = 0; k < nbuckets && e == hipSuccess; ++k) {
    if (buckets[k].n == 0) continue;
    const bool own_branch = buckets[k].n < fills;
    hipStream_t s = own_branch ? p->streams[k + 1] : cap;
    if (own_branch) e = hipStreamWaitEvent(s, p->events[0], 0);
    if (e == hipSuccess) {
      pgx::PrpArgs a = args_of(h, &buckets[k], flags);
      e = launch_tuned(h, S[k], T[k], a, buckets[k].max_bits, p->d_counters + k, s);
    }
    if (own_branch && e == hipSuccess) e = hipEventRecord(p->events[k + 1], s);
  }
  for (uint32_t k = 0; k < nbuckets && e == hipSuccess; ++k)
    if (buckets[k].n > 0 && buckets[k].n < fills) e = hipStreamWaitEvent(cap, p->events[k + 1], 0);
  hipError_t e_end = hipStreamEndCapture(cap, &p->graph);
  if (e == hipSuccess) e = e_end;
  if (e == hipSuccess) e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    plan_release(p);
    delete p;
    return fail(h, PGX_E_HIP, "plan capture: %s", hipGetErrorString(e));
  }
  *out = p;
  return PGX_OK;
}

int pgx_plan_launch(pgx_plan_t* p, void* stream) {
  if (!p) return PGX_E_ARG;
  PGX_HIP(p->h, hipSetDevice(p->h->device));
  PGX_HIP(p->h, hipGraphLaunch(p->exec, stream ? (hipStream_t)stream : p->h->stream));
  return PGX_OK;
}

int pgx_plan_destroy(pgx_plan_t* p) {
  if (!p) return PGX_E_ARG;
  (void)hipSetDevice(p->h->device);
  (void)hipDeviceSynchronize();
  plan_release(p);
  delete p;
  return PGX_OK;
}

int pgx_stream_create_cus(pgx_runner_t* h, uint32_t cu_first, uint32_t cu_count, void** out) {
  if (!h || !out) return PGX_E_ARG;
  *out = nullptr;
  const uint32_t n_cu = (uint32_t)h->n_cu;
  if (cu_count == 0 || cu_first >= n_cu || cu_count > n_cu - cu_first)
    return fail(h, PGX_E_ARG, "compute units [%u, %u) outside the device's %u", cu_first,
                cu_first + cu_count, n_cu);
  PGX_HIP(h, hipSetDevice(h->device));
  std::vector<uint32_t> mask((n_cu + 31) / 32, 0u);
  for (uint32_t cu = cu_first; cu < cu_first + cu_count; ++cu) mask[cu / 32] |= 1u << (cu % 32);
  hipStream_t s = nullptr;
  PGX_HIP(h, hipExtStreamCreateWithCUMask(&s, (uint32_t)mask.size(), mask.data()));
  *out = s;
  return PGX_OK;
}

int pgx_stream_destroy(pgx_runner_t* h, void* stream) {
  if (!h || !stream) return PGX_E_ARG;
  PGX_HIP(h, hipSetDevice(h->device));
  PGX_HIP(h, hipStreamDestroy((hipStream_t)stream));
  return PGX_OK;
}

}  // extern "C"
