// Internal runtime state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pgx.h"
#include "launch.hpp"
#include "slots.hpp"

namespace pgx {
constexpr uint32_t kCounters = 64;  // counter slots per runner
}  // namespace pgx

// One batch of pgx_submit / pgx_submit_mk: its staging, device buffers,
// stream and state.  A runner holds up to two, so that the later bases of one
// batch overlap the first base of the next (the reference keeps three batches
// in flight with one runner each, gap_search_gpu.cpp:887-904).
struct pgx_batch {
  hipStream_t stream = nullptr;
  // pinned staging, candidate-major as at the C ABI (the device transposes)
  uint32_t* h_cand = nullptr;
  uint32_t* h_verdict = nullptr;
  uint32_t* h_res = nullptr;
  // device buffers
  uint32_t* d_cand = nullptr;
  uint32_t* d_verdict = nullptr;
  uint32_t* d_res = nullptr;
  // survivors of the first base (runners with several bases): column list of
  // max_n words (its length in *d_count), then their per-base pass masks
  uint32_t* d_survivors = nullptr;
  uint32_t* d_count = nullptr;
  // device construction (pgx_submit_mk): K cache, multipliers, status word
  std::vector<uint32_t> h_K;  // the K held in d_K
  uint32_t* d_K = nullptr;
  uint64_t* h_m = nullptr;  // pinned
  uint64_t* d_m = nullptr;
  uint32_t* d_status = nullptr;
  uint32_t* h_status = nullptr;  // pinned
  bool status_pending = false;   // the pending batch was built on the device
  bool pending = false;
  size_t pending_n = 0;
  uint32_t pending_flags = 0;
};

struct pgx_runner {
  int device = 0;
  int n_cu = 0;  // compute units of the device
  size_t max_n = 0;
  uint32_t limbs = 0;
  pgx::Bases bases{};
  hipStream_t stream = nullptr;    // the runner's stream (also batch slot 0's)
  // CU mask of the runner's streams (pgx_runner_confine); empty: every CU
  std::vector<uint32_t> cu_mask;
  uint32_t* d_counters = nullptr;  // work counters of persistent launches
  pgx::CounterSlots slots{pgx::kCounters};  // the slot each persistent launch takes
  pgx_batch b0;                    // slot 0: the m-interval search borrows its buffers
  std::unique_ptr<pgx_batch> b1;   // the second batch slot, made when first needed
  std::deque<pgx_batch*> inflight; // submitted batches not yet waited for, oldest first
  // autotuned persistent shapes, keyed by (kernel limb count, mode)
  std::map<uint32_t, pgx_tuning_t> tuned;
  std::string err;
};

namespace pgx {

// Message of the last failing runner create on the calling thread.
std::string& create_error();

// First `count` primes (the reference's generate_primes, miller_rabin.h:272-289).
void first_primes(uint32_t count, uint32_t* out);

inline int fail(pgx_runner* h, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h)
    h->err = buf;
  else
    create_error() = buf;
  return code;
}

// Fills n words at p with v (a kernel, so it may sit inside a captured graph).
hipError_t launch_fill(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
// Appends columns with verdict != INVALID && (verdict & mask) to index.
hipError_t launch_compact(const uint32_t* verdict, uint32_t n, uint32_t mask, uint32_t* index,
                          uint32_t* count, hipStream_t s);

// Plain PRP launch (one lane per column) with the runner's bases; PGX_E_*.
int launch_prp(pgx_runner* h, const uint32_t* d_cand, uint32_t ld, uint32_t limbs, size_t n,
               uint32_t max_bits, uint32_t flags, uint32_t* d_verdict, uint32_t* d_res,
               uint32_t res_ld, hipStream_t stream);

// The first of the runner's bases alone.
Bases first_base(const Bases& b);

// The later bases of a several-base batch over the survivors of its first base
// (verdicts 1 in `verdict`, columns 0..n-1 of cand): k_compact lists them in
// index (its length in *count), k_prp_bases tests every other base on every
// survivor at once into mask (n words, cleared here), k_settle writes each
// survivor's count of leading bases passed.  index, count and mask are the
// caller's scratch; every kernel is a plain launch (no memset), so the
// sequence may be captured into a graph.
hipError_t launch_survivor_bases(const uint32_t* cand, uint32_t ld, uint32_t limbs,
                                 uint32_t check_limbs, uint32_t n, int S, uint32_t mode,
                                 const Bases& later, uint32_t* verdict, uint32_t* index,
                                 uint32_t* count, uint32_t* mask, hipStream_t stream);

// After a first-base launch over columns 0..n-1 of batch b's d_cand (ld = n)
// into d_verdict: the runner's other bases on the columns that passed, one
// (column, base) pair per lane, and the verdicts of those columns set to the
// count of leading bases passed.  Uses b's survivor buffers and count.  PGX_E_*.
int launch_later_bases(pgx_runner* h, pgx_batch& b, uint32_t n, uint32_t max_bits, uint32_t mode,
                       hipStream_t stream);

// Launch over device buffers with the runner's tuned (or default) shape for
// this width and mode: persistent one-lane workgroups or lane groups.
int launch_auto(pgx_runner* h, const PrpArgs& a, uint32_t max_bits, hipStream_t stream);

// Column i of out (out_words rows, leading dimension ld) = m_i * K + sign * x_i
// with m_i = m64[i] (m64 != NULL) or m_base + m_off[i], and x_i = x[i] (x != NULL)
// or x0.  *status |= 1 when a column does not fit out_words or is negative.
hipError_t launch_build(const uint32_t* K, uint32_t k_words, const uint64_t* m64,
                        const uint32_t* m_off, uint64_t m_base, const uint32_t* x, uint32_t x0,
                        int32_t sign, uint32_t n, uint32_t* out, uint32_t ld, uint32_t out_words,
                        uint32_t* status, hipStream_t s);

// verdict[index[i]] = 1 + (trailing one bits of mask[i]) for i < min(*count, n).
hipError_t launch_settle(uint32_t* verdict, const uint32_t* index, const uint32_t* count,
                         uint32_t n, const uint32_t* mask, hipStream_t s);

// dst[i] = src[index[i]] for i < n.
hipError_t launch_gather(const uint32_t* src, const uint32_t* index, uint32_t n, uint32_t* dst,
                         hipStream_t s);

}  // namespace pgx

#define PGX_HIP(h, call)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return pgx::fail((h), PGX_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                                \
  } while (0)
