// Internal runtime state shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>

#include "../../include/pgx.h"
#include "launch.hpp"

struct pgx_runner {
  int device = 0;
  int n_cu = 0;  // compute units of the device
  size_t max_n = 0;
  uint32_t limbs = 0;
  pgx::Bases bases{};
  hipStream_t stream = nullptr;
  // pinned staging (limb-interleaved, ld = max_n)
  uint32_t* h_cand = nullptr;
  uint32_t* h_verdict = nullptr;
  uint32_t* h_res = nullptr;
  // device buffers
  uint32_t* d_cand = nullptr;
  uint32_t* d_verdict = nullptr;
  uint32_t* d_res = nullptr;
  uint32_t* d_counters = nullptr;  // work counters of persistent launches
  // pending batch
  bool pending = false;
  size_t pending_n = 0;
  uint32_t pending_flags = 0;
  // autotuned persistent shapes, keyed by (kernel limb count, mode)
  std::map<uint32_t, pgx_tuning_t> tuned;
  std::string err;
};

namespace pgx {

constexpr uint32_t kCounters = 64;  // counter slots per runner

std::string& create_error();

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
