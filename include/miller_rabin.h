// miller_rabin.h — drop-in C++ surface of the reference's GPU PRP runner,
// backed by the gfx950 C ABI (pgx.h / libpgx.so) instead of CUDA + CGBN.
//
// Kept from the reference (sethtroisi/prime-gap-gpu miller_rabin.h):
//   to_mpz / from_mpz                     miller_rabin.h:39-54
//   mr_params_t<tpi, bits, window_bits>   miller_rabin.h:117-130
//   test_runner_t<params>(n, rounds)      miller_rabin.h:295-349
//   ~test_runner_t()                      miller_rabin.h:351-362
//   run_test(active, tests, results)      miller_rabin.h:364-408
// so gpu_testing's `typedef mr_params_t<TPI, BITS, W> params;
// test_runner_t<params> runner(GPU_BATCH_SIZE, ROUNDS);
// runner.run_test(batch.i, batch.z, batch.result);` compiles unchanged.
//
// Differences, all on error or unused paths:
//  * failures throw std::runtime_error instead of printing and calling exit(1)
//    (the reference's callers catch std::exception and stop the search,
//    gpu_testing.cu:349-353);
//  * results[i] is written for i < min(active, tests.size()) only; the
//    reference also rewrote the inactive tail from stale device memory, which
//    its caller ignores (gpu_testing.cu:110-114);
//  * TPI and WINDOW_BITS are accepted and ignored: one lane owns a candidate
//    and base 2 needs no window table.  BITS is the host limb-buffer width; the
//    kernel's own limb count is picked per batch from the candidates' sizes.
#pragma once

#include <gmp.h>
#include <stdint.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "pgx.h"

inline void to_mpz(mpz_t r, uint32_t* x, uint32_t count) {
  mpz_import(r, count, -1, sizeof(uint32_t), 0, 0, x);
}

inline void from_mpz(mpz_t s, uint32_t* x, uint32_t count) {
  if (mpz_sizeinbase(s, 2) > (size_t)count * 32)
    throw std::runtime_error("from_mpz failed -- result does not fit");
  size_t words = 0;
  if (mpz_sgn(s) != 0) mpz_export(x, &words, -1, sizeof(uint32_t), 0, 0, s);
  while (words < count) x[words++] = 0;
}

template <uint32_t tpi, uint32_t bits, uint32_t window_bits>
class mr_params_t {
 public:
  static const uint32_t TPB = 0;
  static const uint32_t MAX_ROTATION = 4;
  static const uint32_t SHM_LIMIT = 0;
  static const bool CONSTANT_TIME = false;
  static const uint32_t TPI = tpi;
  static const uint32_t BITS = bits;
  static const uint32_t WINDOW_BITS = window_bits;
};

template <class params>
class test_runner_t {
 public:
  // Max number of simultaneous tests.
  const size_t n;
  // Number of Miller-Rabin bases (the first `rounds` primes); 1 = base 2.
  const size_t rounds;

  test_runner_t(size_t n, size_t rounds) : n(n), rounds(rounds), staging(n * kLimbs) {
    static_assert(params::BITS % 32 == 0, "BITS must be a multiple of 32");
    verdict.resize(n);
    int rc = pgx_runner_create(&runner, /*device=*/0, n, kLimbs, nullptr, (uint32_t)rounds);
    if (rc != PGX_OK)
      throw std::runtime_error(std::string("pgx_runner_create: ") + pgx_last_error(nullptr));
  }

  ~test_runner_t() { pgx_runner_destroy(runner); }

  test_runner_t(const test_runner_t&) = delete;
  test_runner_t& operator=(const test_runner_t&) = delete;

  void run_test(const size_t active, const std::vector<mpz_t*>& tests, std::vector<int>& results) {
    if (tests.empty()) return;
    if (tests.size() > n)
      throw std::runtime_error("tests(" + std::to_string(tests.size()) + ") > n(" +
                               std::to_string(n) + ")");
    const size_t to_run = std::min<size_t>(active, tests.size());
    for (size_t i = 0; i < to_run; i++) from_mpz(*tests[i], &staging[i * kLimbs], kLimbs);
    int rc = pgx_run(runner, staging.data(), to_run, PGX_MODE_REFERENCE, verdict.data(), nullptr);
    if (rc != PGX_OK) throw std::runtime_error(std::string("pgx_run: ") + pgx_last_error(runner));
    for (size_t i = 0; i < to_run; i++) results[i] = (verdict[i] == rounds);
  }

 private:
  static constexpr uint32_t kLimbs = params::BITS / 32;
  pgx_runner_t* runner = nullptr;
  std::vector<uint32_t> staging;
  std::vector<uint32_t> verdict;
};
