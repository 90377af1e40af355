// harness.cpp — runs the drop-in gpu_testing.cpp (run_gpu_thread with its
// GPUBatch state machine, fill_batch, process_finished_batch) under the
// reference's own dispatch protocol, GPU_BATCHES threads at once, over one m
// interval whose sieve is read from a file.
//
// It compiles against the reference's headers (gap_search_gpu.h, gpu_testing.h,
// gap_common.h, gap_stats.h) as they are.  The reference's gap_search_gpu.cpp
// needs primesieve and its CLI, so this file supplies what gpu_testing.cpp
// links against from there, written from what those lines specify:
//   * the shared globals (gap_search_gpu.cpp:83-97);
//   * TestData's and GPUBatch's lock / unlock / wait_for_state_and_lock
//     (:177-200, :405-428): a spin flag taken with exchange, waiters parked on
//     the atomic, and a state wait that returns holding the lock once the state
//     is the one asked for or the search stopped;
//   * the X loop of run_testing_thread (:846-1047), with the sieve thread's
//     hand-off replaced by the file: the sieve thread finalizes each X's list
//     ahead of the tests (:711-735: the active m_i not marked composite, from
//     an active list that may predate the last X's primes, OPEN_SIEVES = 4
//     ahead), so the lists are made before the timed passes; per X the loop
//     does try_set_testing_data (:355-395: swap the list in, drop m_i already
//     found, ACTIVE, every batch EMPTY), waits for DONE, merges the batches'
//     stats, remove_vector(active, found) (:340-351), next X.
//
// Input (little-endian): "PGXH", P, d, m_start (u64), m_inc, n_x, words,
// n_active, active[n_active], then per X: X, composites[words] (bit t <->
// m_i = 2t + 1).  Output file: n_x, per X (tests, found), then the words of
// found_prime_m_i.  stdout: one summary line with the timed passes.
//
// usage: harness IN OUT [passes [verbose]]   (every pass starts from the same
// active list; the last pass's results are written; pass 1 is the warm-up)
#include <gmp.h>

#include <algorithm>
#include <atomic>
#include <cassert>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "gap_search_gpu.h"
#include "gpu_testing.h"

// ---- globals of gap_search_gpu.cpp:83-97 -----------------------------------
std::atomic<bool> is_running;
std::atomic<uint8_t> stop_queue{0};
std::mutex overflow_mtx;
std::condition_variable overflow_cv;
std::deque<std::pair<uint64_t, uint32_t>> overflowed;
std::deque<std::pair<uint64_t, uint32_t>> spot_check;

// ---- TestData / GPUBatch members defined in gap_search_gpu.cpp -------------
TestData::TestData(const struct Config config) : stats(high_resolution_clock::now()) {
  m_inc = config.m_inc;
  verbose = config.verbose;
  found_prime_m_i.resize((m_inc / 2) / 32 + 1, 0);  // bit t <-> m_i = 2t + 1
  full_reset();
}

void TestData::full_reset() {
  std::fill(found_prime_m_i.begin(), found_prime_m_i.end(), 0u);
  unknown_m_i.clear();
  test_i = 0;
  state = State::WAITING;
}

void TestData::print_stats() {
  const double t = duration<double>(high_resolution_clock::now() - stats.s_start_t).count();
  printf("harness: %lu tests, %lu primes, %lu batches (%lu partial), %.3f s on the GPU side\n",
         (unsigned long)gpu_stats.total_prp_tests, (unsigned long)gpu_stats.total_primes,
         (unsigned long)gpu_stats.batches_run, (unsigned long)gpu_stats.batches_partial, t);
}

namespace {

template <class Flag>
void spin_lock(Flag& flag) {
  while (flag.exchange(1) == 1) flag.wait(1, std::memory_order_relaxed);
}

template <class Flag>
void spin_unlock(Flag& flag) {
  assert(flag.load() == 1);
  flag = 0;
  flag.notify_one();
}

}  // namespace

// TestData and GPUBatch keep `flag` private; these members are theirs.
void TestData::lock() { spin_lock(flag); }
void TestData::unlock() { spin_unlock(flag); }
void TestData::wait_for_state_and_lock(State desired) {
  for (;;) {
    lock();
    const State current = state.load();
    if (current == desired || !is_running) return;
    unlock();
    state.wait(current);
  }
}

void GPUBatch::lock() { spin_lock(flag); }
void GPUBatch::unlock() { spin_unlock(flag); }
void GPUBatch::wait_for_state_and_lock(State desired) {
  for (;;) {
    lock();
    const State current = state.load();
    if (current == desired || !is_running) return;
    unlock();
    state.wait(current);
  }
}

// ---- the harness ------------------------------------------------------------
namespace {

struct Input {
  uint32_t P = 0, d = 0, m_inc = 0, words = 0;
  uint64_t m_start = 0;
  std::vector<uint32_t> active;
  std::vector<uint32_t> xs;
  std::vector<std::vector<uint32_t>> composites;
};

void read_exact(FILE* f, void* p, size_t bytes) {
  if (fread(p, 1, bytes, f) != bytes) throw std::runtime_error("short input file");
}

Input read_input(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) throw std::runtime_error(std::string("cannot open ") + path);
  Input in;
  char magic[4];
  uint32_t n_x = 0, n_active = 0;
  read_exact(f, magic, 4);
  if (std::string(magic, 4) != "PGXH") throw std::runtime_error("bad magic");
  read_exact(f, &in.P, 4);
  read_exact(f, &in.d, 4);
  read_exact(f, &in.m_start, 8);
  read_exact(f, &in.m_inc, 4);
  read_exact(f, &n_x, 4);
  read_exact(f, &in.words, 4);
  read_exact(f, &n_active, 4);
  in.active.resize(n_active);
  read_exact(f, in.active.data(), 4ull * n_active);
  for (uint32_t k = 0; k < n_x; ++k) {
    uint32_t X = 0;
    read_exact(f, &X, 4);
    in.xs.push_back(X);
    in.composites.emplace_back(in.words);
    read_exact(f, in.composites.back().data(), 4ull * in.words);
  }
  fclose(f);
  return in;
}

bool bit(const std::vector<uint32_t>& bm, uint32_t m_i) {
  const uint32_t t = m_i >> 1;
  return (bm[t >> 5] >> (t & 31)) & 1u;
}

// remove_vector (gap_search_gpu.cpp:340-351): drop every a whose bit is set in
// B, branch-free as there (every a is written, the index advances by 0 or 1).
void remove_found(std::vector<uint32_t>& A, const std::vector<uint32_t>& B) {
  size_t i = 0;
  for (uint32_t a : A) {
    A[i] = a;
    i += !bit(B, a);
  }
  A.resize(i);
}

struct PassResult {
  std::vector<uint32_t> tests, found;
  uint64_t total_tests = 0;
  double seconds = 0;
};

// finalize (gap_search_gpu.cpp:711-735): the tests list of every X from the
// interval's active list, as the sieve thread hands them over ahead of time.
std::vector<std::vector<uint32_t>> finalize_all(const Input& in) {
  std::vector<std::vector<uint32_t>> lists(in.xs.size());
  for (size_t k = 0; k < in.xs.size(); ++k)
    for (uint32_t m_i : in.active)
      if (!bit(in.composites[k], m_i)) lists[k].push_back(m_i);
  return lists;
}

// One pass of run_testing_thread's X loop over every X of the input.
PassResult run_pass(const Input& in, std::vector<std::vector<uint32_t>> lists,
                    TestData& test_data, std::deque<GPUBatch>& batches) {
  PassResult r;
  std::vector<uint32_t> active = in.active;
  test_data.full_reset();
  const auto t0 = high_resolution_clock::now();
  for (size_t k = 0; k < in.xs.size() && is_running; ++k) {
    const uint32_t X = in.xs[k];
    std::vector<uint32_t>& tests = lists[k];
    // try_set_testing_data (:355-395)
    test_data.lock();
    assert(test_data.state == TestData::WAITING);
    assert(test_data.unknown_m_i.empty() && test_data.running_batches == 0);
    const uint64_t primes_before = test_data.gpu_stats.total_primes;
    test_data.m_start = in.m_start;
    test_data.testing_x = X;
    test_data.unknown_m_i.swap(tests);
    remove_found(test_data.unknown_m_i, test_data.found_prime_m_i);
    const size_t n_tests = test_data.unknown_m_i.size();
    r.tests.push_back((uint32_t)n_tests);
    r.total_tests += n_tests;
    if (n_tests == 0) {  // the reference asserts a non-empty list; nothing to run at this X
      test_data.unlock();
      r.found.push_back(0);
      continue;
    }
    test_data.state = TestData::ACTIVE;
    test_data.state.notify_one();
    test_data.unlock();
    for (auto& batch : batches) {  // :928-935
      test_data.active_batches += 1;
      assert(batch.state == GPUBatch::WAITING);
      batch.state = GPUBatch::EMPTY;
      batch.state.notify_one();
    }
    test_data.wait_for_state_and_lock(TestData::DONE);  // :952-954
    if (!is_running) {
      test_data.unlock();
      break;
    }
    assert(test_data.running_batches == 0);
    assert(test_data.test_i == test_data.unknown_m_i.size());
    for (auto& batch : batches) {  // merge the batches' stats (:966-973)
      assert(batch.state == GPUBatch::WAITING);
      batch.lock();
      test_data.gpu_stats.merge(batch.stats);
      batch.stats.reset();
      batch.unlock();
    }
    r.found.push_back((uint32_t)(test_data.gpu_stats.total_primes - primes_before));
    remove_found(active, test_data.found_prime_m_i);  // :977
    test_data.test_i = 0;
    test_data.unknown_m_i.clear();
    test_data.state = TestData::WAITING;
    test_data.unlock();
  }
  r.seconds = duration<double>(high_resolution_clock::now() - t0).count();
  return r;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s IN OUT [passes [verbose]]\n", argv[0]);
    return 2;
  }
  const int passes = argc > 3 ? atoi(argv[3]) : 1;
  const int verbose = argc > 4 ? atoi(argv[4]) : 0;
  Input in;
  try {
    in = read_input(argv[1]);
  } catch (const std::exception& e) {
    fprintf(stderr, "harness: %s\n", e.what());
    return 2;
  }
  mpz_t K;  // K = P# / d (gap_common.cpp:60-65)
  mpz_init(K);
  mpz_primorial_ui(K, in.P);
  if (!mpz_divisible_ui_p(K, in.d)) {
    fprintf(stderr, "harness: d does not divide P#\n");
    return 2;
  }
  mpz_divexact_ui(K, K, in.d);
  gpu_state_and_checks(K, in.m_start + in.m_inc);

  Config config;
  config.p = in.P;
  config.d = in.d;
  config.m_start = in.m_start;
  config.m_inc = in.m_inc;
  config.verbose = verbose;
  is_running = true;
  std::deque<GPUBatch> batches;
  for (size_t i = 0; i < GPU_BATCHES; ++i) batches.emplace_back(GPU_BATCH_SIZE);
  TestData test_data{config};
  std::vector<std::thread> threads;
  for (size_t i = 0; i < GPU_BATCHES; ++i)
    threads.emplace_back(run_gpu_thread, (int)i, verbose, std::ref(test_data),
                         std::ref(batches[i]), std::ref(K));

  const std::vector<std::vector<uint32_t>> lists = finalize_all(in);
  PassResult last;
  std::vector<double> secs;
  for (int p = 0; p < passes && is_running; ++p) {
    last = run_pass(in, lists, test_data, batches);  // the lists are copied per pass
    secs.push_back(last.seconds);
  }
  const bool ok = is_running;
  // wind down as run_testing_thread does (:1030-1047)
  stop_queue = 2;
  for (auto& batch : batches) {
    batch.state = GPUBatch::EMPTY;
    batch.state.notify_all();
  }
  for (auto& t : threads) t.join();
  mpz_clear(K);
  if (!ok) {
    fprintf(stderr, "harness: the search stopped (is_running == false)\n");
    return 1;
  }

  FILE* f = fopen(argv[2], "wb");
  if (!f) {
    fprintf(stderr, "harness: cannot write %s\n", argv[2]);
    return 2;
  }
  const uint32_t n_x = (uint32_t)last.tests.size();
  fwrite(&n_x, 4, 1, f);
  for (uint32_t k = 0; k < n_x; ++k) {
    fwrite(&last.tests[k], 4, 1, f);
    fwrite(&last.found[k], 4, 1, f);
  }
  const uint32_t w = (uint32_t)test_data.found_prime_m_i.size();
  fwrite(&w, 4, 1, f);
  fwrite(test_data.found_prime_m_i.data(), 4, w, f);
  fclose(f);

  printf("harness: %d passes, %zu X, %lu tests per pass, %zu batches of up to %zu on %zu threads;"
         " seconds per pass:", passes, in.xs.size(), (unsigned long)last.total_tests,
         (size_t)test_data.gpu_stats.batches_run, GPU_BATCH_SIZE, GPU_BATCHES);
  for (double s : secs) printf(" %.6f", s);
  // per-stage seconds summed over every batch of every pass (GpuStatsCounters)
  const GpuStatsCounters& g = test_data.gpu_stats;
  printf("; stage seconds over all passes: fill %.6f queued %.6f run %.6f results %.6f"
         " (%lu batches, %lu partial)\n", g.d_fill, g.d_queued_full, g.d_run, g.d_results,
         (unsigned long)g.batches_run, (unsigned long)g.batches_partial);
  return 0;
}
