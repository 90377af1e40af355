// gpu_testing.cpp — replacement for the reference's gpu_testing.cu, driving
// the gfx950 tester through the C ABI (pgx.h).  It keeps the reference's
// interface (gpu_testing.h: GPU_BATCHES, GPU_BATCH_SIZE, run_gpu_thread,
// gpu_state_and_checks) and its TestData / GPUBatch state machine
// (gap_search_gpu.h:64-135, 191-264), so gap_search_gpu.cpp links against it
// unchanged.
//
// What changes for MI355X:
//  * GPU_BATCH_SIZE is 32768 (reference 8192, gpu_testing.cu:57-58), the
//    largest size prime_gap_test accepts (gap_search_gpu.cpp:1076-1077): one
//    lane owns one candidate, so one batch is 512 wavefronts.
//  * GPU_BATCHES is 2 (reference 3), each GPU thread's runner confined to its
//    own half of the CUs (pgx_runner_confine): each batch then has one SIMD
//    per wavefront to itself.  Three unconfined threads (1.5 wavefronts per
//    SIMD, where the older wavefronts keep the issue slots and the batches end
//    one after another) ran the reference's dispatch 25% slower (36.6-41.9
//    against 28.5-30.0 ms per pass of 1.9 M tests, same results).
//  * fill_batch only lists the multipliers m = m_start + m_i; the candidates
//    N = m * K + x are built on the device (pgx_submit_mk), replacing
//    mpz_mul_ui + mpz_add_ui per candidate (gpu_testing.cu:171-172) and the
//    mpz_export pass in run_test (miller_rabin.h:373-378).  8 bytes per
//    candidate cross PCIe instead of BITS/8.  GPUBatch::z is not used.
//  * each of the GPU_BATCHES threads owns a runner (its own stream), so up to
//    three batches run on the device at once, as the reference's threads do.
//  * gpu_state_and_checks no longer asks for a recompile per BITS
//    (gpu_testing.cu:369-385): the kernel limb count is chosen per batch from
//    the candidates' bit length, up to PGX_MAX_BITS bits.
#include <gmp.h>
#include <pthread.h>
#include <unistd.h>

#include <algorithm>
#include <cassert>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <iostream>
#include <mutex>
#include <string>
#include <vector>

#include "gap_common.h"
#include "gap_search_gpu.h"
#include "gap_stats.h"
#include "gpu_testing.h"
#include "pgx.h"

using namespace std::chrono;

#ifdef GPU_BITS
static const uint32_t kBits = GPU_BITS;
#else
static const uint32_t kBits = 1024;
#endif
static const uint32_t kLimbs = kBits / 32;

// Build-time settings, like GPU_BITS: -DPGX_GPU_BATCHES=k, -DPGX_GPU_BATCH_SIZE=n
// (n one of the sizes prime_gap_test accepts).
#ifdef PGX_GPU_BATCHES
const size_t GPU_BATCHES = PGX_GPU_BATCHES;
#else
const size_t GPU_BATCHES = 2;
#endif
#ifdef PGX_GPU_BATCH_SIZE
const size_t GPU_BATCH_SIZE = PGX_GPU_BATCH_SIZE;
#else
const size_t GPU_BATCH_SIZE = 32 * 1024;
#endif
// -DPGX_GPU_CONFINE=0 lets every GPU thread's runner use every CU.  (Confined
// 3, 4 or 6 threads measured slower than 2: a thread's CUs idle through its
// host turnaround between batches.)
#ifdef PGX_GPU_CONFINE
static const bool kConfine = PGX_GPU_CONFINE;
#else
static const bool kConfine = true;
#endif
static const uint32_t kRounds = 1;  // base-2 strong test, as the reference's ROUNDS

namespace {

// K's limbs, little-endian.
std::vector<uint32_t> limbs_of(const mpz_t K) {
  std::vector<uint32_t> out((mpz_sizeinbase(K, 2) + 31) / 32, 0);
  size_t words = 0;
  mpz_export(out.data(), &words, -1, sizeof(uint32_t), 0, 0, K);
  return out;
}

}  // namespace

// Records the primes of a finished batch (reference: gpu_testing.cu:104-133).
static uint32_t process_finished_batch(TestData& test_data, GPUBatch& batch) {
  test_data.lock();
  assert(batch.x == test_data.testing_x);
  uint32_t found = 0, last_m_i = 0;
  for (size_t i = 0; i < batch.i; i++) {
    if (!batch.active[i]) continue;
    assert(batch.result[i] == 0 || batch.result[i] == 1);
    if (batch.result[i] == 1) {
      found++;
      last_m_i = batch.m_i[i];
      test_data.add_found_prime_m_i(last_m_i);
    }
  }
  if (found > 0 && (rand() & 255) == 0) {  // sampled GMP re-check by the overflow threads
    std::lock_guard<std::mutex> lock(overflow_mtx);
    spot_check.emplace_back(test_data.m_start + last_m_i, batch.x);
  }
  test_data.unlock();
  return found;
}

// Takes the next unknown m_i at test_data.testing_x into the batch and lists
// their multipliers in `mult` (test_data.lock must be held).
static void fill_batch(TestData& test_data, GPUBatch& batch, std::vector<uint64_t>& mult) {
  assert(batch.state == GPUBatch::EMPTY);
  assert(test_data.state == TestData::ACTIVE);
  std::fill(batch.active.begin(), batch.active.end(), 0);
  std::fill(batch.result.begin(), batch.result.end(), -1);
  batch.x = test_data.testing_x;
  size_t n = 0, j = test_data.test_i;
  for (; j < test_data.unknown_m_i.size() && n < GPU_BATCH_SIZE; j++) {
    const uint32_t m_i = test_data.unknown_m_i[j];
    const uint32_t half = m_i >> 1;  // found_prime_m_i is a bitset over m_i / 2
    if (test_data.found_prime_m_i[half >> 5] & (1u << (half & 31))) continue;
    mult[n] = test_data.m_start + m_i;
    batch.active[n] = 1;
    batch.m_i[n] = m_i;
    n++;
  }
  test_data.test_i = j;
  batch.i = n;
}

void run_gpu_thread(int runner_num, int verbose, TestData& test_data, GPUBatch& batch,
                    const mpz_t& K_in) {
  pgx_runner_t* runner = nullptr;
  bool holding = false;  // this thread holds batch's lock
  bool running = false;  // ... and counts in test_data.running_batches
  try {
    pthread_setname_np(pthread_self(), ("GPU(" + std::to_string(runner_num) + ")").c_str());
    const std::vector<uint32_t> K = limbs_of(K_in);
    std::vector<uint64_t> mult(GPU_BATCH_SIZE);
    std::vector<uint32_t> verdict(GPU_BATCH_SIZE);
    if (pgx_runner_create(&runner, 0, GPU_BATCH_SIZE, kLimbs, nullptr, kRounds) != PGX_OK)
      throw std::runtime_error(std::string("pgx_runner_create: ") + pgx_last_error(nullptr));
    // each GPU thread's batches on their own share of the CUs
    if (kConfine && GPU_BATCHES > 1 &&
        pgx_runner_confine(runner, (uint32_t)runner_num % GPU_BATCHES, GPU_BATCHES) != PGX_OK)
      throw std::runtime_error(std::string("pgx_runner_confine: ") + pgx_last_error(runner));

    size_t processed = 0;
    while (is_running && stop_queue <= 1) {
      batch.wait_for_state_and_lock(GPUBatch::EMPTY);
      holding = true;
      if (!is_running || stop_queue > 1) {
        holding = false;
        batch.unlock();
        break;
      }
      test_data.lock();
      assert(!test_data.unknown_m_i.empty());
      batch.fill_start = high_resolution_clock::now();
      fill_batch(test_data, batch, mult);
      batch.fill_end = high_resolution_clock::now();
      if (batch.i == 0) {  // nothing left at this X: this batch retires
        assert(test_data.test_i == test_data.unknown_m_i.size());
        batch.state = GPUBatch::WAITING;
        holding = false;
        batch.unlock();
        if (--test_data.active_batches == 0) {
          assert(test_data.running_batches == 0);
          test_data.state = TestData::DONE;
          test_data.state.notify_all();
        }
        test_data.unlock();
        continue;
      }
      test_data.running_batches += 1;
      running = true;
      batch.state = GPUBatch::RUNNING;
      test_data.unlock();

      batch.gpu_start = high_resolution_clock::now();
      if (pgx_submit_mk(runner, K.data(), (uint32_t)K.size(), mult.data(), batch.i, batch.x, 1,
                        PGX_MODE_REFERENCE) != PGX_OK ||
          pgx_wait(runner, verdict.data(), nullptr) != PGX_OK)
        throw std::runtime_error(std::string("pgx: ") + pgx_last_error(runner));
      for (size_t i = 0; i < batch.i; i++) batch.result[i] = (verdict[i] == kRounds);
      batch.gpu_end = high_resolution_clock::now();
      processed++;

      batch.results_start = high_resolution_clock::now();
      batch.primes_in_batch = process_finished_batch(test_data, batch);
      batch.state = GPUBatch::DONE;
      batch.results_end = high_resolution_clock::now();

      batch.stats.total_prp_tests += batch.i;
      batch.stats.total_primes += batch.primes_in_batch;
      batch.stats.batches_run += 1;
      batch.stats.batches_partial += (batch.i < GPU_BATCH_SIZE);
      batch.stats.d_fill += duration<double>(batch.fill_end - batch.fill_start).count();
      batch.stats.d_queued_full += duration<double>(batch.gpu_start - batch.fill_end).count();
      batch.stats.d_run += duration<double>(batch.gpu_end - batch.gpu_start).count();
      batch.stats.d_results += duration<double>(batch.results_end - batch.results_start).count();
      if (verbose >= 4)
        printf("\tGPU(%d) batch %zu: %zu tests at X=%u, %u primes\n", runner_num, processed,
               batch.i, batch.x, batch.primes_in_batch);

      batch.state = GPUBatch::EMPTY;
      holding = false;
      batch.unlock();
      running = false;
      test_data.running_batches -= 1;
    }
    if (verbose >= 2) {
      usleep(runner_num * 10'000);
      printf("GPU(%d): Processed %zu batches\n", runner_num, processed);
    }
    if (!is_running) {
      test_data.state = TestData::DONE;
      test_data.state.notify_all();
    }
  } catch (const std::exception& e) {
    // The reference exits inside CUDA_CHECK (miller_rabin.h:57-68).  Here a
    // failing call stops the search instead: release this thread's batch and
    // move test_data to DONE, so that run_testing_thread, blocked in
    // wait_for_state_and_lock(DONE) (gap_search_gpu.cpp:953-954), wakes, sees
    // is_running == false and winds down.  (An atomic wait only returns once
    // the value changes, so the state itself must change, not just a notify.)
    std::cout << "ERROR in run_gpu_thread\n" << e.what() << std::endl;
    is_running = false;
    if (running) test_data.running_batches -= 1;
    if (holding) {
      batch.state = GPUBatch::DONE;
      batch.unlock();
    }
    test_data.state = TestData::DONE;
    test_data.state.notify_all();
  }
  if (runner) pgx_runner_destroy(runner);
}

void gpu_state_and_checks(const mpz_t& K_in, const uint64_t m_end) {
  // prime_gap_test asserts the same set (gap_search_gpu.cpp:1076-1077).
  static_assert(GPU_BATCH_SIZE == 1024 || GPU_BATCH_SIZE == 2048 || GPU_BATCH_SIZE == 4096 ||
                GPU_BATCH_SIZE == 8192 || GPU_BATCH_SIZE == 16384 || GPU_BATCH_SIZE == 32768);
  const size_t n_bits = mpz_sizeinbase(K_in, 2) + (size_t)std::log2((double)m_end) + 4;
  printf("TESTING PRIMES ON GPU (gfx950, one lane per candidate)\n");
  printf("BITS=%u (host limb buffer)\n", kBits);
  printf("PRP/BATCH=%zu\n", GPU_BATCH_SIZE);
  printf("GPU_BATCHES=%zu\n", GPU_BATCHES);
  if (n_bits > kBits) {
    printf("\nERROR: m*K has up to %zu bits but GPU_BITS=%u\n\n", n_bits, kBits);
    exit(1);
  }
  if (n_bits > PGX_MAX_BITS) {
    printf("\nERROR: m*K has up to %zu bits, the kernels hold %d\n\n", n_bits, PGX_MAX_BITS);
    exit(1);
  }
}
