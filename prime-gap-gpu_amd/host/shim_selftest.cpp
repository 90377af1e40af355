// Drives the drop-in miller_rabin.h surface the way run_gpu_thread does
// (gpu_testing.cu:217-218, 274): test_runner_t<mr_params_t<TPI, BITS, W>>
// runner(n, rounds); runner.run_test(active, z, result) on mpz candidates
// N = m * 337#/2310 + x, and compares every verdict with GMP's
// mpz_probab_prime_p(N, 25).  Exit status 0 when all agree.
#include <gmp.h>

#include <cstdio>
#include <vector>

#include "miller_rabin.h"

int main() {
  typedef mr_params_t<4, 512, 5> params;
  const size_t n = 4096;
  mpz_t K;
  mpz_init(K);
  mpz_primorial_ui(K, 337);
  mpz_tdiv_q_ui(K, K, 2310);

  std::vector<mpz_t> store(n);
  std::vector<mpz_t*> z;
  for (auto& v : store) {
    mpz_init(v);
    z.push_back(&v);
  }
  // odd N: m odd and coprime to 2310, x even
  size_t count = 0;
  for (uint64_t m = 10000019; count < n; m += 2) {
    if (m % 3 == 0 || m % 5 == 0 || m % 7 == 0 || m % 11 == 0) continue;
    for (uint32_t x = 2; x <= 64 && count < n; x += 2) {
      mpz_mul_ui(*z[count], K, m);
      mpz_add_ui(*z[count], *z[count], x);
      ++count;
    }
  }
  int bad = 0, primes = 0;
  for (size_t rounds : {1, 3}) {
    test_runner_t<params> runner(n, rounds);
    std::vector<int> result(n, -1);
    const size_t active = n - 100;  // the tail stays untouched, like an unfilled batch
    runner.run_test(active, z, result);
    for (size_t i = 0; i < n; ++i) {
      if (i >= active) {
        bad += result[i] != -1;
        continue;
      }
      const int want = mpz_probab_prime_p(*z[i], 25) > 0;
      primes += (rounds == 1) && want;
      bad += result[i] != want;
    }
  }
  printf("shim_selftest: %zu candidates, %d primes, %d mismatches\n", n, primes, bad);
  for (auto& v : store) mpz_clear(v);
  mpz_clear(K);
  return bad == 0 && primes > 0 ? 0 : 1;
}
