/* msieve.c — CPU oracle for the m-interval search step.  TEST INFRASTRUCTURE ONLY.
 *
 * Restates, for one X at a time, what the reference's sieve thread and GPU
 * thread do to the interval of multipliers m = m_start + m_i (m_i odd,
 * 0 < m_i < m_inc):
 *   - orc_ms_active   SieveData::is_coprime_and_valid_m (gap_search_gpu.cpp:263-304):
 *                     odd m_i with gcd(m_start + m_i, d) = 1, ascending;
 *   - orc_ms_coprime_x SieveData constructor (gap_search_gpu.cpp:227-256): even X in
 *                     [2, 10P) divisible by no prime <= P that does not divide d;
 *   - orc_ms_sieve    run_sieve_thread (gap_search_gpu.cpp:462-699): bit t of the
 *                     composites bitmap marks m_i = 2t + 1 whose m*K + X has an odd
 *                     factor q < max_prime, q not dividing K: primes dividing d by
 *                     the d wheel (:577-616), primes below 100000 (:619-639),
 *                     the others (:651-689);
 *   - orc_ms_step     finalize (:711-735) + fill_batch (gpu_testing.cu:136-195) +
 *                     process_finished_batch (:104-133) + remove_vector
 *                     (gap_search_gpu.cpp:340-351): tests = active m_i that are
 *                     neither composite nor already found, in ascending order;
 *                     each N = (m_start + m_i) K + X gets the reference verdict for
 *                     every base (GMP); passing m_i join the found bitmap and leave
 *                     the active list.
 *
 * `as_written` selects the reference's arithmetic exactly as written for the
 * primes >= 100000: X * neg_inv_K is a 32-bit product (it wraps once it
 * reaches 2^32), and when X * max_prime >= m_start the start offset is not
 * halved before indexing the half-interval bitmap (:671-689), so those primes
 * mark m_i = 2 t + 1 for t = f, f + p, ... where f is the full-interval
 * offset.  With as_written = 0 every prime marks exactly the m_i it divides.
 */
#include <gmp.h>
#include <stdint.h>
#include <time.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_E_ARG (-1)

#define SMALL_PRIME_LIMIT 100000u

static uint32_t* odd_primes_below(uint64_t limit, size_t* count) {
  /* plain sieve of Eratosthenes over [0, limit) */
  uint8_t* comp = calloc(limit ? limit : 1, 1);
  size_t n = 0;
  for (uint64_t i = 3; i < limit; i += 2)
    if (!comp[i]) {
      ++n;
      for (uint64_t j = i * i; j < limit; j += 2 * i) comp[j] = 1;
    }
  uint32_t* out = malloc((n ? n : 1) * sizeof(uint32_t));
  n = 0;
  for (uint64_t i = 3; i < limit; i += 2)
    if (!comp[i]) out[n++] = (uint32_t)i;
  free(comp);
  *count = n;
  return out;
}

/* modular inverse by the extended Euclidean algorithm, 0 < a < p */
static uint32_t inverse_mod(uint32_t a, uint32_t p) {
  int64_t r0 = p, r1 = a, s0 = 0, s1 = 1;
  while (r1) {
    int64_t q = r0 / r1, t;
    t = r0 - q * r1; r0 = r1; r1 = t;
    t = s0 - q * s1; s0 = s1; s1 = t;
  }
  return (uint32_t)(s0 < 0 ? s0 + p : s0);
}

static void set_bit(uint32_t* bits, uint64_t t) { bits[t >> 5] |= 1u << (t & 31); }
static int get_bit(const uint32_t* bits, uint64_t t) { return (bits[t >> 5] >> (t & 31)) & 1u; }

int orc_ms_active(uint32_t P, uint32_t d, uint64_t m_start, uint32_t m_inc, uint32_t* out,
                  size_t* n) {
  if ((m_start & 1) || (m_inc & 1) || (d & 1) || !n) return ORC_E_ARG;
  const uint32_t half = m_inc / 2;
  uint8_t* bad = calloc(half ? half : 1, 1);
  size_t np = 0;
  uint32_t* pr = odd_primes_below((uint64_t)P + 1, &np);
  for (size_t k = 0; k < np; ++k) {
    const uint32_t p = pr[k];
    if (d % p) continue;
    uint64_t first = (p - m_start % p) % p; /* m_start + first = 0 mod p */
    if (!(first & 1)) first += p;
    for (uint64_t i = first >> 1; i < half; i += p) bad[i] = 1;
  }
  size_t c = 0;
  for (uint32_t i = 0; i < half; ++i)
    if (!bad[i]) {
      if (out) out[c] = 2 * i + 1;
      ++c;
    }
  *n = c;
  free(bad);
  free(pr);
  return ORC_OK;
}

int orc_ms_coprime_x(uint32_t P, uint32_t d, uint32_t* out, size_t cap, size_t* n) {
  if (!n) return ORC_E_ARG;
  size_t np = 0;
  uint32_t* pr = odd_primes_below((uint64_t)P + 1, &np);
  size_t c = 0;
  for (uint32_t x = 2; x < 10 * P; x += 2) {
    int hit = 0;
    for (size_t k = 0; k < np && !hit; ++k) hit = (d % pr[k] != 0) && (x % pr[k] == 0);
    if (!hit) {
      if (out && c < cap) out[c] = x;
      ++c;
    }
  }
  *n = c;
  free(pr);
  return ORC_OK;
}

static double seconds_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* composites: m_inc/64 + 2 words (bit t <-> m_i = 2t + 1), cleared here.
 * *mark_seconds (optional): time of the per-X work alone (clearing, start
 * offsets, marking), i.e. without the prime list and the inverses of K, which
 * the reference computes once per interval (gap_search_gpu.cpp:476-499). */
int orc_ms_sieve(const uint32_t* K, uint32_t k_words, uint32_t P, uint32_t d, uint64_t m_start,
                 uint32_t m_inc, uint64_t max_prime, uint32_t X, int as_written,
                 uint32_t* composites, double* mark_seconds) {
  if (!K || !k_words || (m_start & 1) || (m_inc & 1) || (d & 1) || (X & 1) || !composites ||
      max_prime > 2000000000ull)
    return ORC_E_ARG;
  const uint64_t half = m_inc / 2;
  mpz_t k;
  mpz_init(k);
  mpz_import(k, k_words, -1, sizeof(uint32_t), 0, 0, K);
  size_t np = 0;
  uint32_t* pr = odd_primes_below(max_prime, &np);
  uint32_t* kmod = malloc((np ? np : 1) * sizeof(uint32_t));
  uint32_t* nik = malloc((np ? np : 1) * sizeof(uint32_t));
  for (size_t i = 0; i < np; ++i) {
    const uint32_t p = pr[i];
    kmod[i] = (uint32_t)mpz_fdiv_ui(k, p);
    nik[i] = (p <= P) ? 0 : p - inverse_mod(kmod[i], p);
  }
  const double t_start = seconds_now();
  memset(composites, 0, (half / 32 + 2) * sizeof(uint32_t));
  for (size_t i = 0; i < np; ++i) {
    const uint32_t p = pr[i];
    if (p <= P) {
      if (d % p) continue; /* divides K: X is chosen coprime to it */
      /* d wheel: the first odd m_i whose m K + X is divisible by p, then every p-th bit */
      const uint64_t k_mod = kmod[i];
      const uint64_t shift = m_start % p;
      for (uint64_t mi = 1; mi < 2ull * p + 1; mi += 2)
        if (((shift + mi) * k_mod + X) % p == 0) {
          for (uint64_t t = mi >> 1; t < half; t += p) set_bit(composites, t);
          break;
        }
      continue;
    }
    const uint32_t neg_inv = nik[i];
    /* f: the smallest odd offset with p | (m_start + f) K + X */
    uint64_t f, t0;
    if (p < SMALL_PRIME_LIMIT || !as_written) {
      f = ((uint64_t)X * neg_inv + (p - m_start % p)) % p;
      if (!(f & 1)) f += p;
      t0 = f >> 1;
    } else {
      const uint64_t xk = (uint32_t)(X * neg_inv); /* 32-bit product as written */
      if ((uint64_t)X * max_prime < m_start) {
        f = (m_start - xk) % p;
        f = f ? p - f : 0;
        if (!(f & 1)) f += p;
        t0 = f >> 1;
      } else {
        f = (xk + (p - m_start % p)) % p;
        if (!(f & 1)) f += p;
        t0 = f; /* not halved */
      }
    }
    for (uint64_t t = t0; t < half; t += p) set_bit(composites, t);
  }
  if (mark_seconds) *mark_seconds = seconds_now() - t_start;
  free(kmod);
  free(nik);
  free(pr);
  mpz_clear(k);
  return ORC_OK;
}

/* Reference verdict for one base (miller_rabin.h:221-245), 1 = pass. */
static int ref_pass(const mpz_t N, uint32_t a) {
  mpz_t nm1, e, x, b;
  mpz_inits(nm1, e, x, b, NULL);
  mpz_sub_ui(nm1, N, 1);
  const unsigned long t = mpz_scan1(nm1, 0);
  mpz_tdiv_q_2exp(e, nm1, t);
  mpz_set_ui(b, a);
  mpz_powm(x, b, e, N);
  int pass = 1;
  if (mpz_cmp_ui(x, 1) != 0 && mpz_cmp(x, nm1) != 0) {
    pass = 0;
    if (t > 1) { /* up to t squarings, as written */
      for (unsigned long count = t;;) {
        mpz_mul(x, x, x);
        mpz_mod(x, x, N);
        if (mpz_cmp(x, nm1) == 0) { pass = 1; break; }
        if (--count == 0 || mpz_cmp_ui(x, 1) == 0) break;
      }
    }
  }
  mpz_clears(nm1, e, x, b, NULL);
  return pass;
}

/* One X step.  active (in/out, ascending, *n_active entries), found (in/out
 * bitmap, m_inc/64 + 2 words), composites from orc_ms_sieve.  tests[] receives
 * the tested m_i in order, found_mi[] those that passed every base. */
int orc_ms_step(const uint32_t* K, uint32_t k_words, uint64_t m_start, uint32_t X,
                const uint32_t* composites, uint32_t* active, size_t* n_active, uint32_t* found,
                const uint32_t* bases, uint32_t nbases, uint32_t* tests, size_t* n_tests,
                uint32_t* found_mi, size_t* n_found) {
  if (!K || !composites || !active || !n_active || !found || !bases || !nbases || !n_tests ||
      !n_found)
    return ORC_E_ARG;
  mpz_t k, N;
  mpz_inits(k, N, NULL);
  mpz_import(k, k_words, -1, sizeof(uint32_t), 0, 0, K);
  size_t nt = 0, nf = 0;
  for (size_t i = 0; i < *n_active; ++i) {
    const uint32_t mi = active[i];
    const uint32_t t = mi >> 1;
    if (get_bit(composites, t) || get_bit(found, t)) continue;
    if (tests) tests[nt] = mi;
    ++nt;
    mpz_mul_ui(N, k, m_start + mi);
    mpz_add_ui(N, N, X);
    int pass = 1;
    for (uint32_t b = 0; b < nbases && pass; ++b) pass = ref_pass(N, bases[b]);
    if (pass) {
      if (found_mi) found_mi[nf] = mi;
      ++nf;
      set_bit(found, t);
    }
  }
  size_t w = 0; /* remove_vector */
  for (size_t i = 0; i < *n_active; ++i)
    if (!get_bit(found, active[i] >> 1)) active[w++] = active[i];
  *n_active = w;
  *n_tests = nt;
  *n_found = nf;
  mpz_clears(k, N, NULL);
  return ORC_OK;
}
