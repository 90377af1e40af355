/* oracle.c — CPU oracle for the batched PRP path.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker (or as the timed CPU baseline).  The
 * product (libpgx.so) never calls it.
 *
 * Everything is computed with GMP (6.2.1 from /opt/conda in the build image):
 *   - residues a^(N-1) mod N with mpz_powm (the CPU path timed beside the GPU);
 *   - the reference's verdict chain, restated step by step from
 *     miller_rabin_t::miller_rabin (miller_rabin.h:216-248): d = (N-1)>>t,
 *     x = a^d, early pass on {1, N-1}, immediate fail when t == 1, otherwise
 *     up to t squarings (sqr_wide + rem_wide, here mpz_mul + mpz_mod) with
 *     pass on N-1 and fail on count exhaustion or x == 1;
 *   - the textbook strong test (t-1 squarings);
 *   - mpz_probab_prime_p, the reference's CPU primality call (overflow.cpp:124);
 *   - the candidate generator N = m*K + X, K = P#/d (gap_common.cpp:60-65 uses
 *     mpz_primorial_ui + mpz_tdiv_q_ui; gpu_testing.cu:171-172 mpz_mul_ui +
 *     mpz_add_ui), sieving X with mpz_fdiv_ui residues of m*K.
 * Limb packing follows from_mpz/to_mpz (miller_rabin.h:39-54): mpz_export /
 * mpz_import with order -1, size 4, native endianness.
 */
#include <gmp.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_OK 0
#define ORC_E_ARG (-1)
#define ORC_E_INPUT (-4)

static void to_mpz(mpz_t r, const uint32_t* x, uint32_t count) {
  mpz_import(r, count, -1, sizeof(uint32_t), 0, 0, x);
}

static int from_mpz(const mpz_t s, uint32_t* x, uint32_t count) {
  size_t words = 0;
  if (mpz_sizeinbase(s, 2) > (size_t)count * 32) return ORC_E_INPUT;
  memset(x, 0, count * sizeof(uint32_t));
  if (mpz_sgn(s) != 0) mpz_export(x, &words, -1, sizeof(uint32_t), 0, 0, s);
  return ORC_OK;
}

/* ---- verdicts --------------------------------------------------------- */

/* Reference verdict for one base (miller_rabin.h:221-245): 1 = pass. */
static int ref_one_base(const mpz_t N, uint32_t a) {
  mpz_t nm1, d, x, b;
  mpz_inits(nm1, d, x, b, NULL);
  mpz_sub_ui(nm1, N, 1);
  unsigned long t = mpz_scan1(nm1, 0);
  mpz_tdiv_q_2exp(d, nm1, t);
  mpz_set_ui(b, a);
  mpz_powm(x, b, d, N);
  int pass = 1;
  if (mpz_cmp_ui(x, 1) != 0 && mpz_cmp(x, nm1) != 0) {
    if (t == 1) {
      pass = 0;
    } else {
      unsigned long count = t;
      for (;;) {
        mpz_mul(x, x, x);
        mpz_mod(x, x, N);
        if (mpz_cmp(x, nm1) == 0) break;
        if (--count == 0 || mpz_cmp_ui(x, 1) == 0) {
          pass = 0;
          break;
        }
      }
    }
  }
  mpz_clears(nm1, d, x, b, NULL);
  return pass;
}

/* Textbook strong probable-prime test for one base: 1 = pass. */
static int sprp_one_base(const mpz_t N, uint32_t a) {
  mpz_t nm1, d, x, b;
  mpz_inits(nm1, d, x, b, NULL);
  mpz_sub_ui(nm1, N, 1);
  unsigned long t = mpz_scan1(nm1, 0);
  mpz_tdiv_q_2exp(d, nm1, t);
  mpz_set_ui(b, a);
  mpz_powm(x, b, d, N);
  int pass = (mpz_cmp_ui(x, 1) == 0 || mpz_cmp(x, nm1) == 0);
  for (unsigned long j = 1; !pass && j < t; ++j) {
    mpz_mul(x, x, x);
    mpz_mod(x, x, N);
    if (mpz_cmp(x, nm1) == 0) pass = 1;
    else if (mpz_cmp_ui(x, 1) == 0) break;
  }
  mpz_clears(nm1, d, x, b, NULL);
  return pass;
}

/* residue a^(N-1) mod N into res (S limbs); returns 1 if it equals 1. */
static int fermat_residue(const mpz_t N, uint32_t a, uint32_t* res, uint32_t S) {
  mpz_t e, r, b;
  mpz_inits(e, r, b, NULL);
  mpz_sub_ui(e, N, 1);
  mpz_set_ui(b, a);
  mpz_powm(r, b, e, N);
  if (res) from_mpz(r, res, S);
  int one = mpz_cmp_ui(r, 1) == 0;
  mpz_clears(e, r, b, NULL);
  return one;
}

/* mode 0: reference verdict, count of leading bases passed
 * mode 1: textbook strong, count of leading bases passed
 * mode 2: base-2 flags: 1 Fermat | 2 SPRP | 4 reference
 * residues (optional): a^(N-1) mod N of the first base, S limbs each.
 * Candidates that are even or <= 1 get 0xFFFFFFFF. */
int orc_verdicts(const uint32_t* limbs, size_t n, uint32_t S, const uint32_t* bases,
                 uint32_t nbases, uint32_t mode, uint32_t* out, uint32_t* residues) {
  if (!limbs || !out || S == 0 || (mode != 2 && (!bases || nbases == 0))) return ORC_E_ARG;
  mpz_t N;
  mpz_init(N);
  for (size_t i = 0; i < n; ++i) {
    to_mpz(N, limbs + i * S, S);
    if (mpz_cmp_ui(N, 1) <= 0 || mpz_even_p(N)) {
      out[i] = 0xFFFFFFFFu;
      if (residues) memset(residues + i * S, 0, S * sizeof(uint32_t));
      continue;
    }
    const uint32_t first = (mode == 2) ? 2u : bases[0];
    const int fermat = fermat_residue(N, first, residues ? residues + i * S : NULL, S);
    if (mode == 2) {
      out[i] = (fermat ? 1u : 0u) | (sprp_one_base(N, 2) ? 2u : 0u) |
               (ref_one_base(N, 2) ? 4u : 0u);
    } else {
      uint32_t passed = 0;
      for (uint32_t k = 0; k < nbases; ++k) {
        int ok = (mode == 0) ? ref_one_base(N, bases[k]) : sprp_one_base(N, bases[k]);
        if (!ok) break;
        ++passed;
      }
      out[i] = passed;
    }
  }
  mpz_clear(N);
  return ORC_OK;
}

int orc_probab_prime(const uint32_t* limbs, size_t n, uint32_t S, int reps, int32_t* out) {
  if (!limbs || !out || S == 0) return ORC_E_ARG;
  mpz_t N;
  mpz_init(N);
  for (size_t i = 0; i < n; ++i) {
    to_mpz(N, limbs + i * S, S);
    out[i] = mpz_probab_prime_p(N, reps);
  }
  mpz_clear(N);
  return ORC_OK;
}

/* ---- candidate generator ------------------------------------------------ */

static uint64_t splitmix(uint64_t* x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static uint64_t rotl64(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }

typedef struct {
  uint64_t s[4];
} xoshiro_t;

static uint64_t xoshiro_next(xoshiro_t* g) {
  uint64_t* s = g->s;
  const uint64_t r = rotl64(s[1] * 5, 7) * 9;
  const uint64_t t = s[1] << 17;
  s[2] ^= s[0];
  s[3] ^= s[1];
  s[1] ^= s[2];
  s[0] ^= s[3];
  s[2] ^= t;
  s[3] = rotl64(s[3], 45);
  return r;
}

/* The generator's parts, each checkable against published known answers:
 * splitmix64 outputs from a seed, and xoshiro256** outputs from a given state
 * (the authors' reference algorithms, prng.di.unimi.it). */
int orc_splitmix_stream(uint64_t seed, uint32_t count, uint64_t* out) {
  for (uint32_t i = 0; i < count; ++i) out[i] = splitmix(&seed);
  return ORC_OK;
}

int orc_xoshiro_from_state(const uint64_t* state, uint32_t count, uint64_t* out) {
  xoshiro_t g;
  for (int k = 0; k < 4; ++k) g.s[k] = state[k];
  for (uint32_t i = 0; i < count; ++i) out[i] = xoshiro_next(&g);
  return ORC_OK;
}

/* first outputs of the generator's stream: xoshiro256** seeded with four
 * splitmix64 outputs of `seed` */
int orc_xoshiro_stream(uint64_t seed, uint32_t count, uint64_t* out) {
  xoshiro_t g;
  for (int k = 0; k < 4; ++k) g.s[k] = splitmix(&seed);
  for (uint32_t i = 0; i < count; ++i) out[i] = xoshiro_next(&g);
  return ORC_OK;
}

static uint64_t below(xoshiro_t* g, uint64_t r) {
  const uint64_t reject = (0 - r) % r;
  for (;;) {
    uint64_t v = xoshiro_next(g);
    if (v >= reject) return v % r;
  }
}

static int cmp_u64(const void* a, const void* b) {
  uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return (x > y) - (x < y);
}

int orc_gen_multipliers(uint64_t seed, uint64_t m_lo, uint64_t m_hi, uint32_t d, uint32_t n_m,
                        uint64_t* out_m) {
  if (!out_m || m_hi <= m_lo || d == 0) return ORC_E_ARG;
  xoshiro_t g;
  for (int k = 0; k < 4; ++k) g.s[k] = splitmix(&seed);
  uint32_t have = 0;
  mpz_t mm;
  mpz_init(mm);
  uint64_t tries = 0;
  while (have < n_m) {
    if (++tries > 100000000ull) {
      mpz_clear(mm);
      return ORC_E_ARG;
    }
    uint64_t m = m_lo + below(&g, m_hi - m_lo);
    mpz_set_ui(mm, m);
    if (mpz_gcd_ui(NULL, mm, d) != 1) continue;
    int dup = 0;
    for (uint32_t k = 0; k < have; ++k) dup |= out_m[k] == m;
    if (dup) continue;
    out_m[have++] = m;
  }
  mpz_clear(mm);
  qsort(out_m, n_m, sizeof(uint64_t), cmp_u64);
  return ORC_OK;
}

/* survivors of X in [1, x_max] for each m (N = m*K + X with no prime factor
 * below sieve_limit), first per_m of each in ascending X. */
int orc_gen_sieved_side(uint32_t P, uint32_t d, int32_t sign, const uint64_t* m, uint32_t n_m,
                        uint32_t x_max, uint32_t per_m, uint32_t sieve_limit, uint32_t S,
                        uint32_t* out_limbs, uint32_t* out_m, uint32_t* out_x, size_t* written) {
  if (!m || !out_limbs || !written || S == 0 || x_max == 0 || d == 0 ||
      (sign != 1 && sign != -1))
    return ORC_E_ARG;
  *written = 0;
  mpz_t K, mK, N;
  mpz_inits(K, mK, N, NULL);
  mpz_primorial_ui(K, P);
  if (!mpz_divisible_ui_p(K, d)) {
    mpz_clears(K, mK, N, NULL);
    return ORC_E_INPUT;
  }
  mpz_tdiv_q_ui(K, K, d);
  /* primes below the limit (Eratosthenes) */
  uint8_t* comp_p = calloc(sieve_limit + 1, 1);
  uint32_t* primes = malloc(sizeof(uint32_t) * (sieve_limit / 2 + 2));
  uint32_t np = 0;
  for (uint32_t p = 2; p < sieve_limit; ++p) {
    if (comp_p[p]) continue;
    primes[np++] = p;
    for (uint64_t q = (uint64_t)p * p; q < sieve_limit; q += p) comp_p[q] = 1;
  }
  free(comp_p);
  uint8_t* comp = malloc(x_max + 1);
  size_t w = 0;
  int rc = ORC_OK;
  for (uint32_t mi = 0; mi < n_m && rc == ORC_OK; ++mi) {
    memset(comp, 0, x_max + 1);
    mpz_mul_ui(mK, K, m[mi]);
    for (uint32_t k = 0; k < np; ++k) {
      const uint32_t p = primes[k];
      const uint64_t r = mpz_fdiv_ui(mK, p);
      uint64_t x = (sign > 0) ? (p - r) % p : r; /* m*K + sign*x == 0 mod p */
      if (x == 0) x = p;
      for (; x <= x_max; x += p) comp[x] = 1;
    }
    uint32_t taken = 0;
    for (uint32_t x = 1; x <= x_max && taken < per_m; ++x) {
      if (comp[x]) continue;
      if (sign > 0)
        mpz_add_ui(N, mK, x);
      else
        mpz_sub_ui(N, mK, x);
      if (mpz_sgn(N) <= 0 || from_mpz(N, out_limbs + w * S, S) != ORC_OK) {
        rc = ORC_E_INPUT;
        break;
      }
      if (out_m) out_m[w] = mi;
      if (out_x) out_x[w] = x;
      ++w;
      ++taken;
    }
  }
  free(comp);
  free(primes);
  mpz_clears(K, mK, N, NULL);
  *written = w;
  return rc;
}

int orc_gen_sieved(uint32_t P, uint32_t d, const uint64_t* m, uint32_t n_m, uint32_t x_max,
                   uint32_t per_m, uint32_t sieve_limit, uint32_t S, uint32_t* out_limbs,
                   uint32_t* out_m, uint32_t* out_x, size_t* written) {
  return orc_gen_sieved_side(P, d, 1, m, n_m, x_max, per_m, sieve_limit, S, out_limbs, out_m,
                             out_x, written);
}

/* smallest prime factor of N below `limit`, or 0 (brute-force trial division) */
int orc_small_factor(const uint32_t* limbs, size_t n, uint32_t S, uint32_t limit,
                     uint32_t* out) {
  mpz_t N;
  mpz_init(N);
  for (size_t i = 0; i < n; ++i) {
    to_mpz(N, limbs + i * S, S);
    out[i] = 0;
    for (uint32_t p = 2; p < limit; p += (p == 2) ? 1 : 2) {
      if (mpz_cmp_ui(N, p) <= 0) break;
      if (mpz_divisible_ui_p(N, p)) {
        out[i] = p;
        break;
      }
    }
  }
  mpz_clear(N);
  return ORC_OK;
}

/* ---- CPU baseline: threaded mpz_powm(2, N-1, N) ----------------------- */

typedef struct {
  const uint32_t* limbs;
  size_t lo, hi;
  uint32_t S;
  uint32_t ones;
} shard_t;

static void* shard_run(void* arg) {
  shard_t* s = (shard_t*)arg;
  mpz_t N, e, r, two;
  mpz_inits(N, e, r, two, NULL);
  mpz_set_ui(two, 2);
  uint32_t ones = 0;
  for (size_t i = s->lo; i < s->hi; ++i) {
    to_mpz(N, s->limbs + i * s->S, s->S);
    mpz_sub_ui(e, N, 1);
    mpz_powm(r, two, e, N);
    ones += mpz_cmp_ui(r, 1) == 0;
  }
  s->ones = ones;
  mpz_clears(N, e, r, two, NULL);
  return NULL;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Wall seconds for one pass of mpz_powm(2, N-1, N) over n candidates with
 * `threads` static shards (after one untimed warm-up pass over the first
 * shard's share when warmup != 0).  *fermat_count receives the number of
 * residues equal to 1. */
double orc_time_powm2(const uint32_t* limbs, size_t n, uint32_t S, int threads, int warmup,
                      uint32_t* fermat_count) {
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  shard_t sh[256];
  pthread_t th[256];
  if (warmup) {
    shard_t w = {limbs, 0, n < 64 ? n : 64, S, 0};
    shard_run(&w);
  }
  for (int k = 0; k < threads; ++k) {
    sh[k].limbs = limbs;
    sh[k].lo = n * k / threads;
    sh[k].hi = n * (k + 1) / threads;
    sh[k].S = S;
    sh[k].ones = 0;
  }
  double t0 = now_s();
  for (int k = 0; k < threads; ++k) pthread_create(&th[k], NULL, shard_run, &sh[k]);
  uint32_t ones = 0;
  for (int k = 0; k < threads; ++k) {
    pthread_join(th[k], NULL);
    ones += sh[k].ones;
  }
  double t1 = now_s();
  if (fermat_count) *fermat_count = ones;
  return t1 - t0;
}

/* ---- candidate construction (fill_batch, gpu_testing.cu:136-195) --------- */

/* out[i] (S words, candidate-major) = m[i]*K + sign*x, with fill_batch's GMP
 * calls (mpz_mul_ui, then mpz_add_ui or mpz_sub_ui; gpu_testing.cu:171-172)
 * and from_mpz's packing (mpz_export, miller_rabin.h:39-54).  *seconds gets the
 * wall time of the loop on one thread.  ORC_E_INPUT if a value does not fit S
 * words or is negative. */
int orc_build_mk(const uint32_t* K, uint32_t k_words, const uint64_t* m, size_t n, uint32_t x,
                 int sign, uint32_t S, uint32_t* out, double* seconds) {
  if (!K || !k_words || (n && (!m || !out)) || (sign != 1 && sign != -1) || !S)
    return ORC_E_ARG;
  mpz_t k, z;
  mpz_inits(k, z, NULL);
  to_mpz(k, K, k_words);
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int rc = ORC_OK;
  for (size_t i = 0; i < n && rc == ORC_OK; ++i) {
    mpz_mul_ui(z, k, (unsigned long)m[i]);
    if (sign > 0) mpz_add_ui(z, z, x);
    else mpz_sub_ui(z, z, x);
    if (mpz_sgn(z) < 0) rc = ORC_E_INPUT;
    else rc = from_mpz(z, out + i * S, S);
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (seconds) *seconds = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
  mpz_clears(k, z, NULL);
  return rc;
}

/* ---- overflow search (overflow.cpp:142-160) ------------------------------ */

/* Offset X of the next prime above m*K (sign > 0: mpz_nextprime(m*K), the
 * reference's next side with min_x = 0) or of the previous prime below it
 * (sign < 0: the reference calls mpz_prevprime, which GMP 6.2.1 lacks; here
 * odd y = m*K - 1, m*K - 3, ... with mpz_probab_prime_p(y, 25), the test GMP's
 * own search uses).  *x = 0 when the offset exceeds limit. */
int orc_next_prime_offset(const uint32_t* K, uint32_t k_words, uint64_t m, int sign,
                          uint32_t limit, uint32_t* x) {
  if (!K || !k_words || !x || (sign != 1 && sign != -1)) return ORC_E_ARG;
  mpz_t k, c, y;
  mpz_inits(k, c, y, NULL);
  to_mpz(k, K, k_words);
  mpz_mul_ui(c, k, (unsigned long)m);
  *x = 0;
  if (sign > 0) {
    mpz_nextprime(y, c);
    mpz_sub(y, y, c);
    if (mpz_cmp_ui(y, limit) <= 0) *x = (uint32_t)mpz_get_ui(y);
  } else {
    mpz_sub_ui(y, c, mpz_odd_p(c) ? 2 : 1);
    for (uint32_t off = mpz_odd_p(c) ? 2 : 1; off <= limit; off += 2) {
      if (mpz_probab_prime_p(y, 25)) {
        *x = off;
        break;
      }
      mpz_sub_ui(y, y, 2);
    }
  }
  mpz_clears(k, c, y, NULL);
  return ORC_OK;
}

