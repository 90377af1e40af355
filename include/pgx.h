/* pgx.h — C ABI of the MI355X (gfx950) batched probable-prime tester.
 *
 * Plain C types only (pointers, sizes, status codes).  Every entry point
 * returns PGX_OK (0) or a negative PGX_E_* code; nothing calls exit() and no
 * C++ exception crosses this boundary.  The per-runner message behind a failing
 * call is available from pgx_last_error().
 *
 * The entry points replace the CUDA runner of the reference, one for one:
 *
 *   reference (sethtroisi/prime-gap-gpu)                     here
 *   ------------------------------------------------------   ----------------------------
 *   test_runner_t<params>::test_runner_t(n, rounds)          pgx_runner_create
 *     miller_rabin.h:322-349 (pinned + device buffers,
 *     base list generate_primes(rounds) :272-289, stream)
 *   test_runner_t::~test_runner_t   miller_rabin.h:351-362   pgx_runner_destroy
 *   test_runner_t::run_test         miller_rabin.h:364-408   pgx_submit + pgx_wait, or pgx_run
 *     (pack mpz -> BITS/32 LE u32 limbs :373-378, H2D :383,
 *      kernel_miller_rabin :389-394, D2H :398, sync :402,
 *      results[i] = (passed == rounds) :405-407)
 *   CUDA_CHECK / CGBN_CHECK / from_mpz exit(1)               negative status + pgx_last_error
 *     miller_rabin.h:43-97
 *   kernel_miller_rabin<params>     miller_rabin.h:252-270   pgx_test_device (device-resident
 *                                                            inputs, caller's stream)
 *   fill_batch (N = m*K + x)        gpu_testing.cu:136-195   pgx_gen_* (host candidate builder)
 *
 * Limb layout.  Host arrays passed to pgx_submit/pgx_run are candidate-major,
 * exactly the reference's instance_t limbs: candidate i, limb j (little-endian
 * 32-bit words, mpz_export(order=-1, size=4)) at [i * limbs + j].  Device
 * arrays (pgx_test_device) are limb-interleaved: limb j of candidate i at
 * [j * ld + i], so a wavefront reads each limb row with one coalesced load.
 */
#ifndef PGX_H
#define PGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#include "pgx_defs.h"

typedef struct pgx_runner pgx_runner_t;

/* Create a runner on `device` for up to max_n candidates stored in `limbs`
 * 32-bit words each (the reference's BITS/32).  `bases` lists the Miller-Rabin
 * bases in test order (the reference passes generate_primes(rounds) = the
 * first `rounds` primes, miller_rabin.h:272-289); NULL with nbases = r means
 * the first r primes, as there.  Each base must be in [2, PGX_MAX_BASE].
 * Replaces test_runner_t's constructor, miller_rabin.h:322-349. */
int pgx_runner_create(pgx_runner_t** out, int device, size_t max_n, uint32_t limbs,
                      const uint32_t* bases, uint32_t nbases);

/* Free everything the runner owns.  Replaces miller_rabin.h:351-362. */
int pgx_runner_destroy(pgx_runner_t* h);

/* Confines the runner's kernels to compute units [part*n/parts,
 * (part+1)*n/parts) of its device's n (a CU mask on its streams; parts = 1
 * lifts it).  For callers that keep several runners' batches in flight at
 * once, as the reference's GPU_BATCHES threads do (gpu_testing.cu:203-354, one
 * runner and stream each): unconfined, their kernels share SIMDs and the
 * older wavefronts take the issue slots, so concurrent batches end one after
 * another; confined, each runs on its own CUs.  No counterpart in the
 * reference.  Only between batches (PGX_E_STATE while one is in flight); work
 * queued on the runner's own stream finishes first.  PGX_E_HIP when the
 * process has no hardware queue left for a masked stream. */
int pgx_runner_confine(pgx_runner_t* h, uint32_t part, uint32_t parts);

/* Copy n candidates (candidate-major limbs, see above) into pinned staging,
 * send them to the device as they are (a device kernel makes them
 * limb-interleaved) and launch asynchronously.  The caller's array may be
 * reused as soon as this returns.  Up to two batches may be in flight on one
 * runner (each with its own staging, buffers and stream, so the second runs
 * beside the first); a third submit is PGX_E_STATE until a pgx_wait, which
 * returns the batches in submit order.  A runner is used by one host thread
 * at a time (the reference keeps one runner per GPU thread).  With several bases
 * in the reference or strong mode the first base runs on every candidate and
 * the others, side by side, on the candidates that passed it (compacted on
 * the device); pgx_test_device runs every base in one pass.  The verdicts are
 * the same.  Rejects even N,
 * N <= 1 and N wider than PGX_MAX_BITS with PGX_E_INPUT (the reference
 * asserts / exits, miller_rabin.h:46-49, 375-376). */
int pgx_submit(pgx_runner_t* h, const uint32_t* limbs_le, size_t n, uint32_t flags);

/* Wait for the oldest batch in flight.  verdict[i] (n entries): the mode's verdict
 * word (reference mode: leading bases passed, compare with nbases as
 * miller_rabin.h:406 does).  residues (n * limbs words, candidate-major) is
 * filled when PGX_WANT_RESIDUE was submitted, else may be NULL. */
int pgx_wait(pgx_runner_t* h, uint32_t* verdict, uint32_t* residues);

/* pgx_submit + pgx_wait. */
int pgx_run(pgx_runner_t* h, const uint32_t* limbs_le, size_t n, uint32_t flags,
            uint32_t* verdict, uint32_t* residues);

/* Device-resident batch: candidates already in HBM, limb-interleaved with
 * leading dimension ld >= n and `limbs` rows; max_bits bounds the bit length of
 * every candidate (it selects the kernel's limb count S = ceil((max_bits+2)/32)).
 * verdict: n words; residues: `limbs` rows of res_ld words or NULL.  Launches
 * on `stream` (a hipStream_t; NULL = the runner's stream) and returns without
 * waiting.  Replaces kernel_miller_rabin's launch, miller_rabin.h:388-394. */
int pgx_test_device(pgx_runner_t* h, const uint32_t* d_cand, uint32_t ld, uint32_t limbs,
                    size_t n, uint32_t max_bits, uint32_t flags, uint32_t* d_verdict,
                    uint32_t* d_residues, uint32_t res_ld, void* stream);

/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  pgx_submit's batch on device-resident candidates without a
 * runner: the first of `bases` (NULL: the first nbases primes) on every
 * column, then, with several bases outside FERMAT mode, the columns that
 * passed it compacted on the device and the other bases side by side on
 * them.  Verdicts (and residues of the first base with PGX_WANT_RESIDUE) are
 * those of pgx_test_device / pgx_run.  Layout and max_bits as pgx_test_device;
 * launches on `stream` (NULL = the device's null stream) and returns without
 * waiting.  Allocates nothing: its scratch (survivor list, pass masks,
 * count; 8 MiB) is a device array of the library, one per GPU, held until the
 * process ends; n is limited to PGX_NOALLOC_MAX_N (PGX_E_CAPACITY above).
 * Calls that overlap on one GPU share that scratch and race: serialize them
 * (one stream, or events).  Messages: pgx_last_error(NULL) on this thread. */
int pgx_test_device_noalloc(int device, const uint32_t* d_cand, uint32_t ld, uint32_t limbs,
                            size_t n, uint32_t max_bits, uint32_t flags, const uint32_t* bases,
                            uint32_t nbases, uint32_t* d_verdict, uint32_t* d_residues,
                            uint32_t res_ld, void* stream);

/* fill_batch on the device (gpu_testing.cu:136-195 builds N = m*K + x with
 * mpz_mul_ui + mpz_add_ui per candidate, then run_test exports the limbs).
 * pgx_submit_mk uploads K (k_words little-endian words; cached while the same
 * pointer and size come back) and n multipliers m[i], builds
 * N_i = m[i]*K + sign*x on the device in the runner's limb buffers and launches
 * the test; pgx_wait returns the verdicts as for pgx_submit, or PGX_E_INPUT when
 * a candidate needed more than the runner's `limbs` words or was negative.  An
 * even candidate or one <= 1 (which a sieve never passes) gets
 * PGX_VERDICT_INVALID. */
int pgx_submit_mk(pgx_runner_t* h, const uint32_t* K, uint32_t k_words, const uint64_t* m,
                  size_t n, uint32_t x, int32_t sign, uint32_t flags);

/* The construction kernel alone, on device buffers: column i of d_out
 * (out_words rows, leading dimension ld) = d_m[i]*K + sign*(d_x ? d_x[i] : x0).
 * *d_status (device word, caller-cleared) becomes nonzero when a candidate
 * does not fit out_words words or is negative. */
int pgx_build_candidates(pgx_runner_t* h, const uint32_t* d_K, uint32_t k_words,
                         const uint64_t* d_m, const uint32_t* d_x, uint32_t x0, int32_t sign,
                         uint32_t n, uint32_t* d_out, uint32_t ld, uint32_t out_words,
                         uint32_t* d_status, void* stream);

/* ---- m-interval search: sieve, test and bookkeeping on the device ------------
 * One object per interval of multipliers m = m_start + m_i (m_i odd, m_i <
 * m_inc), driven one X at a time like the reference's sieve thread and GPU
 * thread (gap_search_gpu.cpp:462-804, gpu_testing.cu:104-354).  Every per-m set
 * is a bitmap in HBM, bit t <-> m_i = 2t + 1, m_inc/64 + 2 words:
 *   active      odd m_i with gcd(m_start + m_i, d) = 1 not yet found prime
 *               (is_coprime_and_valid_m :263-304, remove_vector :340-351);
 *   composites  m_i whose m*K + X has an odd factor q < max_prime with q > P or
 *               q | d, for the last sieved X (run_sieve_thread :577-689);
 *   found       m_i with a probable prime m*K + X at some tested X
 *               (TestData::found_prime_m_i, process_finished_batch).
 * The runner supplies the device, stream, bases (rounds) and the batch buffers:
 * a step borrows them, so the runner takes no other batch meanwhile. */
typedef struct pgx_search pgx_search_t;

#define PGX_SEARCH_AS_WRITTEN 1u /* primes >= 100000 use the reference's
                                    arithmetic as written (:651-689): a 32-bit
                                    X*neg_inv_K, and an un-halved start when
                                    X*max_prime >= m_start */
#define PGX_SEARCH_ACTIVE 0u
#define PGX_SEARCH_COMPOSITES 1u
#define PGX_SEARCH_FOUND 2u

typedef struct {
  uint32_t P;         /* K = P# / d */
  uint32_t d;         /* even */
  uint64_t m_start;   /* even */
  uint32_t m_inc;     /* even, < 2^31 */
  uint32_t max_prime; /* sieve limit (exclusive), 1000 .. 2e9 */
  uint32_t flags;     /* PGX_SEARCH_AS_WRITTEN */
} pgx_search_config_t;

/* K: k_words little-endian words, odd.  PGX_E_INPUT when (m_start+m_inc)*K
 * does not fit the runner's words. */
int pgx_search_create(pgx_search_t** out, pgx_runner_t* runner, const uint32_t* K,
                      uint32_t k_words, const pgx_search_config_t* cfg);
int pgx_search_destroy(pgx_search_t* s);
/* Sieve the interval at X (even) and list the tests: active m_i that are
 * neither composite at X nor found, ascending (the reference's next_sieves
 * entry for X).  *n_tests = their count. */
int pgx_search_sieve(pgx_search_t* s, uint32_t X, uint32_t* n_tests);
/* Test every listed m_i: N = (m_start + m_i) K + X with the runner's bases
 * (every base must pass, miller_rabin.h:405-407); passing m_i join found and
 * leave active.  *n_found = how many passed at this X. */
int pgx_search_test(pgx_search_t* s, uint32_t* n_found);
int pgx_search_step(pgx_search_t* s, uint32_t X, uint32_t* n_tests, uint32_t* n_found);
/* The tests list of the last sieved X (cap entries; *n = its length). */
int pgx_search_tests(pgx_search_t* s, uint32_t* out, size_t cap, size_t* n);
/* Every (m_i, X) found so far, ascending in m_i. */
int pgx_search_found(pgx_search_t* s, uint32_t* m_i, uint32_t* x, size_t cap, size_t* n);
/* A copy of one bitmap (PGX_SEARCH_ACTIVE / _COMPOSITES / _FOUND). */
int pgx_search_bitmap(pgx_search_t* s, uint32_t which, uint32_t* out, size_t cap, size_t* n);
const char* pgx_search_last_error(pgx_search_t* s);

/* Message of the last failing call on h (or, when h is NULL, of the last
 * failing pgx_runner_create on the calling thread).  Never NULL. */
const char* pgx_last_error(const pgx_runner_t* h);

/* ---- batched persistent launches ----------------------------------------
 * The reference tests one fixed-BITS batch per launch from three host threads
 * (gpu_testing.cu:203-354, miller_rabin.h:383-402) and needs one binary per
 * BITS (gpu_testing.cu:369-385).  Here candidates of any mix of sizes are
 * grouped into buckets (one kernel limb count each), every bucket runs as a
 * persistent kernel whose resident wavefronts take chunks of candidates from
 * a device counter, and a plan captures all buckets (counters reset by a
 * fill-kernel node) into one hipGraph launch: buckets of at least one
 * wavefront per SIMD one after another, smaller ones in parallel branches. */
typedef struct {
  const uint32_t* d_cand; /* limb-interleaved candidates: limb j of column p at d_cand[j*ld + p] */
  uint32_t ld;            /* leading dimension, >= n */
  uint32_t limbs;         /* rows stored in d_cand */
  uint32_t n;             /* candidates (columns 0..n-1) */
  uint32_t max_bits;      /* bound on every candidate's bit length (selects the kernel) */
  uint32_t* d_verdict;    /* n verdict words (indexed by column) */
  uint32_t* d_residues;   /* limbs rows of res_ld words, or NULL */
  uint32_t res_ld;
} pgx_bucket_t;

typedef struct {
  uint32_t blocks_per_cu;    /* resident workgroups per compute unit; 0 with one lane per
                                candidate: no persistence, one work item per thread over a
                                grid that covers the bucket (the next two fields unused) */
  uint32_t waves_per_block;  /* wavefronts per workgroup, 1..4 */
  uint32_t chunk;            /* candidates a wavefront takes per counter bump, multiple of 64 */
  uint32_t lanes_per_number; /* lanes that share one candidate: 1 (persistent workgroups,
                                the fields above apply), 2, 4, 8, 16 or 32 (a lane group
                                per candidate, one work item per group: for launches with
                                few candidates; the reference spreads a candidate over 4 or
                                8 threads, gpu_testing.cu:51) */
} pgx_tuning_t;

typedef struct pgx_plan pgx_plan_t;

/* One persistent launch over bucket b (the counter is reset on `stream` first).
 * tuning NULL = the runner's autotuned shape for this kernel, or the default. */
int pgx_test_persistent(pgx_runner_t* h, const pgx_bucket_t* b, const pgx_tuning_t* tuning,
                        uint32_t flags, void* stream);

/* Persistent launch over a device-resident work list: work item i tests column
 * d_index[i] of the bucket, for i < *d_count (d_count read on the device, at
 * most b->n).  Used for the extra Miller-Rabin bases on compacted survivors. */
int pgx_test_indexed(pgx_runner_t* h, const pgx_bucket_t* b, const uint32_t* d_index,
                     const uint32_t* d_count, const pgx_tuning_t* tuning, uint32_t flags,
                     void* stream);

/* Stream compaction on the device: appends every column i < n with
 * d_verdict[i] != PGX_VERDICT_INVALID and (d_verdict[i] & mask) != 0 to
 * d_index and its count to *d_count (reset first).  Order within the list is
 * unspecified. */
int pgx_compact(pgx_runner_t* h, const uint32_t* d_verdict, uint32_t n, uint32_t mask,
                uint32_t* d_index, uint32_t* d_count, void* stream);

/* Autotuner: times persistent launch shapes (blocks_per_cu x waves_per_block x
 * chunk), the plain grid and 2/4/8/16/32-lane groups on bucket b with `flags`, keeps
 * the fastest for b's kernel in the
 * runner (used when tuning is NULL) and returns it in *best with the measured
 * candidates per second in *per_sec (either may be NULL). */
int pgx_autotune(pgx_runner_t* h, const pgx_bucket_t* b, uint32_t flags, pgx_tuning_t* best,
                 double* per_sec);

/* Capture every bucket into one hipGraph (tuning[i] for bucket i, or NULL for
 * the runner's tuned shapes).  The buffers the buckets point to must outlive
 * the plan; each pgx_plan_launch re-tests them with their current contents. */
int pgx_plan_create(pgx_plan_t** out, pgx_runner_t* h, const pgx_bucket_t* buckets,
                    uint32_t nbuckets, const pgx_tuning_t* tuning, uint32_t flags);
int pgx_plan_launch(pgx_plan_t* p, void* stream);
int pgx_plan_destroy(pgx_plan_t* p);

/* A stream of the runner's device whose kernels run only on compute units
 * [cu_first, cu_first + cu_count) (a CU mask; the stream gets a hardware
 * queue of its own), for callers that split the GPU between concurrent
 * stages: e.g. batch k+1's base-2 Fermat launch on most compute units and
 * batch k's Miller-Rabin survivors on the rest, so that the latency-bound
 * stage neither waits behind the big launch nor stretches it.  No
 * counterpart in the reference (one stage per GPU thread there).  A process
 * has few hardware queues: creation fails with PGX_E_HIP when none is left.
 * The caller destroys the stream with pgx_stream_destroy once its work is
 * done. */
int pgx_stream_create_cus(pgx_runner_t* h, uint32_t cu_first, uint32_t cu_count, void** stream);
int pgx_stream_destroy(pgx_runner_t* h, void* stream);

/* ---- candidate construction: N = m * K + X with K = P# / d ----------------
 * Host-side builder for synthetic sieved candidates, the role of fill_batch
 * (gpu_testing.cu:136-195) and init_K (gap_common.cpp:60-65).
 *
 * pgx_gen_sieved: for each of the n_m multipliers m[k] (already chosen by the
 * caller), sieves X in [1, x_max] with every prime below sieve_limit and keeps
 * the first per_m survivors in ascending X (N has no prime factor <
 * sieve_limit).  Writes them in m order: out_limbs (candidate-major, `limbs`
 * words each), out_m (index k into m[]), out_x.  Returns the number written
 * (<= n_m * per_m) through *written.  Fails with PGX_E_INPUT if d does not
 * divide P# or a candidate does not fit in `limbs` words. */
int pgx_gen_sieved(uint32_t P, uint32_t d, const uint64_t* m, uint32_t n_m, uint32_t x_max,
                   uint32_t per_m, uint32_t sieve_limit, uint32_t limbs, uint32_t* out_limbs,
                   uint32_t* out_m, uint32_t* out_x, size_t* written);

/* Same for either side of m*K: N = m*K + sign*X, sign = +1 (the next-prime
 * side) or -1 (the prev-prime side, the reference's mpz_prevprime direction,
 * overflow.cpp:158). */
int pgx_gen_sieved_side(uint32_t P, uint32_t d, int32_t sign, const uint64_t* m, uint32_t n_m,
                        uint32_t x_max, uint32_t per_m, uint32_t sieve_limit, uint32_t limbs,
                        uint32_t* out_limbs, uint32_t* out_m, uint32_t* out_x, size_t* written);

/* Draw n_m distinct multipliers uniformly from [m_lo, m_hi) coprime to d with
 * the generator's seeded xoshiro256** stream (rejection sampling, ascending
 * order on return). */
int pgx_gen_multipliers(uint64_t seed, uint64_t m_lo, uint64_t m_hi, uint32_t d, uint32_t n_m,
                        uint64_t* out_m);

/* Bit length of K = P# / d (0 if d does not divide P#). */
uint32_t pgx_k_bits(uint32_t P, uint32_t d);

/* ---- measurement --------------------------------------------------------
 * VALU issue-rate microbenchmark (no counterpart in the reference, which
 * publishes no roofline): lane-instructions per second of one instruction
 * kind on `device`, 8 independent chains per lane, `blocks` x 256 threads x
 * `iters` x 8 instructions.  op: 0 v_mad_u64_u32 (8 independent chains),
 * 1 v_mad_u64_u32 in one dependent chain (what one chain sustains), 2 v_add_u32,
 * 3 v_mul_lo_u32, 4 v_fma_f64.  Gives the denominator of the INT32
 * multiply-add roofline. */
int pgx_valu_rate(int device, int op, uint32_t blocks, uint32_t iters, double* per_sec);

/* ---- serving: next-prime / prev-prime searches, continuous batching ------
 * Per-m requests in the role of the reference's overflow path
 * (mpz_nextprime / mpz_prevprime, overflow.cpp:148, 158) and its found-prime
 * bookkeeping (gpu_testing.cu:104-133): every request carries the sieved
 * candidates of both sides of m*K in ascending X; the answer per side is the
 * first candidate passing the reference's verdict with base 2 (early exit).
 * Candidates stay in HBM pools, one per kernel width; requests are admitted
 * while the byte budget and pool room allow and join the next iteration of
 * their width class; classes run on their own streams and overlap.  No
 * counterpart exists in the reference (which has no serving path). */
typedef struct {
  uint64_t hbm_budget;   /* bytes of admitted candidate + verdict storage */
  uint32_t chunk;        /* candidates per side per iteration at 512 bits (scaled by width) */
  uint32_t max_requests; /* admitted requests per width class */
  uint32_t pool_columns; /* candidate columns per width class pool */
  uint32_t max_side;     /* most candidates one side may carry */
  uint32_t stage_columns;/* columns admitted per class per iteration (pinned staging) */
  uint32_t group_min_bits;  /* classes at least this wide (kernel capacity, bits) test ... */
  uint32_t group_max_items; /* ... iterations of at most this many candidates with ... */
  uint32_t group_lanes;     /* ... 2, 4, 8, 16 or 32 lanes per candidate (lower latency for the wide
                               classes on a lightly loaded GPU); group_max_items 0 = never */
  uint32_t wide_cu_percent; /* 1..99: the classes at least group_min_bits wide run on this
                               share of the CUs and the narrower ones on the rest, so a
                               narrow class's wavefronts never share a SIMD with a wide
                               class's; 0 = every class may use every CU */
  uint32_t slots_per_class; /* iterations of one class in flight at once (1..4; 0 = 1): with
                               2, requests arriving during an iteration start at once in
                               the other slot instead of waiting for it to finish */
} pgx_serve_config_t;

typedef struct {
  uint64_t id;
  uint64_t arrival_ns;   /* CLOCK_MONOTONIC ns of arrival (0 = at submit) */
  uint32_t limbs;        /* 32-bit words per candidate in next/prev */
  uint32_t max_bits;     /* bound on every candidate's bit length */
  uint32_t n_next, n_prev;
  const uint32_t* next;  /* n_next candidates m*K + X, ascending X, candidate-major */
  const uint32_t* prev;  /* n_prev candidates m*K - X, ascending X */
} pgx_serve_request_t;

typedef struct {
  uint64_t id;
  int32_t next_first;    /* index into next of the first PRP, -1 if none */
  int32_t prev_first;
  uint32_t tests;        /* candidates tested for this request */
  uint64_t latency_ns;   /* arrival to completion */
} pgx_serve_result_t;

typedef struct {
  uint64_t bytes_in_use;
  uint64_t tests;
  uint32_t active, waiting, in_flight, finished_unread;
  /* device time of the iterations that tested candidates (HIP events on the
   * iteration's stream), and the device's idle time between one iteration's
   * end and the next one's start on the same stream: the host's scheduling
   * turnaround (poll, apply, admit, launch) plus any time with no work */
  uint64_t iterations, gaps;
  double iter_ms, gap_ms;
} pgx_serve_stats_t;

typedef struct pgx_server pgx_server_t;

int pgx_serve_create(pgx_server_t** out, int device, const pgx_serve_config_t* cfg);
/* Queues a request (its candidates are copied; the caller's arrays may be
 * reused at once).  PGX_E_CAPACITY if it can never fit the budget. */
int pgx_serve_submit(pgx_server_t* s, const pgx_serve_request_t* req);
/* Non-blocking: advances every width class whose iteration finished (retire,
 * admit, launch) and returns up to max_out completed requests. */
int pgx_serve_poll(pgx_server_t* s, pgx_serve_result_t* out, uint32_t max_out, uint32_t* n_out);
int pgx_serve_stats(const pgx_server_t* s, pgx_serve_stats_t* st);
int pgx_serve_destroy(pgx_server_t* s);
const char* pgx_serve_last_error(const pgx_server_t* s);

#ifdef __cplusplus
}
#endif

#endif /* PGX_H */
