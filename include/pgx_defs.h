/* pgx_defs.h — constants of the C ABI (include/pgx.h), in a header of their
 * own so that device code depends on these values only. */
#ifndef PGX_DEFS_H
#define PGX_DEFS_H

/* status codes */
#define PGX_OK 0
#define PGX_E_ARG (-1)      /* bad argument (null pointer, zero size, ...) */
#define PGX_E_HIP (-2)      /* HIP runtime failure (message names the call) */
#define PGX_E_CAPACITY (-3) /* n larger than the runner's max_n (miller_rabin.h:368-371) */
#define PGX_E_INPUT (-4)    /* candidate even, <= 1, or too wide (miller_rabin.h:46-49, 375-376) */
#define PGX_E_STATE (-5)    /* wait without submit, submit while busy */
#define PGX_E_NODEV (-6)    /* no usable gfx950 device */

/* test modes (flags & PGX_MODE_MASK) */
#define PGX_MODE_REFERENCE 0u /* reference verdict: passed = number of leading bases that pass
                                 the strong test as written in miller_rabin.h:216-248 (which
                                 also accepts a^(N-1) == -1 when ctz(N-1) >= 2) */
#define PGX_MODE_STRONG 1u    /* textbook strong test per base: passed = leading bases passed */
#define PGX_MODE_FERMAT 2u    /* base-2 Fermat only: verdict word = PGX_FLAG_* bits of base 2 */
#define PGX_MODE_MASK 3u
#define PGX_WANT_RESIDUE 4u   /* also return a^(N-1) mod N of the first base (usually 2) */

/* bits of the verdict word in PGX_MODE_FERMAT */
#define PGX_FLAG_FERMAT 1u /* 2^(N-1) == 1 (mod N) */
#define PGX_FLAG_SPRP 2u   /* strong probable prime to base 2 */
#define PGX_FLAG_QUIRK 4u  /* reference verdict for base 2 */

#define PGX_VERDICT_INVALID 0xFFFFFFFFu /* candidate rejected by the kernel */
#define PGX_MAX_BASES 16
#define PGX_MAX_BASE 255 /* largest Miller-Rabin base (the first 16 primes end at 53) */
#define PGX_MAX_BITS 2068 /* widest candidate the kernels take (74 radix-2^28 limbs) */
#define PGX_NOALLOC_MAX_N (1u << 20) /* most candidates per pgx_test_device_noalloc call */

#endif /* PGX_DEFS_H */
