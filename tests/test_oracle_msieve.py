"""Pins the oracle of the m-interval search step (oracle/msieve.c) against
brute force: gcd for the active multipliers, trial division for the
composites bitmap, CPython's pow for the verdicts, and an independent
restatement (via pow(K, -1, p)) of the reference's as-written start offsets for
the primes >= 100000 (gap_search_gpu.cpp:651-689)."""
import math

import numpy as np
import pytest

import oracle as O


def primorial_over(P, d):
    K = math.prod(q for q in range(2, P + 1) if all(q % r for r in range(2, math.isqrt(q) + 1)))
    assert K % d == 0
    return K // d


def k_words(K):
    n = max(1, (K.bit_length() + 31) // 32)
    return np.array([(K >> (32 * j)) & 0xFFFFFFFF for j in range(n)], dtype=np.uint32)


def odd_primes_below(n):
    s = np.ones(n, dtype=bool)
    s[:2] = False
    s[4::2] = False
    for i in range(3, math.isqrt(n) + 1, 2):
        if s[i]:
            s[i * i::2 * i] = False
    return [int(p) for p in np.nonzero(s)[0] if p > 2]


def bit(bits, t):
    return (int(bits[t >> 5]) >> (t & 31)) & 1


def brute_composites(K, P, d, m_start, m_inc, max_prime, X):
    """bit t set iff an odd prime q < max_prime not dividing K divides
    (m_start + 2t + 1) K + X (trial division, vectorised over m)."""
    half = m_inc // 2
    m = m_start + 2 * np.arange(half, dtype=np.int64) + 1
    comp = np.zeros(half, dtype=bool)
    for q in odd_primes_below(max_prime):
        if q <= P and d % q:
            continue  # q divides K
        comp |= ((m % q) * (K % q) + X) % q == 0
    return comp


def as_bits(words, half):
    return np.array([bit(words, t) for t in range(half)], dtype=bool)


CFG = dict(P=37, d=6, m_inc=6000, max_prime=200_003)


def test_active_is_gcd_filter():
    for P, d, m_start in ((37, 6, 10**6), (337, 2310, 10**7), (211, 210, 123456)):
        got = O.ms_active(P, d, m_start, 4000)
        want = [mi for mi in range(1, 4000, 2) if math.gcd(m_start + mi, d) == 1]
        assert got.tolist() == want


def test_coprime_x_brute():
    for P, d in ((37, 6), (337, 2310)):
        kp = [q for q in odd_primes_below(P + 1) + [2] if d % q]
        want = [x for x in range(2, 10 * P, 2) if all(x % q for q in kp)]
        assert O.ms_coprime_x(P, d).tolist() == want


@pytest.mark.parametrize("m_start", [10**6, 10**13 + 10**6])
def test_sieve_matches_trial_division(m_start):
    """Both branches of the large-prime start (X * max_prime below / above
    m_start) mark exactly the divisible multipliers in the corrected mode."""
    P, d, m_inc, mp = CFG["P"], CFG["d"], CFG["m_inc"], CFG["max_prime"]
    K = primorial_over(P, d)
    for X in (2, 4, 184, 368):
        got = as_bits(O.ms_sieve(k_words(K), P, d, m_start, m_inc, mp, X), m_inc // 2)
        want = brute_composites(K, P, d, m_start, m_inc, mp, X)
        assert np.array_equal(got, want), X


def quirk_marks(K, P, d, m_start, m_inc, max_prime, X):
    """Marks of the as-written reference: primes < 100000 and the d wheel are
    exact; each prime p >= 100000 starts at t0 from a 32-bit X * neg_inv_K
    and, when X * max_prime >= m_start, indexes the half bitmap with the
    un-halved full offset."""
    half = m_inc // 2
    marks = brute_composites(K, P, d, m_start, m_inc, min(max_prime, 100_000), X)
    for p in odd_primes_below(max_prime):
        if p < 100_000:
            continue
        neg_inv = p - pow(K % p, -1, p)
        xk = (X * neg_inv) & 0xFFFFFFFF
        if X * max_prime < m_start:
            f = (-(m_start - xk)) % p  # p - ((m_start - xk) mod p), or 0
        else:
            f = (xk - m_start) % p
        if f % 2 == 0:
            f += p
        t0 = f // 2 if X * max_prime < m_start else f
        marks[t0::p] = True
    return marks


@pytest.mark.parametrize("m_start,max_prime", [(10**6, 200_003), (10**13, 200_003),
                                               (10**13, 3_000_000)])
def test_as_written_quirks(m_start, max_prime):
    P, d, m_inc = CFG["P"], CFG["d"], CFG["m_inc"]
    K = primorial_over(P, d)
    for X in (2, 184, 368):
        aw = as_bits(O.ms_sieve(k_words(K), P, d, m_start, m_inc, max_prime, X, as_written=True),
                     m_inc // 2)
        assert np.array_equal(aw, quirk_marks(K, P, d, m_start, m_inc, max_prime, X)), X
        ok = as_bits(O.ms_sieve(k_words(K), P, d, m_start, m_inc, max_prime, X), m_inc // 2)
        if X * max_prime < m_start and X * max_prime < 2**32:  # nothing as written differs
            assert np.array_equal(aw, ok)


def test_as_written_differs_where_the_reference_does():
    """README's example shape (m_start 1e7, max_prime 1e6): the un-halved start
    changes the bitmap for X >= 10."""
    P, d, m_start, m_inc, mp = 337, 2310, 10**7, 20000, 10**6
    K = primorial_over(P, d)
    a = O.ms_sieve(k_words(K), P, d, m_start, m_inc, mp, 2, as_written=True)
    b = O.ms_sieve(k_words(K), P, d, m_start, m_inc, mp, 2)
    assert np.array_equal(a, b)  # X * max_prime < m_start: the halved branch
    a = O.ms_sieve(k_words(K), P, d, m_start, m_inc, mp, 16, as_written=True)
    b = O.ms_sieve(k_words(K), P, d, m_start, m_inc, mp, 16)
    assert not np.array_equal(a, b)


def py_ref_pass(N, a):
    return O.py_ref_verdict(N, a)


def test_step_tests_found_and_prune():
    P, d, m_start, m_inc, mp = 37, 6, 10**6, 2000, 20_000
    K = primorial_over(P, d)
    kw = k_words(K)
    active = O.ms_active(P, d, m_start, m_inc)
    found = np.zeros(O.bitmap_words(m_inc), dtype=np.uint32)
    seen_found = set()
    for X in O.ms_coprime_x(P, d)[:6]:
        X = int(X)
        comp = O.ms_sieve(kw, P, d, m_start, m_inc, mp, X)
        before = active.copy()
        tests, fmi, active = O.ms_step(kw, m_start, X, comp, active, found, bases=(2, 3))
        want_tests = [int(mi) for mi in before
                      if not bit(comp, int(mi) >> 1) and int(mi) not in seen_found]
        assert tests.tolist() == want_tests
        want_found = [mi for mi in want_tests
                      if all(py_ref_pass((m_start + mi) * K + X, a) for a in (2, 3))]
        assert fmi.tolist() == want_found
        seen_found |= set(want_found)
        assert active.tolist() == [int(mi) for mi in before if int(mi) not in seen_found]
        assert all(bit(found, mi >> 1) for mi in seen_found)
    assert seen_found  # the interval holds primes at these X
