"""GPU parity of the m-interval search (sieve at X, tests list, PRP tests,
found bitmap, pruning of the active list) against the CPU oracle's
restatement of the reference's sieve and GPU threads (oracle/msieve.c), in the
corrected arithmetic and as the reference writes it."""
import sys

import numpy as np
import pytest

import oracle as O
import pgx

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-search] {msg}", flush=True)
    sys.stdout.flush()


def kwords(K):
    return pgx.to_limbs([K], max(1, (K.bit_length() + 31) // 32))[0]


def run_case(P, d, m_start, m_inc, max_prime, as_written, bases, n_x, limbs=16):
    K = pgx.primorial_over(P, d)
    kw = kwords(K)
    half = m_inc // 2
    xs = [int(x) for x in O.ms_coprime_x(P, d)[:n_x]]
    active = O.ms_active(P, d, m_start, m_inc)
    found = np.zeros(O.bitmap_words(m_inc), dtype=np.uint32)
    with pgx.Runner(1 << 16, limbs=limbs, bases=bases) as run, \
            pgx.Search(run, K, P, d, m_start, m_inc, max_prime, as_written) as s:
        assert np.array_equal(s.active(), active)
        for X in xs:
            say(f"P={P} d={d} m_start={m_start} m_inc={m_inc} max_prime={max_prime} "
                f"as_written={as_written} bases={bases} X={X}")
            comp = O.ms_sieve(kw, P, d, m_start, m_inc, max_prime, X, as_written)
            n_tests = s.sieve(X)
            got_comp = s.bitmap(pgx.SEARCH_COMPOSITES)
            assert np.array_equal(got_comp, comp), f"composites differ at X={X}"
            tests, fmi, active = O.ms_step(kw, m_start, X, comp, active, found, bases)
            assert n_tests == len(tests)
            assert np.array_equal(s.tests(), tests)
            n_found = s.test()
            assert n_found == len(fmi)
            assert np.array_equal(s.active(), active)
            assert np.array_equal(s.bitmap(pgx.SEARCH_FOUND), found)
        mi, x = s.found()
    # every found pair, ascending in m_i, at the X where the oracle found it
    want = {}
    a = O.ms_active(P, d, m_start, m_inc)
    f2 = np.zeros(O.bitmap_words(m_inc), dtype=np.uint32)
    for X in xs:
        comp = O.ms_sieve(kw, P, d, m_start, m_inc, max_prime, X, as_written)
        _, fm, a = O.ms_step(kw, m_start, X, comp, a, f2, bases)
        for m in fm:
            want[int(m)] = X
    assert [int(v) for v in mi] == sorted(want)
    assert [int(v) for v in x] == [want[int(v)] for v in mi]
    return len(mi), half


@pytest.mark.parametrize("as_written", [False, True])
def test_readme_interval(as_written):
    """The reference README's run: P=337 d=2310 m_start=1e7 m_inc=2e5, primes to 1e6."""
    nf, half = run_case(337, 2310, 10**7, 200_000, 10**6, as_written, [2], 6)
    assert nf > 0


def test_first_branch_and_wrap_as_written():
    """m_start above X * max_prime (the halved branch) with a 32-bit wrap of
    X * neg_inv_K (max_prime 3e6), two bases."""
    run_case(211, 210, 10**13, 60_000, 3_000_000, True, [2, 3], 4)
    run_case(211, 210, 10**13, 60_000, 3_000_000, False, [2, 3], 2)


def test_odd_tail_and_tiny_interval():
    """half not a multiple of 32, a single tile, d with a prime above 64."""
    run_case(907, 2190, 2 * 10**6, 2 * 1001 + 38, 200_000, False, [2], 3, limbs=48)
    run_case(37, 6, 10**6, 6_000, 200_003, True, [2], 4, limbs=4)


def test_errors():
    K = pgx.primorial_over(337, 2310)
    with pgx.Runner(1024, limbs=16) as run:
        with pytest.raises(pgx.PgxError):
            pgx.Search(run, K, 337, 2310, 10**7 + 1, 200_000, 10**6)  # odd m_start
        with pytest.raises(pgx.PgxError):
            pgx.Search(run, K, 337, 2310, 10**7, 200_000, 10)  # max_prime below 1000
        with pgx.Runner(1024, limbs=8) as small:
            with pytest.raises(pgx.PgxError) as e:
                pgx.Search(small, K, 337, 2310, 10**7, 200_000, 10**6)
            assert e.value.code == pgx.PGX_E_INPUT
        with pgx.Search(run, K, 337, 2310, 10**7, 20_000, 10**6) as s:
            with pytest.raises(pgx.PgxError):
                s.test()  # nothing sieved
            with pytest.raises(pgx.PgxError):
                s.sieve(3)  # odd X

    say("batches smaller than the tests list")
    m_start, m_inc, mp = 10**7, 20_000, 10**6
    kw = kwords(K)
    with pgx.Runner(128, limbs=16) as run, pgx.Search(run, K, 337, 2310, m_start, m_inc, mp) as s:
        nt, nf = s.step(2)
        comp = O.ms_sieve(kw, 337, 2310, m_start, m_inc, mp, 2)
        found = np.zeros(O.bitmap_words(m_inc), dtype=np.uint32)
        tests, fmi, _ = O.ms_step(kw, m_start, 2, comp, O.ms_active(337, 2310, m_start, m_inc),
                                  found, (2,))
        assert nt == len(tests) > 3 * 128 and nf == len(fmi)
        assert np.array_equal(s.found()[0], np.sort(fmi))


def test_deep_sieve_device_primes():
    """Primes up to 2e8 generated on the device (11 M sieving primes): the
    composites bitmap still equals the oracle's, in both arithmetic modes."""
    run_case(337, 2310, 10**12, 20_000, 200_000_000, False, [2], 1)
    run_case(337, 2310, 10**7, 20_000, 200_000_000, True, [2], 1)
