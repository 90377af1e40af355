"""Pins the CPU oracle (GMP) against the golden vectors and against the
mathematics; no GPU involved."""
import numpy as np
import pytest

import oracle as O
from conftest import limbs_needed


def _limbs(values, S):
    out = np.zeros((len(values), S), dtype=np.uint32)
    for i, v in enumerate(values):
        out[i] = np.frombuffer(v.to_bytes(4 * S, "little"), dtype="<u4")
    return out


def _ints(arr):
    return [int.from_bytes(np.ascontiguousarray(r, dtype="<u4").tobytes(), "little") for r in arr]


def _check_records(recs):
    ns = [int(r["n"], 16) for r in recs]
    S = max(limbs_needed(n) for n in ns)
    L = _limbs(ns, S)
    f, res = O.verdicts(L, O.MODE_FERMAT, residues=True)
    for i, r in enumerate(recs):
        assert _ints(res[i:i + 1])[0] == int(r["res2"], 16), r["n"]
        assert bool(f[i] & 1) == r["fermat2"]
        assert bool(f[i] & 2) == r["sprp"]["2"]
        assert bool(f[i] & 4) == r["ref"]["2"]
    for mode, key in ((O.MODE_REFERENCE, "ref"), (O.MODE_STRONG, "sprp")):
        bases = [2, 3, 5, 7]
        cnt = O.verdicts(L, mode, bases)
        for i, r in enumerate(recs):
            want = 0
            for a in bases:
                if not r[key][str(a)]:
                    break
                want += 1
            assert cnt[i] == want, (r["n"], mode)


def test_oracle_kat(vectors):
    _check_records(vectors["kat"])


def test_oracle_widths(vectors):
    _check_records(vectors["widths"])


def test_oracle_sieved_records(vectors):
    for s in vectors["sieved"].values():
        _check_records(s["cands"])


def test_oracle_tiny_brute_force(vectors):
    t = vectors["tiny_res2"]
    ns = list(range(t["first"], t["first"] + t["step"] * len(t["res"]), t["step"]))
    L = _limbs(ns, 1)
    _, res = O.verdicts(L, O.MODE_FERMAT, residues=True)
    assert [int(x) for x in res[:, 0]] == [int(h, 16) for h in t["res"]]


def test_oracle_prime_records_pass_every_base(vectors):
    """Primes pass every test; probab_prime agrees with the fixture's sympy verdict."""
    recs = vectors["kat"] + vectors["widths"]
    ns = [int(r["n"], 16) for r in recs]
    S = max(limbs_needed(n) for n in ns)
    pp = O.probab_prime(_limbs(ns, S), 25)
    for r, p in zip(recs, pp):
        assert (p > 0) == r["prime"], r["n"]
        if r["prime"]:  # bases that are multiples of N (N = 3, 5, 7) give 0
            N = int(r["n"], 16)
            assert r["fermat2"] and all(v for a, v in r["sprp"].items() if int(a) % N)


def test_quirk_equals_sprp():
    """a^(N-1) == -1 (mod N) is impossible for odd N > 1 (every prime factor p
    would satisfy p == 1 mod 2^(v2(N-1)+1), hence so would N), so the
    reference's extra squaring never changes its verdict: ref == SPRP."""
    rng = np.random.default_rng(7)
    ns = [int(x) | 1 for x in rng.integers(3, 1 << 40, size=3000)] + list(range(3, 20001, 2))
    L = _limbs(ns, 2)
    for a in (2, 3, 5):
        ref = O.verdicts(L, O.MODE_REFERENCE, [a])
        st = O.verdicts(L, O.MODE_STRONG, [a])
        assert np.array_equal(ref, st)
    # and no odd N below 2^16 has 2^(N-1) == N-1
    assert not any(pow(2, n - 1, n) == n - 1 for n in range(3, 1 << 16, 2))


def test_python_restatement_matches_oracle():
    rng = np.random.default_rng(3)
    ns = [int(x) | 1 for x in rng.integers(3, 1 << 62, size=500)] + [2047, 3277, 341, 561]
    L = _limbs(ns, 2)
    f = O.verdicts(L, O.MODE_FERMAT)
    for n, v in zip(ns, f):
        assert bool(v & 1) == (O.py_residue(n) == 1)
        assert bool(v & 2) == O.py_sprp(n, 2)
        assert bool(v & 4) == O.py_ref_verdict(n, 2)


def test_prng_parts_match_published_known_answers():
    """The generator's two parts against the known answers published with
    their reference algorithms: splitmix64 from seed 0, and xoshiro256** from
    the state {1, 2, 3, 4} (the vectors other implementations test against);
    the seeded stream is xoshiro256** on four splitmix64 outputs."""
    sm = O.splitmix_stream(0, 4)
    assert [int(v) for v in sm] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                    0x06C45D188009454F, 0xF88BB8A8724C81EC]
    xo = O.xoshiro_from_state([1, 2, 3, 4], 10)
    assert [int(v) for v in xo] == [11520, 0, 1509978240, 1215971899390074240,
                                    1216172134540287360, 607988272756665600,
                                    16172922978634559625, 8476171486693032832,
                                    10595114339597558777, 2904607092377533576]
    for seed in (0, 0x5EED0001, 2**64 - 1):
        assert np.array_equal(O.xoshiro_stream(seed, 16),
                              O.xoshiro_from_state(O.splitmix_stream(seed, 4), 16))


def test_prng_matches_fixture(vectors):
    p = vectors["prng"]
    got = O.xoshiro_stream(p["seed"], len(p["first"]))
    assert [hex(int(x)) for x in got] == p["first"]


def test_generator_matches_fixture(vectors):
    for s in vectors["sieved"].values():
        m = O.gen_multipliers(s["seed"], 10**7, 10**7 + 10**4, s["d"], len(s["m"]))
        assert [int(x) for x in m] == s["m"]
        L, mi, x = O.gen_sieved(s["P"], s["d"], m, s["x_max"], s["per_m"], s["limbs"])
        assert _ints(L) == [int(c["n"], 16) for c in s["cands"]]
        assert list(x) == [c["x"] for c in s["cands"]]
        assert list(mi) == [c["m_index"] for c in s["cands"]]


def test_generator_sieve_by_trial_division():
    m = O.gen_multipliers(11, 10**7, 10**7 + 10**4, 2310, 1)
    L, _, x = O.gen_sieved(337, 2310, m, 3000, 10**9, 16, sieve_limit=20000)
    assert len(x) > 0
    assert not O.small_factor(L, 20000).any()  # survivors: no factor < limit
    # every dropped X has one
    K = 1
    for p in range(2, 338):
        if all(p % q for q in range(2, int(p**0.5) + 1)):
            K *= p
    K //= 2310
    kept = set(int(v) for v in x)
    dropped = [int(m[0]) * K + X for X in range(1, 3001) if X not in kept]
    D = _limbs(dropped, 16)
    assert O.small_factor(D, 20000).all()


def test_planted_residue_mistake_is_caught(vectors):
    """A corrupted residue (one flipped bit) must fail the comparison used by
    the parity tests."""
    rec = vectors["kat"][12]
    N = int(rec["n"], 16)
    S = limbs_needed(N)
    _, res = O.verdicts(_limbs([N], S), O.MODE_FERMAT, residues=True)
    bad = res.copy()
    bad[0, 0] ^= 1
    assert _ints(res)[0] == int(rec["res2"], 16)
    assert _ints(bad)[0] != int(rec["res2"], 16)


def test_cpu_baseline_timer_counts_fermat():
    m = O.gen_multipliers(5, 10**7, 10**7 + 10**4, 2310, 1)
    L, _, _ = O.gen_sieved(337, 2310, m, 20000, 200, 16)
    sec, cnt = O.time_powm2(L, threads=2)
    f = O.verdicts(L, O.MODE_FERMAT)
    assert sec > 0 and cnt == int((f & 1).sum())


def test_next_prime_offset_vs_sympy():
    sympy = pytest.importorskip("sympy")
    import pgx
    K = pgx.primorial_over(337, 2310)
    kw = pgx.to_limbs([K], (K.bit_length() + 31) // 32)[0]
    for m in (10000019, 10000079):
        c = m * K
        assert O.next_prime_offset(kw, m, 1, 30000) == sympy.nextprime(c) - c
        assert O.next_prime_offset(kw, m, -1, 30000) == c - sympy.prevprime(c)
    assert O.next_prime_offset(kw, 10000019, 1, 10) == 0  # beyond the limit


def test_build_mk_vs_python_bignums():
    """fill_batch's construction in the oracle (mpz_mul_ui, mpz_add_ui /
    mpz_sub_ui, mpz_export) against CPython integers, both signs, and the
    errors for values that do not fit or go negative."""
    import pgx
    K = pgx.primorial_over(337, 2310)
    kw = pgx.to_limbs([K], 14)[0]
    rng = np.random.default_rng(3)
    m = rng.integers(1, 1 << 40, size=500, dtype=np.uint64)
    for sign, x in ((1, 449999), (-1, 17)):
        out, sec = O.build_mk(kw, m, x, sign, 16)
        want = pgx.to_limbs([int(v) * K + sign * x for v in m], 16)
        assert np.array_equal(out, want) and sec >= 0
    with pytest.raises(ValueError):  # 439 + 40 bits do not fit 14 words
        O.build_mk(kw, m, 1, 1, 14)
    with pytest.raises(ValueError):  # 0*K - 1 < 0
        O.build_mk(kw, np.zeros(1, dtype=np.uint64), 1, -1, 16)
