"""GPU parity: the gfx950 kernels, called through the C ABI, against the golden
vectors and the GMP oracle on the same inputs.  Bit-exact: every verdict word
and every residue 2^(N-1) mod N."""
import sys

import numpy as np
import pytest

import oracle as O
import pgx
from conftest import limbs_needed

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


def expect_counts(recs, key, bases):
    out = []
    for r in recs:
        c = 0
        for a in bases:
            if not r[key][str(a)]:
                break
            c += 1
        out.append(c)
    return out


def check_records(recs, limbs, label):
    ns = [int(r["n"], 16) for r in recs]
    L = pgx.to_limbs(ns, limbs)
    say(f"{label}: {len(ns)} records, {limbs} limbs, fermat+residues")
    with pgx.Runner(len(ns), rounds=4, limbs=limbs) as run:
        v, res = run.run(L, pgx.MODE_FERMAT, residues=True)
        assert pgx.from_limbs(res) == [int(r["res2"], 16) for r in recs]
        assert [bool(x & 1) for x in v] == [r["fermat2"] for r in recs]
        assert [bool(x & 2) for x in v] == [r["sprp"]["2"] for r in recs]
        assert [bool(x & 4) for x in v] == [r["ref"]["2"] for r in recs]
        say(f"{label}: reference + strong modes, bases 2,3,5,7")
        ref = run.run(L, pgx.MODE_REFERENCE)
        assert list(ref) == expect_counts(recs, "ref", (2, 3, 5, 7))
        st = run.run(L, pgx.MODE_STRONG)
        assert list(st) == expect_counts(recs, "sprp", (2, 3, 5, 7))


def test_kat_vectors(vectors):
    recs = vectors["kat"]
    check_records(recs, max(limbs_needed(int(r["n"], 16)) for r in recs), "kat")


def test_every_kernel_width(vectors):
    # one runner per stored width so each kernel limb count S is selected
    by_limbs = {}
    for r in vectors["widths"]:
        by_limbs.setdefault(limbs_needed(int(r["n"], 16)), []).append(r)
    for limbs, recs in sorted(by_limbs.items()):
        check_records(recs, limbs, f"widths[{limbs}]")


def test_sieved_fixture_sets(vectors):
    for name, s in vectors["sieved"].items():
        check_records(s["cands"], s["limbs"], name)


def test_tiny_brute_force(vectors):
    t = vectors["tiny_res2"]
    ns = list(range(t["first"], t["first"] + t["step"] * len(t["res"]), t["step"]))
    say(f"tiny: {len(ns)} odd N < 4096")
    with pgx.Runner(len(ns), limbs=1) as run:
        _, res = run.run(pgx.to_limbs(ns, 1), pgx.MODE_FERMAT, residues=True)
    assert [int(x) for x in res[:, 0]] == [int(h, 16) for h in t["res"]]


def test_tiny_n_in_wide_kernels_vs_oracle():
    """Odd N < 4096 batched with one wide N, so that the kernel of the wide
    one tests them all: R mod N then takes up to ~2,000 doublings (16 per
    small-multiplier step) from 2^(L-1)."""
    import random
    rng = random.Random(5)
    for bits in (987, 2019):
        wide = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        ns = list(range(3, 4096, 2)) + [wide]
        words = (bits + 31) // 32
        L = pgx.to_limbs(ns, words)
        say(f"tiny N beside a {bits}-bit N: {len(ns)} candidates, {words} words")
        ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
        with pgx.Runner(len(ns), limbs=words, bases=[2, 3, 5]) as run:
            v, res = run.run(L, pgx.MODE_FERMAT, residues=True)
            assert np.array_equal(v, ov) and np.array_equal(res, orr), bits
            assert np.array_equal(run.run(L, pgx.MODE_STRONG),
                                  O.verdicts(L, O.MODE_STRONG, (2, 3, 5))), bits


def test_all_odd_below_2_16_vs_oracle():
    ns = list(range(3, 1 << 16, 2))
    L = pgx.to_limbs(ns, 1)
    say(f"odd N < 2^16: {len(ns)} candidates, all modes")
    ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
    with pgx.Runner(len(ns), rounds=3, limbs=1) as run:
        v, res = run.run(L, pgx.MODE_FERMAT, residues=True)
        assert np.array_equal(v, ov) and np.array_equal(res, orr)
        for mode in (pgx.MODE_REFERENCE, pgx.MODE_STRONG):
            assert np.array_equal(run.run(L, mode), O.verdicts(L, mode, (2, 3, 5)))


def test_headline_2e20_all_vs_gmp():
    """The headline workload whole: 2^20 sieved candidates (P=337, d=2310, 64
    multipliers, X <= 450000, 16-limb buffers, the benchmark's seed) in 16
    plain launches of 65,536 on HBM-resident buffers; every verdict and
    residue against GMP (the oracle on all host cores, chunk by chunk)."""
    import os
    from concurrent.futures import ThreadPoolExecutor

    import torch
    m = pgx.gen_multipliers(0x5EED0002, 10**7, 10**7 + 10**4, 2310, 64)
    L, _, _ = pgx.gen_sieved(337, 2310, m, 450000, 16384, 16)
    n = len(L)
    assert n == 1 << 20
    say(f"headline: {n} candidates, 16 launches of 65536")
    b = pgx.DeviceBatch(L)
    with pgx.Runner(1, limbs=16) as run:
        for c0 in range(0, n, 65536):
            run.test_device(b.cand.data_ptr() + 4 * c0, b.ld, b.limbs, 65536, b.max_bits,
                            pgx.MODE_FERMAT, b.verdict.data_ptr() + 4 * c0,
                            b.res.data_ptr() + 4 * c0, b.ld)
        torch.cuda.synchronize()
    v, res = b.verdicts(), b.residues()
    chunks = [(c0, min(c0 + 32768, n)) for c0 in range(0, n, 32768)]
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    say(f"oracle on {threads} threads")
    with ThreadPoolExecutor(threads) as ex:
        out = list(ex.map(lambda c: O.verdicts(L[c[0]:c[1]], O.MODE_FERMAT, residues=True),
                          chunks))
    ov = np.concatenate([o[0] for o in out])
    orr = np.concatenate([o[1] for o in out])
    assert np.array_equal(v, ov)
    assert np.array_equal(res, orr)


@pytest.mark.parametrize("P,d,limbs,per_m", [(337, 2310, 16, 1024), (211, 210, 16, 256),
                                             (503, 210, 24, 256), (701, 210, 32, 256),
                                             (907, 30030, 64, 128)])
def test_sieved_candidates_vs_gmp(P, d, limbs, per_m):
    """Configs' generators at reduced counts: residues and verdicts vs mpz_powm."""
    m = pgx.gen_multipliers(0x5EED0000 + P, 10**7, 10**7 + 10**4, d, 16)
    L, _, _ = pgx.gen_sieved(P, d, m, 450000, per_m, limbs)
    say(f"sieved P={P} d={d}: {len(L)} candidates in {limbs} limbs")
    ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
    with pgx.Runner(len(L), rounds=3, limbs=limbs) as run:
        v, res = run.run(L, pgx.MODE_FERMAT, residues=True)
        assert np.array_equal(res, orr)
        assert np.array_equal(v, ov)
        assert np.array_equal(run.run(L, pgx.MODE_REFERENCE),
                              O.verdicts(L, O.MODE_REFERENCE, (2, 3, 5)))
    # the reference's verdict equals the strong test's on every input
    assert np.array_equal((v >> 1) & 1, (v >> 2) & 1)
    assert (v & 1).sum() > 0


def test_config1_16384_candidates_vs_gmp():
    """P=337, d=2310, 512-bit buffers, 16384 sieved candidates (X widened)."""
    m = pgx.gen_multipliers(0x5EED0001, 10**7, 10**7 + 10**4, 2310, 16)
    L, _, _ = pgx.gen_sieved(337, 2310, m, 450000, 1024, 16)
    assert len(L) == 16384
    say("config-1 set: 16384 candidates")
    ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
    with pgx.Runner(16384, limbs=16) as run:
        v, res = run.run(L, pgx.MODE_FERMAT, residues=True)
    assert np.array_equal(v, ov)
    assert np.array_equal(res, orr)


def test_run_test_reference_semantics():
    """run_test(active, tests, results): only the first `active` entries are
    tested and written; results = (passed == rounds) (miller_rabin.h:364-408)."""
    primes = [2**61 - 1, 2**89 - 1, 2**107 - 1, 1000003]
    comps = [3215031751, 2047, 561, 2**61 + 1]
    tests = primes + comps
    with pgx.Runner(16, rounds=1, limbs=4) as run:
        res = [-1] * len(tests)
        run.run_test(6, tests, res)
        # 3215031751 and 2047 are base-2 strong pseudoprimes
        assert res == [1, 1, 1, 1, 1, 1, -1, -1]
    with pgx.Runner(16, rounds=4, limbs=4) as run:
        res = [-1] * len(tests)
        run.run_test(len(tests), tests, res)
        # 3215031751 passes bases 2,3,5,7 (it is a strong pseudoprime to all four)
        assert res == [1, 1, 1, 1, 1, 0, 0, 0]
        run.run_test(0, tests, res)  # nothing tested
        run.run_test(3, [], res)     # empty batch returns early


def test_submit_wait_and_errors():
    with pgx.Runner(8, limbs=2) as run:
        L = pgx.to_limbs([7, 9, 11], 2)
        L2 = pgx.to_limbs([13, 15], 2)
        run.submit(L, pgx.MODE_FERMAT, residues=True)
        run.submit(L2, pgx.MODE_FERMAT)  # a second batch in flight
        with pytest.raises(pgx.PgxError) as e:
            run.submit(L, pgx.MODE_FERMAT)
        assert e.value.code == pgx.PGX_E_STATE
        v, r = run.wait()  # oldest first
        assert list(v & 1) == [1, 0, 1] and pgx.from_limbs(r) == [1, 4, 1]
        assert list(run.wait() & 1) == [1, 0]
        with pytest.raises(pgx.PgxError) as e:
            run.wait()
        assert e.value.code == pgx.PGX_E_STATE
        for bad in ([8], [1], [0]):
            with pytest.raises(pgx.PgxError) as e:
                run.run(pgx.to_limbs(bad, 2), pgx.MODE_FERMAT)
            assert e.value.code == pgx.PGX_E_INPUT
        with pytest.raises(pgx.PgxError) as e:
            run.run(pgx.to_limbs([7] * 9, 2), pgx.MODE_FERMAT)
        assert e.value.code == pgx.PGX_E_CAPACITY
        # the runner still works after rejected calls
        assert list(run.run(pgx.to_limbs([13], 2), pgx.MODE_FERMAT) & 1) == [1]


def test_rejection_past_the_first_staging_piece():
    # the host checks and stages candidates about 1 MiB at a time, copying
    # each piece to the device before the next is checked: a bad candidate in
    # a later piece is still reported by index, and the runner stays usable
    n = (1 << 20) // 8 + 5
    with pgx.Runner(n, limbs=2) as run:
        L = pgx.to_limbs([7] * n, 2)
        L[n - 2, 0] = 8
        with pytest.raises(pgx.PgxError) as e:
            run.run(L, pgx.MODE_FERMAT)
        assert e.value.code == pgx.PGX_E_INPUT and f"candidate {n - 2} " in str(e.value)
        v = run.run(pgx.to_limbs([7] * (n - 1) + [9], 2), pgx.MODE_FERMAT)
        assert (v[:-1] & 1).all() and not (v[-1] & 1)


def test_device_resident_interleaved_layout():
    """pgx_test_device on torch-owned HBM buffers, limb-interleaved, ld > n."""
    import torch
    m = pgx.gen_multipliers(0x5EED0002, 10**7, 10**7 + 10**4, 2310, 4)
    L, _, _ = pgx.gen_sieved(337, 2310, m, 450000, 1000, 16)
    n, limbs = L.shape
    ld = n + 64
    say(f"device-resident: n={n} ld={ld}")
    host = np.zeros((limbs, ld), dtype=np.uint32)
    host[:, :n] = L.T
    dev = torch.from_numpy(host.view(np.int32)).cuda()
    verdict = torch.zeros(n, dtype=torch.int32, device="cuda")
    res = torch.zeros((limbs, ld), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    with pgx.Runner(1, limbs=limbs) as run:
        run.test_device(dev.data_ptr(), ld, limbs, n, 462, pgx.MODE_FERMAT, verdict.data_ptr(),
                        res.data_ptr(), ld, stream)
        torch.cuda.synchronize()
    v = verdict.cpu().numpy().view(np.uint32)
    r = res.cpu().numpy().view(np.uint32)[:, :n].T
    ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
    assert np.array_equal(v, ov) and np.array_equal(r, orr)


def test_host_rows_wider_than_one_tile_and_ragged_batches():
    """pgx_run moves candidate-major rows through a device transpose in tiles of
    64 candidates x 64 words: buffers of 67 words (a second, partial word
    tile), batch sizes that end mid-tile, and batches smaller than the runner,
    one after another on the same runner."""
    rng = np.random.default_rng(11)
    ns = [int.from_bytes(rng.bytes(250), "little") | (1 << 1990) | 1 for _ in range(200)]
    ns += [3, 5, 2047, 2**61 - 1]
    L = pgx.to_limbs(ns, 67)
    ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
    with pgx.Runner(len(ns), limbs=67) as run:
        for k in (len(ns), 1, 63, 65, 131):
            say(f"rows: {k} candidates x 67 words")
            v, r = run.run(L[-k:], pgx.MODE_FERMAT, residues=True)
            assert np.array_equal(v, ov[-k:]) and np.array_equal(r, orr[-k:]), k
            assert np.array_equal(run.run(L[:k], pgx.MODE_FERMAT), ov[:k]), k


def test_three_runners_driven_from_three_threads():
    """The reference's deployment: three runner instances, each owned by its
    own host thread and called synchronously at the same time
    (run_gpu_thread, gpu_testing.cu:203-354). Every call returns GMP's
    verdicts and residues for its own batch."""
    import threading
    m = pgx.gen_multipliers(0x5EED0007, 10**7, 10**7 + 10**4, 2310, 8)
    L, _, _ = pgx.gen_sieved(337, 2310, m, 450000, 768, 16)
    ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
    chunks = np.array_split(np.arange(len(L)), 9)
    out, errors = {}, []

    def worker(t):
        try:
            with pgx.Runner(max(len(c) for c in chunks), limbs=16) as run:
                for rep in range(2):
                    for k in range(t, len(chunks), 3):
                        out[(k, rep)] = run.run(L[chunks[k]], pgx.MODE_FERMAT, residues=True)
        except Exception as e:  # reported in the main thread
            errors.append(e)

    say(f"three threads x three runners, {len(L)} candidates, 2 passes")
    threads = [threading.Thread(target=worker, args=(t,)) for t in range(3)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not errors, errors
    assert len(out) == 2 * len(chunks)
    for (k, rep), (v, r) in out.items():
        idx = chunks[k]
        assert np.array_equal(v, ov[idx]) and np.array_equal(r, orr[idx]), (k, rep)


@pytest.mark.parametrize("S", [5, 17, 31, 46, 74])
def test_bit_lengths_below_kernel_top(S):
    """Random odd N of every bit length from 60 bits below the widest the
    kernel of S limbs holds up to it, in one batch (the kernel is picked by the
    widest, so R/N runs from 2^4 to 2^64 across the batch); residues and
    verdicts against GMP."""
    import random
    rb = 29 if S <= 31 else 28
    top = rb * S - 4
    rng = random.Random(S)
    vals = [rng.getrandbits(b) | (1 << (b - 1)) | 1
            for b in range(top - 60, top + 1) for _ in range(8)]
    words = (top + 31) // 32
    L = pgx.to_limbs(vals, words)
    say(f"S={S}: {len(vals)} candidates of {top - 60}..{top} bits")
    ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
    with pgx.Runner(len(vals), limbs=words, bases=[2, 3]) as run:
        v, r = run.run(L, pgx.MODE_FERMAT, residues=True)
        assert np.array_equal(v, ov) and np.array_equal(r, orr)
        assert np.array_equal(run.run(L, pgx.MODE_STRONG), O.verdicts(L, O.MODE_STRONG, (2, 3)))


def test_two_batches_in_flight_several_bases():
    """Two batches in flight on one runner (slot 0 and the second slot, each
    with its own stream), seven bases: the first base of one runs beside the
    later bases of the other; verdicts equal to one batch at a time, in
    submit order, through pgx_submit and pgx_submit_mk alike."""
    m = pgx.gen_multipliers(0x5EED0C01, 10**7, 10**7 + 10**4, 2310, 8)
    cands, _, _ = pgx.gen_sieved(337, 2310, m, 450000, 2048, 16)
    a, b = cands[:9000], cands[9000:16001]
    K = pgx.primorial_over(337, 2310)
    mm = np.asarray(m, dtype=np.uint64).repeat(600)
    with pgx.Runner(1 << 14, rounds=7, limbs=16) as run:
        want_a = run.run(a, pgx.MODE_REFERENCE)
        want_b = run.run(b, pgx.MODE_REFERENCE)
        want_mk = run.run_mk(K, mm, 4, 1, pgx.MODE_REFERENCE)
        for _ in range(3):
            run.submit(a, pgx.MODE_REFERENCE)
            run.submit(b, pgx.MODE_REFERENCE)
            assert np.array_equal(run.wait(), want_a)
            run.submit_mk(K, mm, 4, 1, pgx.MODE_REFERENCE)
            assert np.array_equal(run.wait(), want_b)
            assert np.array_equal(run.wait(), want_mk)
    assert np.array_equal(want_a, O.verdicts(a, O.MODE_REFERENCE, bases=pgx.first_primes(7)))
    assert set(np.unique(want_a)) >= {0, 7}
