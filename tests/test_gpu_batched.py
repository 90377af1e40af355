"""GPU parity of the batched path: persistent kernels, hipGraph plans over
mixed-size buckets, device compaction and the extra Miller-Rabin bases on
Fermat survivors.  Checked against the plain kernel and the GMP oracle."""
import sys

import numpy as np
import pytest

import oracle as O
import pgx

pytestmark = pytest.mark.gpu

# (P, d, buffer words) of the mixed-size configuration (d = 210)
MIXED = [(211, 210, 16), (337, 210, 16), (503, 210, 24), (701, 210, 32)]


def say(msg):
    print(f"[gpu-batched] {msg}", flush=True)
    sys.stdout.flush()


def gen(P, d, words, n_m, per_m, seed_off=0):
    m = pgx.gen_multipliers(0x5EED0100 + P + seed_off, 10**7, 10**7 + 10**4, d, n_m)
    L, _, _ = pgx.gen_sieved(P, d, m, 450000, per_m, words)
    return L


def test_persistent_equals_plain_and_oracle():
    import torch
    with pgx.Runner(1, limbs=16) as run:
        for P, d, words in MIXED:
            L = gen(P, d, words, 8, 256)
            say(f"persistent P={P}: {len(L)} candidates, {words} words")
            a = pgx.DeviceBatch(L)
            run.test_persistent(a.bucket(), pgx.MODE_FERMAT, residues=True)
            torch.cuda.synchronize()
            ov, ores = O.verdicts(L, O.MODE_FERMAT, residues=True)
            assert np.array_equal(a.verdicts(), ov)
            assert np.array_equal(a.residues(), ores)
            # a non-default shape, and the plain grid (blocks_per_cu 0), give
            # the same answer
            for t in (pgx.Tuning(2, 1, 256, 1), pgx.Tuning(0, 0, 0, 1)):
                b = pgx.DeviceBatch(L)
                run.test_persistent(b.bucket(), pgx.MODE_FERMAT, residues=True, tuning=t)
                torch.cuda.synchronize()
                assert np.array_equal(b.verdicts(), ov)
                assert np.array_equal(b.residues(), ores)


def test_plan_graph_replays_match_eager():
    """One hipGraph over four buckets; replayed three times with eager work and
    synchronizations in between, every replay equal to the eager result."""
    import torch
    batches, eager = [], []
    with pgx.Runner(1, limbs=16) as run:
        for P, d, words in MIXED:
            L = gen(P, d, words, 4, 256, seed_off=7)
            batches.append(pgx.DeviceBatch(L))
            e = pgx.DeviceBatch(L)
            run.test_persistent(e.bucket(), pgx.MODE_FERMAT, residues=True)
            eager.append(e)
        torch.cuda.synchronize()
        say(f"plan over {len(batches)} buckets")
        plan = pgx.Plan(run, [b.bucket() for b in batches], pgx.MODE_FERMAT, residues=True)
        for rep in range(3):
            for b in batches:  # clear outputs between replays
                b.verdict.zero_()
                b.res.zero_()
            torch.cuda.synchronize()
            plan.launch()
            torch.cuda.synchronize()
            for b, e in zip(batches, eager):
                assert np.array_equal(b.verdicts(), e.verdicts()), rep
                assert np.array_equal(b.residues(), e.residues()), rep
            # eager work between replays
            run.test_persistent(eager[0].bucket(), pgx.MODE_FERMAT, residues=True)
            torch.cuda.synchronize()
        plan.close()
    # and the eager result itself is the oracle's
    L0 = gen(*MIXED[0], 4, 256, seed_off=7)
    assert np.array_equal(eager[0].verdicts(), O.verdicts(L0, O.MODE_FERMAT))


def test_plan_with_a_gpu_filling_bucket_beside_small_ones():
    """A bucket of more than one wavefront per SIMD (81,920 candidates) runs on
    the plan's capture stream, two small buckets on branches beside it; every
    replay equals the eager result, and a sample of the large bucket the
    oracle's."""
    import torch
    specs = [(337, 2310, 16, 20, 4096), (211, 210, 16, 4, 256), (503, 210, 24, 4, 256)]
    data = [gen(P, d, w, n_m, per_m, seed_off=11) for P, d, w, n_m, per_m in specs]
    assert len(data[0]) == 81920
    batches = [pgx.DeviceBatch(L) for L in data]
    eager = [pgx.DeviceBatch(L) for L in data]
    with pgx.Runner(1, limbs=16) as run:
        for e in eager:
            run.test_persistent(e.bucket(), pgx.MODE_FERMAT, residues=True)
        torch.cuda.synchronize()
        say(f"plan over buckets of {[len(L) for L in data]}")
        plan = pgx.Plan(run, [b.bucket() for b in batches], pgx.MODE_FERMAT, residues=True)
        for rep in range(3):
            for b in batches:
                b.verdict.zero_()
                b.res.zero_()
            torch.cuda.synchronize()
            plan.launch()
            torch.cuda.synchronize()
            for b, e in zip(batches, eager):
                assert np.array_equal(b.verdicts(), e.verdicts()), rep
                assert np.array_equal(b.residues(), e.residues()), rep
            run.test_persistent(eager[1].bucket(), pgx.MODE_FERMAT, residues=True)
            torch.cuda.synchronize()
        plan.close()
    ov, ores = O.verdicts(data[0][:512], O.MODE_FERMAT, residues=True)
    assert np.array_equal(eager[0].verdicts()[:512], ov)
    assert np.array_equal(eager[0].residues()[:512], ores)


def test_fermat_then_mr_on_compacted_survivors():
    """Base-2 Fermat on every candidate, on-device compaction of the PRPs, then
    strong tests to bases 3 and 5 on the survivors only (P=907, d=30030)."""
    import torch
    L = gen(907, 30030, 64, 16, 128)
    n = len(L)
    say(f"fermat->compact->MR: {n} candidates of ~1260 bits")
    a = pgx.DeviceBatch(L)
    index = torch.zeros(n, dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int32, device="cuda")
    mr = pgx.DeviceBatch(L, residues=False)  # same candidates; MR verdicts land here
    mr.verdict.fill_(-7)                       # marker: untouched columns keep it
    # one stream orders the three steps (each runner's default stream is its own)
    s = torch.cuda.Stream().cuda_stream
    with pgx.Runner(1, limbs=64) as fermat, pgx.Runner(1, limbs=64, bases=[3, 5]) as extra:
        fermat.test_persistent(a.bucket(), pgx.MODE_FERMAT, residues=True, stream=s)
        fermat.compact(a.verdict.data_ptr(), n, pgx.FLAG_FERMAT, index.data_ptr(),
                       count.data_ptr(), stream=s)
        extra.test_indexed(mr.bucket(), index.data_ptr(), count.data_ptr(), pgx.MODE_STRONG,
                           stream=s)
        torch.cuda.synchronize()
    v = a.verdicts()
    ov = O.verdicts(L, O.MODE_FERMAT)
    assert np.array_equal(v, ov)
    surv = np.nonzero(ov & 1)[0]
    k = int(count.item())
    assert k == len(surv) and k > 0
    assert sorted(index.cpu().numpy()[:k].tolist()) == surv.tolist()
    got = mr.verdicts()
    want = O.verdicts(L[surv], O.MODE_STRONG, (3, 5))
    assert np.array_equal(got[surv], want)
    others = np.setdiff1d(np.arange(n), surv)
    assert np.all(got[others] == np.uint32(-7 & 0xFFFFFFFF))
    say("the same work list on the plain grid")
    mr2 = pgx.DeviceBatch(L, residues=False)
    mr2.verdict.fill_(-7)
    torch.cuda.synchronize()
    with pgx.Runner(1, limbs=64, bases=[3, 5]) as extra:
        extra.test_indexed(mr2.bucket(), index.data_ptr(), count.data_ptr(), pgx.MODE_STRONG,
                           tuning=pgx.Tuning(0, 0, 0, 1))
        torch.cuda.synchronize()
    assert np.array_equal(mr2.verdicts(), got)


def test_indexed_with_empty_list():
    import torch
    L = gen(337, 2310, 16, 1, 64)
    b = pgx.DeviceBatch(L, residues=False)
    b.verdict.fill_(5)
    index = torch.zeros(1, dtype=torch.int32, device="cuda")
    count = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pgx.Runner(1, limbs=16) as run:
        run.test_indexed(b.bucket(), index.data_ptr(), count.data_ptr(), pgx.MODE_FERMAT)
        torch.cuda.synchronize()
    assert np.all(b.verdicts() == 5)


# strong pseudoprimes to the first k primes for growing k (2047: base 2;
# 1373653: 2, 3; ... 3825123056546413051: 2..23), base-3 ones (121, 703,
# 1891) and Carmichael numbers: candidates that pass the first bases of a list
# and fail a later one
PSEUDOPRIMES = [2047, 1373653, 25326001, 3215031751, 2152302898747, 3474749660383,
                341550071728321, 3825123056546413051, 121, 703, 1891, 561, 1105, 41041]


def two_fills():
    """Two device fills: two wavefronts of 64 on each of the 4 SIMDs of every
    compute unit."""
    import torch
    return 2 * 4 * 64 * torch.cuda.get_device_properties(0).multi_processor_count


def one_pass(L, bases, mode):
    """Verdicts of a single launch over every base (pgx_test_device)."""
    import torch
    b = pgx.DeviceBatch(L, residues=False)
    with pgx.Runner(1, limbs=L.shape[1], bases=bases) as r:
        r.test_device(b.cand.data_ptr(), b.ld, b.limbs, b.n, b.max_bits, mode,
                      b.verdict.data_ptr())
        torch.cuda.synchronize()
    return b.verdicts()


@pytest.mark.parametrize("bases", [[2, 3, 5, 7, 11, 13, 17], [3, 2, 5], [2, 3]])
def test_runner_runs_extra_bases_on_survivors(bases):
    """pgx_submit with several bases tests the first base on every candidate
    and the others, one (survivor, base) pair per lane, on the device-compacted
    list of those that passed it.  The verdicts equal a single pass over all
    bases (pgx_test_device) on every column and the oracle's on the
    pseudoprimes and a sample, with the first base's residues; the list and
    the pass masks are rebuilt for the next batch, a smaller one, a batch with
    no survivors and candidates built on the device."""
    fills = two_fills()
    P = pgx.to_limbs(PSEUDOPRIMES, 16)
    sieved = gen(337, 2310, 16, 64, fills // 64 + 64)
    L = np.concatenate([P, sieved[:fills + 100], P])
    n = len(L)
    sample = np.r_[0:len(P) + 1500, len(P) + 1500:n - len(P):97, n - len(P):n]
    with pgx.Runner(n + 4096, limbs=16, bases=bases) as run:
        for mode in (pgx.MODE_STRONG, pgx.MODE_REFERENCE):
            say(f"bases {bases} mode {mode}: {n} candidates (two fills = {fills})")
            v, r = run.run(L, mode, residues=True)
            assert np.array_equal(v, one_pass(L, bases, mode))
            ov, ores = O.verdicts(L[sample], mode, tuple(bases), residues=True)
            assert np.array_equal(v[sample], ov) and np.array_equal(r[sample], ores)
            assert int((v >= 1).sum()) > int((v == len(bases)).sum())  # later bases fail some
        say(f"bases {bases}: next batch, a smaller one, no survivors")
        L2 = gen(337, 2310, 16, 64, fills // 64 + 64, seed_off=1)[:fills + 3000]
        assert np.array_equal(run.run(L2, pgx.MODE_STRONG), one_pass(L2, bases, pgx.MODE_STRONG))
        small = L2[:4096]
        assert np.array_equal(run.run(small, pgx.MODE_STRONG),
                              O.verdicts(small, O.MODE_STRONG, tuple(bases)))
        none = np.tile(pgx.to_limbs([9, 15, 21, 25, 27, 33], 16), (fills // 6 + 1, 1))
        assert np.all(run.run(none, pgx.MODE_STRONG) == 0)
        say(f"bases {bases}: candidates built on the device")
        K = pgx.primorial_over(337, 2310)
        m = np.asarray(pgx.gen_multipliers(0x5EED0300, 10**7, 10**7 + 10**6, 2310, fills + 50),
                       dtype=np.uint64)
        v = run.run_mk(K, m, 7920, 1, pgx.MODE_STRONG)
        Lm = pgx.to_limbs([int(mm) * K + 7920 for mm in m], 16)
        odd = (Lm[:, 0] & 1) == 1
        assert np.all(v[~odd] == pgx.VERDICT_INVALID)
        assert np.array_equal(v[odd], one_pass(Lm[odd], bases, pgx.MODE_STRONG))


def test_autotune_picks_a_valid_shape():
    import torch
    L = gen(337, 2310, 16, 16, 1024)
    b = pgx.DeviceBatch(L)
    with pgx.Runner(1, limbs=16) as run:
        t, rate = run.autotune(b.bucket(), pgx.MODE_FERMAT)
        say(f"autotune: {t.as_dict()} {rate / 1e6:.1f} M/s")
        assert rate > 0 and t.lanes_per_number == 1 and t.chunk % 64 == 0
        b.verdict.zero_()
        run.test_persistent(b.bucket(), pgx.MODE_FERMAT, residues=True)  # uses the tuned shape
        torch.cuda.synchronize()
    ov, ores = O.verdicts(L[:4096], O.MODE_FERMAT, residues=True)
    assert np.array_equal(b.verdicts()[:4096], ov)
    assert np.array_equal(b.residues()[:4096], ores)


def test_bad_tuning_and_bucket_are_rejected():
    L = gen(337, 2310, 16, 1, 64)
    b = pgx.DeviceBatch(L)
    with pgx.Runner(1, limbs=16) as run:
        for bad in (pgx.Tuning(1, 1, 100, 1), pgx.Tuning(1, 5, 64, 1), pgx.Tuning(1, 1, 64, 3),
                    pgx.Tuning(0, 4, 64, 2)):  # the plain grid is for one lane per candidate
            with pytest.raises(pgx.PgxError) as e:
                run.test_persistent(b.bucket(), pgx.MODE_FERMAT, tuning=bad)
            assert e.value.code == pgx.PGX_E_ARG
        wide = b.bucket()
        wide.max_bits = 5000
        with pytest.raises(pgx.PgxError) as e:
            run.test_persistent(wide, pgx.MODE_FERMAT)
        assert e.value.code == pgx.PGX_E_INPUT


def test_streams_on_a_share_of_the_compute_units():
    """pgx_stream_create_cus: launches on a stream confined to a few compute
    units give the oracle's results; ranges outside the device are rejected.
    (Each such stream takes a hardware queue of its own, of which a process
    has few: one stream here.)"""
    import torch
    L = gen(337, 2310, 16, 4, 1024, seed_off=3)
    ov, ores = O.verdicts(L[:1024], O.MODE_FERMAT, residues=True)
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    with pgx.Runner(1, limbs=16) as run:
        for first, count in ((n_cu - 8, 8),):
            say(f"stream on compute units [{first}, {first + count}) of {n_cu}")
            st = run.cu_stream(first, count)
            b = pgx.DeviceBatch(L)
            run.test_persistent(b.bucket(), pgx.MODE_FERMAT, residues=True, stream=st)
            run.test_device(b.cand.data_ptr(), b.ld, b.limbs, 1024, b.max_bits, pgx.MODE_FERMAT,
                            b.verdict.data_ptr(), b.res.data_ptr(), b.ld, st)
            torch.cuda.synchronize()
            run.destroy_stream(st)
            assert np.array_equal(b.verdicts()[:1024], ov)
            assert np.array_equal(b.residues()[:1024], ores)
        for first, count in ((0, 0), (n_cu, 1), (n_cu - 1, 2)):
            with pytest.raises(pgx.PgxError) as e:
                run.cu_stream(first, count)
            assert e.value.code == pgx.PGX_E_ARG


def test_runner_confined_to_a_share_of_the_compute_units():
    """pgx_runner_confine: batches of a runner confined to half the compute
    units (both batch slots, two batches in flight) give the oracle's verdicts
    and residues; it is refused while a batch is in flight and for parts out of
    range; parts = 1 lifts it."""
    L = gen(337, 2310, 16, 4, 2048, seed_off=5)
    a, b = L[:4096], L[4096:8192]
    va, ra = O.verdicts(a, O.MODE_FERMAT, residues=True)
    vb, rb = O.verdicts(b, O.MODE_FERMAT, residues=True)
    with pgx.Runner(4096, limbs=16) as run:
        for part, parts in ((1, 2), (0, 1)):
            say(f"runner on part {part} of {parts} of the compute units, two batches in flight")
            run.confine(part, parts)
            run.submit(a, pgx.MODE_FERMAT, residues=True)
            run.submit(b, pgx.MODE_FERMAT, residues=True)
            with pytest.raises(pgx.PgxError) as e:
                run.confine(0, 2)
            assert e.value.code == pgx.PGX_E_STATE
            got_a, got_ra = run.wait()
            got_b, got_rb = run.wait()
            assert np.array_equal(got_a, va) and np.array_equal(got_ra, ra)
            assert np.array_equal(got_b, vb) and np.array_equal(got_rb, rb)
        for part, parts in ((0, 0), (2, 2), (0, 1 << 20)):
            with pytest.raises(pgx.PgxError) as e:
                run.confine(part, parts)
            assert e.value.code == pgx.PGX_E_ARG


def test_base_range_is_checked():
    with pytest.raises(pgx.PgxError) as e:
        pgx.Runner(16, limbs=16, bases=[2, 256])
    assert e.value.code == pgx.PGX_E_ARG
    with pytest.raises(pgx.PgxError):
        pgx.Runner(16, limbs=16, bases=[1])
    with pgx.Runner(16, limbs=16, bases=[2, 255]) as r:  # the largest base works
        L = pgx.to_limbs([1000003, 3215031751, 2047], 16)
        v = r.run(L, pgx.MODE_STRONG)
        assert np.array_equal(v, O.verdicts(L, O.MODE_STRONG, (2, 255)))
