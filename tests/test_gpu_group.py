"""GPU parity of the kernels that spread one candidate over a group of 2, 4, 8,
16 or 32 lanes (lanes_per_number in pgx_tuning_t): verdicts of every mode and
residues against the GMP oracle and against the one-lane kernels, through
persistent (bucket) and indexed launches, at every group kernel width."""
import sys

import numpy as np
import pytest

import oracle as O
import pgx

pytestmark = pytest.mark.gpu

# (P, d, buffer words): sieved candidates of ~296, 462, 710, 987, 1260 bits
SIZES = [(211, 210, 16), (337, 2310, 16), (503, 210, 24), (701, 210, 32), (907, 30030, 64)]


def say(msg):
    print(f"[gpu-group] {msg}", flush=True)
    sys.stdout.flush()


def gen(P, d, words, n_m=4, per_m=96, seed=0):
    m = pgx.gen_multipliers(0x5EED0300 + P + seed, 10**7, 10**7 + 10**4, d, n_m)
    L, _, _ = pgx.gen_sieved(P, d, m, 450000, per_m, words)
    return L


def run_bucket(run, L, mode, lanes, residues=True):
    import torch
    a = pgx.DeviceBatch(L)
    t = pgx.Tuning(8, 4, 64, lanes)
    run.test_persistent(a.bucket(), mode, residues=residues, tuning=t)
    torch.cuda.synchronize()
    return a.verdicts(), (a.residues() if residues else None)


@pytest.mark.parametrize("lanes", [2, 4, 8, 16, 32])
def test_group_fermat_residues_all_sizes(lanes):
    with pgx.Runner(1, limbs=16) as run:
        for P, d, words in SIZES:
            L = gen(P, d, words)
            say(f"lanes={lanes} P={P} {len(L)} candidates, Fermat + residues")
            v, r = run_bucket(run, L, pgx.MODE_FERMAT, lanes)
            ov, ores = O.verdicts(L, O.MODE_FERMAT, residues=True)
            assert np.array_equal(v, ov), P
            assert np.array_equal(r, ores), P


@pytest.mark.parametrize("lanes", [2, 4, 8, 16, 32])
def test_group_strong_and_reference_bases(lanes):
    with pgx.Runner(1, limbs=16, bases=[2, 3, 5, 7]) as run:
        for P, d, words in SIZES[:4]:
            L = gen(P, d, words, n_m=2, per_m=64, seed=1)
            for mode, omode in ((pgx.MODE_STRONG, O.MODE_STRONG),
                                (pgx.MODE_REFERENCE, O.MODE_REFERENCE)):
                say(f"lanes={lanes} P={P} mode={mode} bases 2,3,5,7")
                v, r = run_bucket(run, L, mode, lanes)
                assert np.array_equal(v, O.verdicts(L, omode, (2, 3, 5, 7))), (P, mode)
                assert np.array_equal(r, O.verdicts(L, O.MODE_FERMAT, residues=True)[1]), P


def test_group_kats_and_tiny_n():
    """Strong pseudoprimes and every odd N < 2^12 in the narrowest group
    kernels (N much smaller than the kernel)."""
    Ns = list(range(3, 1 << 12, 2)) + [2047, 3277, 1373653, 25326001, 3215031751]
    L = pgx.to_limbs(Ns, 16)
    with pgx.Runner(1, limbs=16, bases=[2, 3, 5, 7]) as run:
        for lanes in (2, 4, 8, 16, 32):
            say(f"lanes={lanes} tiny N and KATs")
            v, r = run_bucket(run, L, pgx.MODE_STRONG, lanes)
            assert np.array_equal(v, O.verdicts(L, O.MODE_STRONG, (2, 3, 5, 7))), lanes
            v, r = run_bucket(run, L, pgx.MODE_FERMAT, lanes)
            ov, ores = O.verdicts(L, O.MODE_FERMAT, residues=True)
            assert np.array_equal(v, ov) and np.array_equal(r, ores), lanes


def test_group_widest_and_indexed():
    """2000-bit candidates (the 76-limb four-lane, 80-limb eight- and
    sixteen-lane and 96-limb thirty-two-lane kernels) and an indexed launch
    over a compacted list, equal to the one-lane kernels."""
    import torch
    rng = np.random.default_rng(7)
    big = [int.from_bytes(rng.bytes(250), "little") | (1 << 1999) | 1 for _ in range(300)]
    L = pgx.to_limbs(big, 64)
    with pgx.Runner(1, limbs=64) as run:
        ov, ores = O.verdicts(L, O.MODE_FERMAT, residues=True)
        for lanes in (4, 8, 16, 32):
            say(f"lanes={lanes} 2000-bit Fermat")
            v, r = run_bucket(run, L, pgx.MODE_FERMAT, lanes)
            assert np.array_equal(v, ov) and np.array_equal(r, ores), lanes
    P, d, words = SIZES[3]
    L = gen(P, d, words, n_m=4, per_m=128, seed=2)
    st = torch.cuda.Stream().cuda_stream  # one stream orders both runners' launches
    with pgx.Runner(1, limbs=words) as fermat, pgx.Runner(1, limbs=words, bases=[3, 5]) as mr:
        a = pgx.DeviceBatch(L)
        index = torch.zeros(len(L), dtype=torch.int32, device="cuda")
        count = torch.zeros(1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        fermat.test_persistent(a.bucket(), pgx.MODE_FERMAT, stream=st)
        fermat.compact(a.verdict.data_ptr(), len(L), pgx.FLAG_FERMAT, index.data_ptr(),
                       count.data_ptr(), stream=st)
        out = {}
        for lanes in (1, 2, 4, 8, 16, 32):
            vv = torch.full((len(L),), -1, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            b = pgx.Bucket(a.cand.data_ptr(), a.ld, a.limbs, len(L), a.max_bits, vv.data_ptr(),
                           0, a.ld)
            say(f"indexed MR(3,5) lanes={lanes}")
            mr.test_indexed(b, index.data_ptr(), count.data_ptr(), pgx.MODE_STRONG,
                            tuning=pgx.Tuning(8, 4, 64, lanes), stream=st)
            torch.cuda.synchronize()
            out[lanes] = vv.cpu().numpy().view(np.uint32)
        k = int(count.cpu()[0])
        surv = np.sort(index.cpu().numpy()[:k].view(np.uint32))
        assert k == int(((a.verdicts() & pgx.FLAG_FERMAT) != 0).sum())
        want = O.verdicts(L[surv], O.MODE_STRONG, (3, 5))
        for lanes in (1, 2, 4, 8, 16, 32):
            assert np.array_equal(out[lanes][surv], want), lanes
            rest = np.setdiff1d(np.arange(len(L)), surv)
            assert np.all(out[lanes][rest] == 0xFFFFFFFF), lanes  # untouched columns


def test_group_autotune_and_plan():
    with pgx.Runner(1, limbs=32) as run:
        L = gen(701, 210, 32, n_m=2, per_m=256, seed=3)
        a = pgx.DeviceBatch(L)
        t, rate = run.autotune(a.bucket(), pgx.MODE_FERMAT)
        say(f"autotune over lanes: {t.as_dict()} {rate / 1e6:.2f} M/s")
        assert t.lanes_per_number in (1, 2, 4, 8, 16, 32) and rate > 0
        b = pgx.DeviceBatch(L)
        plan = pgx.Plan(run, [b.bucket()], pgx.MODE_FERMAT, residues=True,
                        tunings=[pgx.Tuning(8, 4, 64, 4)])
        import torch
        for _ in range(2):
            plan.launch()
            torch.cuda.synchronize()
            ov, ores = O.verdicts(L, O.MODE_FERMAT, residues=True)
            assert np.array_equal(b.verdicts(), ov) and np.array_equal(b.residues(), ores)
        plan.close()
        for bad in (3, 64):
            with pytest.raises(pgx.PgxError):
                run.test_persistent(a.bucket(), pgx.MODE_FERMAT, tuning=pgx.Tuning(8, 4, 64, bad))
