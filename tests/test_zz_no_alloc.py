"""pgx_test_device_noalloc (no runner, nothing allocated: its scratch is a
device array of the library) against pgx_run, whose kernels run the same
first-base / compaction / side-by-side pipeline in a runner's scratch: two
calls in a row, the second with a smaller input, bit-exact verdicts and
residues; the stale scratch of the first call must not leak into the second."""
import sys

import numpy as np
import pytest

import oracle as O
import pgx

pytestmark = pytest.mark.gpu

BASES = [2, 3, 5, 7]
# strong pseudoprimes: to base 2 (2047, 3277), to 2 and 3 (1373653), to 2, 3
# and 5 (25326001), to 2, 3, 5 and 7 (3215031751): survivors of the first base
# that fail a later one, and one that passes all four
SPSP = [2047, 3277, 1373653, 25326001, 3215031751]


def say(msg):
    print(f"[no-alloc] {msg}", flush=True)
    sys.stdout.flush()


def candidates(seed, n, spsp=()):
    m = pgx.gen_multipliers(seed, 10**7, 10**7 + 10**4, 2310, 8)
    c, _, _ = pgx.gen_sieved(337, 2310, m, 450000, (n + 7) // 8, 16)
    c = c[:n].copy()
    for k, v in enumerate(spsp):  # spread the pseudoprimes over the batch
        c[(k * 997) % n] = pgx.to_limbs([v], 16)[0]
    return c


def run_noalloc(cands, mode, bases, residues):
    import torch
    n, limbs = cands.shape
    d = torch.from_numpy(np.ascontiguousarray(cands.T).view(np.int32)).cuda()
    v = torch.full((n,), -7, dtype=torch.int32, device="cuda")  # stale values to overwrite
    r = torch.full((limbs, n), -7, dtype=torch.int32, device="cuda") if residues else None
    max_bits = max(int(x).bit_length() for x in pgx.from_limbs(cands))
    stream = torch.cuda.Stream()
    pgx.test_device_noalloc(d.data_ptr(), n, limbs, n, max_bits, mode, bases, v.data_ptr(),
                            r.data_ptr() if residues else 0, n if residues else 0,
                            stream.cuda_stream)
    torch.cuda.synchronize()
    vh = v.cpu().numpy().view(np.uint32)
    rh = r.cpu().numpy().view(np.uint32).T.copy() if residues else None
    return vh, rh


def run_existing(cands, mode, bases, residues):
    with pgx.Runner(len(cands), limbs=cands.shape[1], bases=bases) as run:
        return run.run(cands, mode, residues=True) if residues else (run.run(cands, mode), None)


@pytest.mark.parametrize("mode", [pgx.MODE_REFERENCE, pgx.MODE_STRONG])
def test_two_calls_several_bases_match_pgx_run(mode):
    first = candidates(0x5EED00A1, 20011, SPSP)
    second = candidates(0x5EED00A2, 3001, SPSP[::-1])
    for k, c in enumerate((first, second)):
        say(f"mode={mode} call {k + 1}: n={len(c)} bases={BASES}")
        got, _ = run_noalloc(c, mode, BASES, residues=False)
        want, _ = run_existing(c, mode, BASES, residues=False)
        assert np.array_equal(got, want), f"call {k + 1}: {int(np.sum(got != want))} differ"
        assert set(np.unique(got)) >= {0, 1, 4}
    # the second call against the oracle as well (GMP, every base)
    assert np.array_equal(got, O.verdicts(second, O.MODE_REFERENCE if mode == pgx.MODE_REFERENCE
                                          else O.MODE_STRONG, bases=BASES))


def test_two_calls_fermat_residues_match_pgx_run():
    for k, n in enumerate((8192, 1000)):
        c = candidates(0x5EED00B0 + k, n)
        say(f"FERMAT call {k + 1}: n={n} residues")
        got_v, got_r = run_noalloc(c, pgx.MODE_FERMAT, [2], residues=True)
        want_v, want_r = run_existing(c, pgx.MODE_FERMAT, [2], residues=True)
        assert np.array_equal(got_v, want_v)
        assert np.array_equal(got_r, want_r)


def test_rejects_more_than_the_maximum():
    import torch
    d = torch.zeros(16, dtype=torch.int32, device="cuda")
    with pytest.raises(pgx.PgxError) as e:
        pgx.test_device_noalloc(d.data_ptr(), pgx.NOALLOC_MAX_N + 1, 16, pgx.NOALLOC_MAX_N + 1,
                                462, pgx.MODE_REFERENCE, BASES, d.data_ptr())
    assert e.value.code == pgx.PGX_E_CAPACITY
    assert "maximum" in str(e.value)
