"""GPU parity of device-side candidate construction (fill_batch on the device,
gpu_testing.cu:136-195): N = m*K + sign*x built by pgx_build_candidates /
pgx_submit_mk, checked word for word against Python's bignums and, through
the PRP kernel, against the GMP oracle."""
import sys

import numpy as np
import pytest

import oracle as O
import pgx

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-build] {msg}", flush=True)
    sys.stdout.flush()


def primorial_over(P, d):
    """K = P# / d (gap_common.cpp:60-65)."""
    K = 1
    for q in range(2, P + 1):
        if all(q % r for r in range(2, int(q ** 0.5) + 1)):
            K *= q
    assert K % d == 0
    return K // d


def host_build(K, m, xs, sign):
    return [int(mi) * K + sign * int(x) for mi, x in zip(m, xs)]


def device_build(run, K, m, xs, sign, out_words, ld=None, per_candidate=True):
    """Run the construction kernel on torch buffers; (columns as ints, status)."""
    import torch
    n = len(m)
    ld = ld or n
    kw = max(1, (K.bit_length() + 31) // 32)
    dK = torch.from_numpy(pgx.to_limbs([K], kw)[0].view(np.int32)).cuda()
    dm = torch.from_numpy(np.asarray(m, dtype=np.uint64).view(np.int64)).cuda()
    dx = torch.from_numpy(np.asarray(xs, dtype=np.uint32).view(np.int32)).cuda()
    out = torch.full((out_words, ld), -1, dtype=torch.int32, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()  # the runner's stream does not wait for torch's
    run.build_candidates(dK.data_ptr(), kw, dm.data_ptr(), dx.data_ptr() if per_candidate else 0,
                         int(xs[0]), sign, n, out.data_ptr(), ld, out_words, status.data_ptr())
    torch.cuda.synchronize()
    cols = out.cpu().numpy().view(np.uint32)[:, :n].T
    return pgx.from_limbs(np.ascontiguousarray(cols)), int(status.cpu()[0])


@pytest.mark.parametrize("P,d", [(211, 210), (337, 2310), (907, 2190)])
@pytest.mark.parametrize("sign", [1, -1])
def test_build_matches_bignum(P, d, sign):
    K = primorial_over(P, d)
    words = (K.bit_length() + 64 + 31) // 32 + 1
    rng = np.random.default_rng(P * 7 + (sign > 0))
    m = rng.integers(1, 2**63, size=1000, dtype=np.uint64)
    m[:4] = [1, 2**64 - 1, 2**32 - 1, 2**32]  # carry corners of the 64-bit multiplier
    xs = rng.integers(0, 2**32, size=1000, dtype=np.uint64).astype(np.uint32)
    xs[:2] = [2**32 - 1, 0]
    say(f"build P={P} d={d} sign={sign} words={words}")
    with pgx.Runner(1, limbs=16) as run:
        got, status = device_build(run, K, m, xs, sign, words, ld=1024)
    assert status == 0
    assert got == host_build(K, m, xs, sign)


def test_build_single_x_and_status():
    K = primorial_over(337, 2310)
    kw = (K.bit_length() + 31) // 32
    m = np.arange(1, 300, dtype=np.uint64) * 1_000_003
    with pgx.Runner(1, limbs=16) as run:
        say("build with one x for every candidate")
        got, status = device_build(run, K, m, [12345] * len(m), 1, kw + 2, per_candidate=False)
        assert status == 0 and got == host_build(K, m, [12345] * len(m), 1)
        say("build into too few words: status set")
        _, status = device_build(run, K, m, [1] * len(m), 1, kw)
        assert status != 0
        say("negative candidate (m = 0, sign -1): status set")
        _, status = device_build(run, K, np.zeros(64, dtype=np.uint64), [5] * 64, -1, kw + 2)
        assert status != 0


@pytest.mark.parametrize("sign", [1, -1])
def test_submit_mk_matches_oracle(sign):
    """The prev/next sides of the reference's search: verdicts and residues of
    m*K + sign*x equal GMP's on the same numbers."""
    P, d = 337, 2310
    K = primorial_over(P, d)
    m = pgx.gen_multipliers(0x5EED0200 + (sign > 0), 10**7, 10**7 + 10**5, d, 2000)
    m = np.asarray(m, dtype=np.uint64)
    with pgx.Runner(len(m), limbs=16) as run:
        for x in (2, 7920, 30010, 1):  # m*K is odd: x = 1 makes every candidate even
            say(f"submit_mk P={P} sign={sign} x={x} n={len(m)}")
            v, r = run.run_mk(K, m, x, sign, pgx.MODE_STRONG, residues=True)
            L = pgx.to_limbs(host_build(K, m, [x] * len(m), sign), 16)
            ov, ores = O.verdicts(L, O.MODE_STRONG, residues=True)
            # even candidates (m*K + x with both odd) come back invalid
            even = (L[:, 0] & 1) == 0
            assert np.all(v[even] == pgx.VERDICT_INVALID)
            assert np.array_equal(v[~even], ov[~even])
            assert np.array_equal(r[~even], ores[~even])


def test_submit_mk_k_change_and_errors():
    K1 = primorial_over(211, 210)
    K2 = K1 + 2 * 3 * 5  # same word count, other value: the cached K must be replaced
    m = np.arange(101, 1101, 2, dtype=np.uint64)
    with pgx.Runner(len(m), limbs=16) as run:
        for K in (K1, K2, K1):
            say(f"submit_mk K of {K.bit_length()} bits")
            v = run.run_mk(K, m, 2, 1, pgx.MODE_FERMAT)
            L = pgx.to_limbs(host_build(K, m, [2] * len(m), 1), 16)
            odd = (L[:, 0] & 1) == 1
            assert odd.all()
            assert np.array_equal(v[odd], O.verdicts(L[odd], O.MODE_FERMAT))
        say("submit_mk wider than the runner: PGX_E_INPUT from wait")
        big = primorial_over(503, 210)  # ~700 bits > 16 words with m up to 2^64
        with pytest.raises(pgx.PgxError) as e:
            run.run_mk(big, np.full(8, 2**64 - 1, dtype=np.uint64), 1, 1, pgx.MODE_FERMAT)
        assert e.value.code == pgx.PGX_E_INPUT
        say("runner usable after the error")
        v = run.run_mk(K1, m, 1, 1, pgx.MODE_FERMAT)
        assert len(v) == len(m)
        with pytest.raises(pgx.PgxError):
            run.run_mk(K1, m, 1, 0, pgx.MODE_FERMAT)  # sign must be +-1
        assert len(run.run_mk(K1, m[:0], 1, 1, pgx.MODE_FERMAT)) == 0
