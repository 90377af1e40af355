"""CPU-side tests of the product's host code: candidate builder vs the oracle,
limb packing, the C ABI's exported symbols, loud failure without a GPU, and the
host build of the kernels' per-lane routine vs the oracle."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
import pgx
from conftest import PKG, ROOT, limbs_needed


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "pgx.h")).read()
    declared = set(re.findall(r"\b(pgx_[a-z_]+)\s*\(", hdr))
    lib = ctypes.CDLL(pgx.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert declared >= set(pgx.EXPORTED)


def test_first_primes_matches_generate_primes():
    # miller_rabin.h:272-289: list[0] = 2, then odd numbers not divisible by an
    # earlier odd prime
    assert pgx.first_primes(10) == [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]


def test_limb_packing_round_trip():
    rng = np.random.default_rng(1)
    vals = [int(x) for x in rng.integers(1, 1 << 62, size=50)] + [2**511 + 1, 2**512 - 1, 1]
    arr = pgx.to_limbs(vals, 16)
    assert pgx.from_limbs(arr) == vals
    # same layout as mpz_export(order=-1, size=4): limb 0 is the low word
    assert arr[-3, 0] == 1 and arr[-3, 15] == 0x80000000
    with pytest.raises(pgx.PgxError):
        pgx.to_limbs([2**512], 16)


@pytest.mark.parametrize("P,d,limbs", [(337, 2310, 16), (211, 210, 16), (503, 210, 24),
                                       (701, 210, 32), (907, 30030, 64)])
def test_generator_matches_oracle(P, d, limbs):
    seed = 0x5EED0000 + P
    m = pgx.gen_multipliers(seed, 10**7, 10**7 + 10**4, d, 3)
    mo = O.gen_multipliers(seed, 10**7, 10**7 + 10**4, d, 3)
    assert np.array_equal(m, mo)
    a = pgx.gen_sieved(P, d, m, 30000, 300, limbs)
    b = O.gen_sieved(P, d, mo, 30000, 300, limbs)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert len(a[0]) > 0


@pytest.mark.parametrize("P,d,limbs", [(337, 2310, 16), (907, 2310, 48)])
def test_prev_side_generator_matches_oracle(P, d, limbs):
    m = pgx.gen_multipliers(17 + P, 10**7, 10**7 + 10**4, d, 2)
    a = pgx.gen_sieved(P, d, m, 20000, 400, limbs, sign=-1)
    b = O.gen_sieved(P, d, m, 20000, 400, limbs, sign=-1)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    # N = m*K - X, and no survivor has a factor below the sieve limit
    # with no sieving primes, X = 1 survives: m*K + 1
    mK = pgx.from_limbs(pgx.gen_sieved(P, d, m[:1], 1, 1, limbs, sieve_limit=2)[0])[0] - 1
    assert pgx.from_limbs(a[0][:3]) == [mK - int(x) for x in a[2][:3]]
    assert not O.small_factor(a[0][:50], 3000).any()


def test_generator_rejects_bad_input():
    assert pgx.k_bits(337, 7 * 1009) == 0  # 1009 does not divide 337#
    with pytest.raises(pgx.PgxError):
        pgx.gen_sieved(337, 1009, np.array([10**7 + 1], dtype=np.uint64), 100, 10, 16)
    with pytest.raises(pgx.PgxError):  # N (~462 bits) does not fit 8 limbs
        pgx.gen_sieved(337, 2310, np.array([10**7 + 1], dtype=np.uint64), 1000, 10, 8)
    with pytest.raises(pgx.PgxError):  # not enough coprime multipliers
        pgx.gen_multipliers(1, 10, 12, 2310, 5)


def test_k_bits_matches_survey_table():
    assert pgx.k_bits(337, 2310) == 439
    assert pgx.k_bits(907, 30030) == 1236


def test_runner_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(pgx.PgxError) as e:
        pgx.Runner(16, rounds=1, limbs=16)
    assert e.value.code == pgx.PGX_E_NODEV


# ---- host build of the per-lane kernel routine -----------------------------

EMU_PATH = os.path.join(ROOT, "prime-gap-gpu_amd", "libpgx_lane_emu.so")


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU_PATH):
        pytest.fail("libpgx_lane_emu.so is not built")
    lib = ctypes.CDLL(EMU_PATH)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    lib.pgx_lane_emu.argtypes = [u32p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint32, u32p, ctypes.c_uint32, u32p, u32p]
    lib.pgx_lane_emu.restype = ctypes.c_uint32

    def run(L, S, mode, bases=(2,)):
        L = np.ascontiguousarray(L, dtype=np.uint32)
        n, limbs = L.shape
        v = np.zeros(n, np.uint32)
        r = np.zeros_like(L)
        b = np.asarray(bases, np.uint32)
        s = lib.pgx_lane_emu(pgx._u32(L), n, limbs, S, mode, pgx._u32(b), len(b), pgx._u32(v),
                             pgx._u32(r))
        assert s > 0
        return v, r
    return run


def test_lane_routine_tiny_brute_force(emu):
    L = pgx.to_limbs(range(3, 1 << 13, 2), 1)
    for S in (1, 2, 3):
        v, r = emu(L, S, pgx.MODE_FERMAT)
        ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
        assert np.array_equal(v, ov) and np.array_equal(r, orr)


def test_lane_routine_fixture_records(emu, vectors):
    recs = [r for r in vectors["kat"] + vectors["widths"] if int(r["n"], 16).bit_length() <= 734]
    ns = [int(r["n"], 16) for r in recs]
    L = pgx.to_limbs(ns, 24)
    v, r = emu(L, 26, pgx.MODE_FERMAT)  # 26 radix-2^29 limbs hold 750 bits
    assert pgx.from_limbs(r) == [int(x["res2"], 16) for x in recs]
    for i, rec in enumerate(recs):
        assert bool(v[i] & 1) == rec["fermat2"] and bool(v[i] & 2) == rec["sprp"]["2"]
    for mode, key in ((pgx.MODE_REFERENCE, "ref"), (pgx.MODE_STRONG, "sprp")):
        v, _ = emu(L, 26, mode, (2, 3, 5, 7))
        for i, rec in enumerate(recs):
            want = 0
            for a in "2357":
                if not rec[key][a]:
                    break
                want += 1
            assert v[i] == want


def test_lane_routine_sieved_vs_oracle(emu):
    m = pgx.gen_multipliers(0x5EED0001, 10**7, 10**7 + 10**4, 2310, 2)
    L, _, _ = pgx.gen_sieved(337, 2310, m, 200000, 200, 16)
    ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
    for S in (17, 20):  # 462-bit candidates: 17 limbs is the narrowest that holds them
        for mode in (pgx.MODE_FERMAT, pgx.MODE_FERMAT | 0x100):  # both folded-squaring forms
            v, r = emu(L, S, mode)
            assert np.array_equal(v, ov) and np.array_equal(r, orr), (S, mode)
    v, _ = emu(L, 17, pgx.MODE_REFERENCE, (2, 3, 5))
    assert np.array_equal(v, O.verdicts(L, O.MODE_REFERENCE, (2, 3, 5)))


def test_lane_routine_every_width_gap(emu):
    """Random odd N of every bit length from 60 bits below a kernel's top up
    to its top (R/N from 2^64 down to 2^4: Mont(1) = R mod N after 5 to 65
    doublings), residues and verdicts against GMP."""
    import random
    rng = random.Random(5)
    for S in (5, 12, 17, 26):
        top = 29 * S - 4
        vals = [rng.getrandbits(b) | (1 << (b - 1)) | 1
                for b in range(top - 60, top + 1) for _ in range(4)]
        L = pgx.to_limbs(vals, (29 * S + 31) // 32)
        v, r = emu(L, S, pgx.MODE_FERMAT)
        ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
        assert np.array_equal(v, ov) and np.array_equal(r, orr), S
        v, _ = emu(L, S, pgx.MODE_STRONG, (2, 3, 5))
        assert np.array_equal(v, O.verdicts(L, O.MODE_STRONG, (2, 3, 5))), S


def test_lane_routine_top_limb_zero_squaring(emu):
    """The base-2 squaring that skips a zero top limb (mode bit 0x200) holds
    for N of at most 29(S-1) - 1 bits; a lane with a longer N runs the
    ordinary folded loop.  Random odd N from 60 bits below that bound up to
    the kernel's top (both loops, and the bound itself on either side),
    residues and verdicts against GMP in every base-2 mode.  S = 26 has no
    folded squaring (the column bound) and must give the same results too."""
    import random
    rng = random.Random(11)
    for S in (2, 3, 5, 12, 17, 26):
        bound = 29 * (S - 1) - 1
        top = 29 * S - 4
        lens = range(max(3, bound - 60), top + 1)
        vals = [rng.getrandbits(b) | (1 << (b - 1)) | 1 for b in lens for _ in range(3)]
        vals += [(1 << bound) - 1, (1 << bound) + 1, (1 << (bound - 1)) + 1]
        L = pgx.to_limbs(vals, (29 * S + 31) // 32)
        v, r = emu(L, S, pgx.MODE_FERMAT | 0x200)
        ov, orr = O.verdicts(L, O.MODE_FERMAT, residues=True)
        assert np.array_equal(v, ov) and np.array_equal(r, orr), S
        for mode, omode in ((pgx.MODE_REFERENCE, O.MODE_REFERENCE),
                            (pgx.MODE_STRONG, O.MODE_STRONG)):
            v, _ = emu(L, S, mode | 0x200, (2,))
            assert np.array_equal(v, O.verdicts(L, omode, (2,))), (S, mode)
    # the 462-bit search candidates (17 limbs: bound 463 bits) and small N
    m = pgx.gen_multipliers(0x5EED0006, 10**7, 10**7 + 10**4, 2310, 2)
    L, _, _ = pgx.gen_sieved(337, 2310, m, 200000, 300, 16)
    for arr in (L, pgx.to_limbs(list(range(3, 1 << 12, 2)), 16)):
        v, r = emu(arr, 17, pgx.MODE_FERMAT | 0x200)
        ov, orr = O.verdicts(arr, O.MODE_FERMAT, residues=True)
        assert np.array_equal(v, ov) and np.array_equal(r, orr)


def test_lane_routine_rejects_what_does_not_fit(emu):
    # even, one, and a value too wide for one radix-2^29 limb (R = 2^29 >= 16N
    # allows 25-bit N)
    L = pgx.to_limbs([10, 1, (1 << 26) - 1, (1 << 25) - 1], 1)
    v, _ = emu(L, 1, pgx.MODE_FERMAT)
    assert list(v[:3]) == [pgx.VERDICT_INVALID] * 3
    assert v[3] != pgx.VERDICT_INVALID


def test_lane_routine_small_bases_tiny_n(emu):
    """Miller-Rabin bases other than 2 (multiply-by-a steps) on every odd
    N < 4096, in kernels much wider than N, against CPython's pow."""
    Ns = list(range(3, 4096, 2))
    L = pgx.to_limbs(Ns, 4)
    bases = (3, 5, 251)
    want = []
    for N in Ns:
        passed = 0
        for a in bases:
            if a % N == 0 or not O.py_ref_verdict(N, a):
                break
            passed += 1
        want.append(passed)
    for S in (1, 2, 3, 10):
        v, _ = emu(L, S, pgx.MODE_REFERENCE, bases)
        assert v.tolist() == want, S


def test_bench_refuses_several_gpus():
    import subprocess
    import sys as _sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([_sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_lane_routine_base_two_alone_in_every_mode(emu):
    """The base list {2} runs the base-2 routine in every mode (as the kernels
    are launched); its verdict counts equal the general routine's (list
    {2, 3}, first base) and GMP's."""
    m = pgx.gen_multipliers(0x5EED0005, 10**7, 10**7 + 10**4, 2310, 2)
    L, _, _ = pgx.gen_sieved(337, 2310, m, 200000, 300, 16)
    Ns = list(range(3, 4096, 2))
    for arr in (L, pgx.to_limbs(Ns, 16)):
        for mode, omode in ((pgx.MODE_REFERENCE, O.MODE_REFERENCE),
                            (pgx.MODE_STRONG, O.MODE_STRONG)):
            v, r = emu(arr, 17, mode, (2,))
            want = O.verdicts(arr, omode, (2,))
            assert np.array_equal(v, want), mode
            v23, _ = emu(arr, 17, mode, (2, 3))
            assert np.array_equal(np.minimum(v23, 1), v), mode
            _, ores = O.verdicts(arr, O.MODE_FERMAT, residues=True)
            assert np.array_equal(r, ores), mode


def test_counter_slots_two_threads_share_one_runner(tmp_path):
    """Persistent launches take work-counter slots from their runner
    (csrc/slots.hpp).  Two host threads launching on one runner must get
    different slots: the first kCounters - 1 = 63 slots handed out, whichever
    thread took them, are all distinct and lie in 1..63 (slot 0 is the survivor
    pass's)."""
    src = tmp_path / "slots.cpp"
    src.write_text(
        '#include "slots.hpp"\n'
        "#include <cstdio>\n#include <thread>\n#include <vector>\n#include <atomic>\n"
        "int main() {\n"
        "  pgx::CounterSlots slots(64);\n"
        "  std::vector<unsigned> a(31), b(32);\n"
        "  std::atomic<int> go{0};\n"
        "  auto run = [&](std::vector<unsigned>& out) {\n"
        "    go.fetch_add(1); while (go.load() < 2) {}\n"
        "    for (auto& v : out) v = slots.next(); };\n"
        "  std::thread t1(run, std::ref(a)), t2(run, std::ref(b));\n"
        "  t1.join(); t2.join();\n"
        "  for (unsigned v : a) printf(\"%u \", v);\n"
        "  for (unsigned v : b) printf(\"%u \", v);\n"
        "  for (int k = 0; k < 126; ++k) printf(\"%u \", slots.next());\n"
        "  return 0; }\n")
    exe = tmp_path / "slots"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", f"-I{os.path.join(PKG, 'csrc')}",
                    str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]
    first = vals[:63]
    assert sorted(first) == list(range(1, 64))
    # afterwards the slots keep coming round in order, each once per 63
    rest = vals[63:]
    assert all(rest[i] % 63 + 1 == rest[i + 1] for i in range(len(rest) - 1))


def _serve_load(*args):
    if not os.path.exists(pgx.SERVE_LOAD):
        pytest.skip("prime-gap-gpu_amd/serve_load is not built (build())")
    return subprocess.run([pgx.SERVE_LOAD, *map(str, args)], capture_output=True, text=True,
                          timeout=60)


def test_serve_load_rejects_bad_arguments_before_any_gpu_call(tmp_path):
    """The C++ serving load generator exits 2 on bad arguments or template
    files, before it creates a server (so this runs without a GPU)."""
    assert _serve_load().returncode == 2
    assert _serve_load(tmp_path / "missing.bin", tmp_path / "o", 1000, 1, 1).returncode == 2
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"XXXX\x00\x00\x00\x00")
    assert _serve_load(bad, tmp_path / "o", 1000, 1, 1).returncode == 2
    good = tmp_path / "t.bin"
    cands = pgx.to_limbs([2**61 - 1, 2**89 - 1], 4)
    pgx.write_serve_templates(str(good), [(89, cands, cands[:1])])
    r = _serve_load(good, tmp_path / "o", 0, 1, 1)  # no requests
    assert r.returncode == 2 and "RATE" in r.stderr
    r = _serve_load(good, tmp_path / "o", 1000, 1, 1, "colour=blue")
    assert r.returncode == 2 and "unknown key" in r.stderr
    # a truncated template (header promises more candidates than the file holds)
    raw = good.read_bytes()
    short = tmp_path / "short.bin"
    short.write_bytes(raw[:-4])
    assert _serve_load(short, tmp_path / "o", 1000, 1, 1).returncode == 2


def test_serve_templates_layout(tmp_path):
    """write_serve_templates: "PGXS", count, then per template limbs, max_bits,
    n_next, n_prev and the candidate-major words of each side."""
    import struct
    nxt = pgx.to_limbs([2**61 - 1, 2**89 - 1, 2**107 - 1], 4)
    prv = pgx.to_limbs([2**31 - 1], 4)
    path = tmp_path / "t.bin"
    pgx.write_serve_templates(str(path), [(107, nxt, prv), (107, prv, nxt[:0])])
    raw = path.read_bytes()
    assert raw[:4] == b"PGXS" and struct.unpack_from("<I", raw, 4)[0] == 2
    assert struct.unpack_from("<IIII", raw, 8) == (4, 107, 3, 1)
    words = np.frombuffer(raw, dtype="<u4", count=16, offset=24)
    assert np.array_equal(words[:12].reshape(3, 4), nxt)
    assert np.array_equal(words[12:].reshape(1, 4), prv)
    assert struct.unpack_from("<IIII", raw, 24 + 64) == (4, 107, 1, 0)
    assert len(raw) == 24 + 64 + 16 + 16
