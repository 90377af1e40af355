"""The drop-in C++ surface: include/miller_rabin.h (test_runner_t over the C
ABI) and the gpu_testing.cu replacement, run by tests/dropin/harness under the
reference's own dispatch protocol (GPU_BATCHES GPU threads, GPUBatch state machine)."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from conftest import PKG, ROOT

REF = "/root/reference"
HARNESS = os.path.join(ROOT, "tests", "dropin", "harness")


def write_harness_input(path, P, d, m_start, m_inc, xs, active, composites):
    """The harness's input file: the interval, its active m_i and one
    composites bitmap per X (bit t <-> m_i = 2t + 1)."""
    words = len(composites[0])
    with open(path, "wb") as f:
        f.write(b"PGXH")
        f.write(struct.pack("<IIQIIII", P, d, m_start, m_inc, len(xs), words, len(active)))
        f.write(np.ascontiguousarray(active, dtype="<u4").tobytes())
        for X, c in zip(xs, composites):
            assert len(c) == words
            f.write(struct.pack("<I", X))
            f.write(np.ascontiguousarray(c, dtype="<u4").tobytes())


def read_harness_output(path):
    """(tests per X, found per X, found_prime_m_i words)."""
    with open(path, "rb") as f:
        raw = f.read()
    (n_x,) = struct.unpack_from("<I", raw, 0)
    pairs = np.frombuffer(raw, dtype="<u4", count=2 * n_x, offset=4).reshape(n_x, 2)
    (w,) = struct.unpack_from("<I", raw, 4 + 8 * n_x)
    found = np.frombuffer(raw, dtype="<u4", count=w, offset=8 + 8 * n_x)
    return pairs[:, 0].copy(), pairs[:, 1].copy(), found.copy()


def test_harness_is_built():
    """build() makes tests/dropin/harness where the reference is mounted; the
    GPU tests below run that binary."""
    if not os.path.isdir(REF) and not os.path.exists(HARNESS):
        pytest.skip("reference tree not mounted and no prebuilt harness")
    assert os.path.exists(HARNESS), "tests/dropin/harness is not built (build())"


def test_harness_stops_when_the_runner_cannot_be_created(tmp_path):
    """With no device visible, every run_gpu_thread fails at runner creation.
    The drop-in must stop the search (is_running = false, test_data DONE)
    instead of leaving the dispatch loop waiting for DONE forever."""
    if not os.path.exists(HARNESS):
        pytest.skip("harness not built")
    import oracle as O
    P, d, m_start, m_inc = 337, 2310, 10**7, 20_000
    active = O.ms_active(P, d, m_start, m_inc)
    comp = [np.zeros(O.bitmap_words(m_inc), dtype=np.uint32)]
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    write_harness_input(inp, P, d, m_start, m_inc, [2], active, comp)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="999", ROCR_VISIBLE_DEVICES="999")
    r = subprocess.run([HARNESS, str(inp), str(out)], capture_output=True, text=True,
                       errors="replace", timeout=120, env=env)
    assert r.returncode == 1, r.stdout + r.stderr
    assert "ERROR in run_gpu_thread" in r.stdout
    assert "the search stopped" in r.stderr


def test_gpu_testing_replacement_compiles_against_reference_headers(tmp_path):
    """host/gpu_testing.cpp must compile against the reference's own
    gap_search_gpu.h / gap_common.h / gap_stats.h / gpu_testing.h (TestData,
    GPUBatch, run_gpu_thread, gpu_state_and_checks) with the drop-in
    miller_rabin.h and pgx.h first on the include path."""
    if not os.path.isdir(REF):
        pytest.skip("reference tree not mounted")
    gmp_inc = "/opt/conda/include"
    cmd = ["g++", "-std=c++20", "-Wall", "-Werror", "-fsyntax-only",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{REF}", f"-I{gmp_inc}",
           os.path.join(PKG, "host", "gpu_testing.cpp")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_shim_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text(
        '#include "miller_rabin.h"\n'
        "typedef mr_params_t<8, 1024, 5> params;\n"
        "static_assert(params::BITS == 1024 && params::TPI == 8 && params::WINDOW_BITS == 5);\n"
        "void f(std::vector<mpz_t*>& z, std::vector<int>& r) {\n"
        "  test_runner_t<params> runner(8192, 1);\n"
        "  runner.run_test(z.size(), z, r);\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only",
                        f"-I{os.path.join(ROOT, 'include')}", "-I/opt/conda/include", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_shim_selftest_on_gpu():
    """test_runner_t::run_test on mpz candidates agrees with mpz_probab_prime_p
    (rounds 1 and 3; inactive tail untouched)."""
    exe = os.path.join(PKG, "shim_selftest")
    assert os.path.exists(exe), "shim_selftest is not built (make in prime-gap-gpu_amd)"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 mismatches" in r.stdout


def run_harness_case(tmp_path, P, d, m_start, m_inc, max_prime, xs):
    """The drop-in over an interval against the oracle's search step
    (orc_ms_step: finalize, fill_batch, the verdicts, process_finished_batch,
    remove_vector) X by X: per-X tests and found counts and the found bitmap."""
    import oracle as O
    import pgx
    K = pgx.primorial_over(P, d)
    kw = pgx.to_limbs([K], (K.bit_length() + 31) // 32)[0]
    active = O.ms_active(P, d, m_start, m_inc)
    comps = [O.ms_sieve(kw, P, d, m_start, m_inc, max_prime, X) for X in xs]
    want_tests, want_found = [], []
    found = np.zeros(O.bitmap_words(m_inc), dtype=np.uint32)
    a = active
    for X, c in zip(xs, comps):
        tests, fmi, a = O.ms_step(kw, m_start, X, c, a, found, [2])
        want_tests.append(len(tests))
        want_found.append(len(fmi))
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    write_harness_input(inp, P, d, m_start, m_inc, xs, active, comps)
    print(f"[dropin] harness P={P} d={d} m_start={m_start} m_inc={m_inc} X={xs}", flush=True)
    sys.stdout.flush()
    r = subprocess.run([HARNESS, str(inp), str(out), "2"], capture_output=True, text=True,
                       errors="replace", timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    tests, nfound, fwords = read_harness_output(out)
    assert list(tests) == want_tests
    assert list(nfound) == want_found
    w = len(fwords)  # (m_inc/2)/32 + 1 words; the oracle keeps one more, always zero
    assert np.array_equal(fwords, found[:w])
    assert not found[w:].any()
    return sum(want_tests), sum(want_found)


@pytest.mark.gpu
def test_dropin_readme_interval(tmp_path):
    """README's interval (P=337, d=2310, m_start=1e7, m_inc=2e5, primes to 1e6),
    6 X steps: one batch per X, the other GPU thread retires empty."""
    assert os.path.exists(HARNESS), "tests/dropin/harness is not built"
    nt, nf = run_harness_case(tmp_path, 337, 2310, 10**7, 200_000, 10**6, [2, 4, 6, 8, 10, 12])
    assert nt > 30_000 and nf > 2_000


@pytest.mark.gpu
def test_dropin_two_batches_in_flight(tmp_path):
    """58,614 tests at one X: two batches of up to 32,768 run on two GPU threads
    at once, and found_prime_m_i collects both."""
    assert os.path.exists(HARNESS), "tests/dropin/harness is not built"
    nt, _ = run_harness_case(tmp_path, 337, 2310, 10**7, 1_200_000, 10**6, [6])
    assert nt > 32_768
