"""The drop-in C++ surface: include/miller_rabin.h (test_runner_t over the C
ABI) and the gpu_testing.cu replacement."""
import os
import shutil
import subprocess

import pytest

from conftest import PKG, ROOT

REF = "/root/reference"


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
