"""CPU tests of the alignment pass (prime-gap-gpu_amd/csrc/align_isa.py) that
the build runs over the device assembly of the lane-group kernels and of the
widest one-lane kernels: on a small kernel compiled here, the aligned object
holds the compiler's instructions with every 8-byte instruction 8-byte
aligned, and the checker rejects an object with an instruction dropped or a
branch moved."""
import os
import re
import subprocess

import pytest

from conftest import PKG

ROCM = "/opt/rocm"
HIPCC = os.path.join(ROCM, "bin", "hipcc")
CLANG = os.path.join(ROCM, "llvm", "bin", "clang")
ALIGN = os.path.join(PKG, "csrc", "align_isa.py")

SRC = r"""
#include <hip/hip_runtime.h>
#include <stdint.h>
__global__ void k_mix(const uint32_t* a, uint32_t* out, int n) {
  uint32_t x = a[threadIdx.x], y = a[threadIdx.x + 64];
  uint64_t acc = 0;
  for (int i = 0; i < n; ++i) {
    acc = (uint64_t)x * y + acc;
    x = (uint32_t)(acc >> 7) & 0x1fffffffu;
    y ^= (x << (i & 7)) | 1u;
    if (y & 0x100u) acc += y;
  }
  out[threadIdx.x] = (uint32_t)acc ^ x;
}
"""

pytestmark = pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)),
                                reason="ROCm compiler not installed")


def _run(*args, **kw):
    return subprocess.run(list(args), capture_output=True, text=True, **kw)


@pytest.fixture(scope="module")
def built(tmp_path_factory):
    d = tmp_path_factory.mktemp("align")
    src = d / "k.hip"
    src.write_text(SRC)
    dev_s = d / "k.s"
    r = _run(HIPCC, "--offload-arch=gfx950", "-O3", "--cuda-device-only", "-S", "-o", str(dev_s),
             str(src))
    assert r.returncode == 0, r.stderr
    out_o = d / "k_al.o"
    r = _run("python3", ALIGN, "build", str(dev_s), str(out_o))
    assert r.returncode == 0, r.stdout + r.stderr
    return d, dev_s, out_o, r.stdout


def test_aligned_object_matches_the_compiler(built):
    d, dev_s, out_o, log = built
    assert "1 functions match the compiler's, 0 kept as compiled" in log
    # and the checker alone agrees
    r = _run("python3", ALIGN, "verify", str(out_o) + ".orig.o", str(out_o))
    assert r.returncode == 0, r.stderr


def _mutate_and_verify(built, mutate):
    d, dev_s, out_o, _ = built
    al_s = str(out_o) + ".al.s"
    lines = open(al_s).read().split("\n")
    bad_s = d / "bad.s"
    bad_s.write_text("\n".join(mutate(lines)))
    bad_o = d / "bad.o"
    r = _run(CLANG, "-target", "amdgcn-amd-amdhsa", "-mcpu=gfx950", "-c", str(bad_s), "-o",
             str(bad_o))
    assert r.returncode == 0, r.stderr
    return _run("python3", ALIGN, "verify", str(out_o) + ".orig.o", str(bad_o))


def test_checker_rejects_a_dropped_instruction(built):
    def drop(lines):
        i = next(i for i, l in enumerate(lines) if l.strip().startswith("v_mad_u64_u32"))
        return lines[:i] + lines[i + 1:]
    r = _mutate_and_verify(built, drop)
    assert r.returncode != 0 and "instruction" in (r.stdout + r.stderr)


def test_checker_rejects_a_moved_branch(built):
    def move(lines):
        # send the first conditional branch to another block's label
        labels = [l.split(":")[0].strip() for l in lines if re.match(r"^\.LBB\d+_\d+:", l)]
        i = next(i for i, l in enumerate(lines) if re.match(r"\s+s_cbranch_\w+\s+\.LBB", l))
        op, tgt = lines[i].split()[:2]
        other = next(x for x in labels if x != tgt)
        return lines[:i] + [f"\t{op} {other}"] + lines[i + 1:]
    r = _mutate_and_verify(built, move)
    assert r.returncode != 0


def test_checker_rejects_a_misaligned_object(built):
    # the compiler's own object passes the instruction check but not the
    # alignment check (its loop mixes 4- and 8-byte encodings)
    d, dev_s, out_o, _ = built
    r = _run("python3", ALIGN, "verify", str(out_o) + ".orig.o", str(out_o) + ".orig.o")
    assert r.returncode != 0 and "4 mod 8" in (r.stdout + r.stderr)
