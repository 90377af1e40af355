"""CPU tests of the generator of the assembly exponent loop (csrc/gen_sqr.py).

Every instruction list the generator emits is interpreted on Python integers
and compared with x^2 2^bit R^-1 mod N (Montgomery squaring with the folded
base-2 step, mont.hpp's mont_sqr_dbl), the lazy bound x < 2N and the 64-bit
range of every accumulator.  The GPU parity tests run the assembly itself."""
import importlib.util
import os
import random

import pytest

from conftest import PKG

_spec = importlib.util.spec_from_file_location("gen_sqr", os.path.join(PKG, "csrc", "gen_sqr.py"))
gen = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(gen)

# kernel limb counts the one-lane kernels instantiate (prime-gap-gpu_amd/Makefile)
WIDTHS = list(range(1, 49)) + [56, 64, 74]


def _form(S, top0):
    return next((f for f in ("rdbl", "fold", "sqrshift") if gen.bound_ok(S, top0, f)), None)


@pytest.mark.parametrize("S", [S for S in WIDTHS if 4 <= S <= gen.MAX_S])
def test_every_generated_width_interprets_correctly(S):
    for top0 in (False, True):
        form = _form(S, top0)
        if form is None:
            continue
        gen.verify(S, top0, form, trials=60, seed=7)


@pytest.mark.parametrize("S", [4, 5, 9, 13, 17, 20, 25, 31, 32, 36])
def test_small_base_form_interprets_correctly(S):
    # squaring, then a' x mod N with the quotient estimated in double precision
    # (interpreted with correctly rounded fused multiply-adds and products)
    assert gen.bound_ok(S, False, "mr") and not gen.bound_ok(S, True, "mr")
    gen.verify(S, False, "mr", trials=60, seed=11)


def test_small_base_form_stops_where_one_double_cannot_hold_the_limbs():
    assert gen.bound_ok(36, False, "mr") and not gen.bound_ok(37, False, "mr")


def test_forms_match_the_column_bounds_of_the_cpp_loop():
    # the C++ loop folds the doubling for RB = 28 and for RB = 29 up to 21
    # limbs (static_assert in mont.hpp); there a folded form is generated
    for S in WIDTHS:
        if 4 <= S <= gen.MAX_S and (S >= 32 or S <= 21):
            assert _form(S, False) in ("rdbl", "fold"), S
    # radix 2^29 beyond 21 limbs fits neither folded form with all limbs in
    # use: squaring then doubling, as the C++ loop's mont_sqr + dbl_lazy
    for S in range(22, 32):
        assert not gen.bound_ok(S, False, "fold") and _form(S, False) == "sqrshift", S
        assert not gen.bound_ok(S, True, "sqrshift")


def test_bound_check_is_exact_at_the_edge():
    # rdbl at 17 limbs (radix 2^29) holds by a margin below 2^36; at 18 limbs
    # its first upper column exceeds 2^64 for all-ones limbs
    assert gen.bound_ok(17, False, "rdbl")
    assert not gen.bound_ok(18, False, "rdbl")
    L = (1 << 29) - 1
    # first upper column of the 18-limb squaring: 8 cross products (doubled
    # operands), 17 reduction products, the whole column doubled
    col = 8 * (2 * L) * L + 17 * L * L
    assert 2 * col > gen.M64


def test_interpreter_catches_a_dropped_product():
    S, top0 = 17, True
    g = gen.Gen(S, top0, "rdbl")
    ops = g.prologue() + g.build()
    # a cross product in the middle of a column's chain (its accumulator is
    # read and written by the same instruction)
    idx = next(i for i, op in enumerate(ops)
               if op[0] == "mad" and len(op[2]) == 3 and op[2][2] == op[1] and op[2][0].startswith("xd"))
    broken = ops[:idx] + ops[idx + 1:]
    rng = random.Random(3)
    RB, V = g.RB, g.V
    mask = (1 << RB) - 1
    N = rng.getrandbits(460) | (1 << 459) | 1
    x = rng.randrange(0, 2 * N)
    ninv = (-pow(N, -1, 1 << 32)) & 0xFFFFFFFF
    xl = [(x >> (RB * i)) & mask for i in range(V)]
    nl = [(N >> (RB * i)) & mask for i in range(V)]
    R = 1 << (RB * S)
    want = (x * x * pow(R, -1, N)) % N
    good = gen.interpret(S, top0, "rdbl", ops, xl, nl, ninv, 0)
    assert sum(v << (RB * i) for i, v in enumerate(good)) % N == want
    bad = gen.interpret(S, top0, "rdbl", broken, xl, nl, ninv, 0)
    assert sum(v << (RB * i) for i, v in enumerate(bad)) % N != want


def _size(line):
    """Encoded bytes of an instruction the generator emits (gfx950)."""
    op = line.split(" ", 1)[0]
    if op.startswith("s_"):
        return 4
    if op.endswith("_e32"):
        srcs = line.split(",")[1:]
        lit = any(x.strip().startswith("0x") for x in srcs)
        return 8 if lit else 4
    return 8  # VOP3 / VOP3b


@pytest.mark.parametrize("S,top0,small", [(17, True, False), (17, False, False), (20, False, False),
                                          (25, False, False), (36, True, False), (17, False, True)])
def test_generated_loop_is_aligned_with_scalar_work_away_from_the_mads(S, top0, small):
    text = gen.function_text(S, top0, small)
    body = [l.strip().strip('"').replace("\\n", "") for l in text.splitlines() if l.strip().startswith('"')]
    assert sum(1 for l in body if l.startswith("s_cbranch_scc1 1b")) == 1
    assert body[body.index("1:") - 1] == ".p2align 6"
    # every 8-byte instruction of the loop starts on an 8-byte boundary
    off = 0
    for l in body[body.index("1:") + 1:]:
        n = _size(l)
        if n == 8:
            assert off % 8 == 0, l
        off += n
    # the scalar loop control comes after at least 8 vector instructions that
    # follow the previous iteration's last SGPR-writing multiply-add
    start = body.index("1:")
    salu = next(i for i in range(start, len(body)) if body[i].startswith("s_sub_u32"))
    last_mad = max(i for i, l in enumerate(body) if l.startswith(("v_mad_u64_u32", "v_mad_i64_i32")))
    tail = len(body) - 1 - last_mad - 1  # vector instructions after the last mad, before the branch
    assert tail + (salu - start - 1) >= 8
