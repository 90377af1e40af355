#!/usr/bin/env python3
"""Generates the exponent loop of the base-2 one-lane kernels as inline assembly.

For a kernel with S radix-2^RB limbs the generated function runs the bits
b_hi .. b_lo of one 32-bit exponent word: per bit, x <- 2^bit * x^2 / R mod N
(R = 2^(RB*S), values lazy in [0, 2N)), exactly the operation of mont_sqr_dbl /
mont_sqr_dbl_top0 in mont.hpp, so the residues do not depend on which path ran.

Why assembly.  At one wavefront per SIMD (a 65,536-candidate launch on 1,024
SIMDs) every instruction takes an issue slot and gfx950 has two costly
patterns that the compiler does not avoid:
  * a scalar instruction issued within a few vector instructions of a
    v_mad_u64_u32 (which writes an SGPR carry-out) waits ~110 cycles; the
    compiled loop computes its bit index and loop condition right after the
    last multiply-add of the squaring;
  * an 8-byte instruction that does not start on an 8-byte boundary costs
    about one cycle more (measured on extracted loop bodies: the 987-bit
    squaring 10,681 cycles with its 4-byte operand shifts leaving the rest of
    the loop on 4 mod 8, 8,585 with them in 8-byte VOP3 form); compiled code
    mixes 4- and 8-byte encodings freely.
The generated loop puts its scalar work after the operand shifts at the top of
each iteration (no multiply-add for at least 8 vector instructions before it),
computes the loop condition there so the closing branch does not wait, and
keeps every instruction 8-byte aligned: vector instructions are 8 bytes (VOP3
forms, or 4-byte forms with a 32-bit literal), the scalar ones come in a pair,
the loop head is 64-byte aligned.  Dependencies between vector instructions
cost nothing on this target (measured), so the column order of the C++
routine is kept and no latency scheduling is needed.

Two forms of the folded doubling:
  * "rdbl": the reduction divides by R / 2^bit (the last reduction digit has
    RB - bit bits and the upper columns are shifted by bit as they join the
    carry: v_lshl_add_u64 with the bit as shift); only xd = 2x is formed.
  * "fold": the operands are shifted (xd = x << (bit + 1) for the cross
    products, x << bit = xd >> 1 formed just before each square product), as
    mont_sqr_dbl does.
rdbl needs fewer instructions; it is used where its column sums provably stay
below 2^64 (checked exactly below for all-ones limbs), then fold, then
sqrshift; a width where none fits keeps the C++ loop, as do widths above MAX_S (the
register allocator's time grows steeply with the operand count of one
statement: 25 s at 36 limbs, 45 s at 46, over 15 minutes at 64; temporaries
in fixed registers instead of operands took longer still).

The script also interprets every instruction list it emits on random inputs
(Python integers) and checks the result against x^2 2^bit R^-1 mod N, the lazy
bound and 64-bit overflow, so a wrong list fails at build time.

Usage: gen_sqr.py OUT.inc S1,S2,...
"""
import math
import random
import struct
import sys
from fractions import Fraction

M64 = (1 << 64) - 1
MAX_S = 48


def radix(S):
    return 29 if S <= 31 else 28


def columns(S, V):
    """Products of each column of x^2 + m*N, as (kind, i, j) triples."""
    cols = []
    for k in range(S + V - 1):
        i0 = 0 if k < V else k - V + 1
        prods = []
        for i in range(i0, k):
            if i < k - i:
                prods.append(('cross', i, k - i))
        if k % 2 == 0 and k // 2 < V:
            prods.append(('square', k // 2, k // 2))
        i1 = k if k < S else S
        for i in range(i0, i1):
            if k - i < V:
                prods.append(('red', i, k - i))
        cols.append(prods)
    return cols


def bound_ok(S, top0, form):
    """Exact worst-case column sums (all limbs 2^RB - 1, bit = 1) stay < 2^64."""
    RB = radix(S)
    V = S - 1 if top0 else S
    L = (1 << RB) - 1
    if form in ('sqrshift', 'mr') and top0:
        return False  # the doubled value needs the top limb; mr keeps all limbs
    if form == 'mr' and RB * S > 1016:
        return False  # the quotient estimate sums the limbs in one double
    xd = (4 * L) if form == 'fold' else (2 * L)   # cross operand
    xs = (2 * L) if form == 'fold' else L         # square operand
    carry = 0
    for k, prods in enumerate(columns(S, V)):
        s = 0
        for kind, i, j in prods:
            s += xd * L if kind == 'cross' else (xs * L if kind == 'square' else L * L)
        if k < S:
            t = s + carry + L * L       # + m_k n_0
            sh = RB - 1 if (form == 'rdbl' and k == S - 1) else RB
        else:
            t = (2 * s if form == 'rdbl' else s) + carry
            sh = RB
        if t > M64:
            return False
        carry = t >> sh
    return True


class Gen:
    """Instruction list of one squaring (columns in order) over named values.

    Names: x{i} (limbs, updated in place), xd{i}/xs{i} (shifted operands),
    n{i}, m{i} (reduction digits), acc (the unshifted sum of an upper column
    of rdbl), c0/c1 (carries), T0/T1 (fixed 64-bit pairs whose low words are
    addressable: column sums accumulate there).
    """

    def __init__(self, S, top0, form):
        self.S, self.top0, self.form = S, top0, form
        self.RB = radix(S)
        self.V = S - 1 if top0 else S
        self.ops = []

    def emit(self, kind, dst, srcs, extra=None):
        self.ops.append((kind, dst, srcs, extra))

    def build(self):
        S, V, RB, form = self.S, self.V, self.RB, self.form
        cols = columns(S, V)
        carry = None
        for k, prods in enumerate(cols):
            T = 'T0' if k % 2 == 0 else 'T1'
            # The carry of the previous column seeds the first multiply-add of
            # the column (a dependent chain costs nothing here), except for the
            # upper columns of rdbl, whose sum is shifted by the bit first.
            shifted = (k >= S and form == 'rdbl')
            acc = 'acc' if shifted else T
            seed = None if (k == 0 or shifted) else carry
            first = True
            for kind, i, j in prods:
                if kind == 'cross':
                    a, b = f'xd{i}', f'x{j}'
                elif kind == 'square':
                    if form in ('rdbl', 'sqrshift', 'mr'):
                        a, b = f'x{i}', f'x{j}'
                    else:  # x_i << bit, formed from xd_i just before its use
                        self.emit('shr1', 'xsq', [f'xd{i}'])
                        a, b = 'xsq', f'x{j}'
                else:
                    a, b = f'm{i}', f'n{j}'
                if first and seed is None:
                    self.emit('mad', acc, [a, b], 'init')
                else:
                    self.emit('mad', acc, [a, b, seed if first else acc])
                first = False
            if k < S:
                last = (k == S - 1 and form == 'rdbl')
                self.emit('mul', f'm{k}', [T + '.lo', 'ninv'])
                self.emit('and', f'm{k}', [f'm{k}'], 'mlast' if last else 'mask')
                self.emit('mad', T, [f'm{k}', 'n0', T])
                carry = f'c{k % 2}'
                self.emit('shr64', carry, [T], 'shl' if last else RB)
            else:
                if shifted:
                    self.emit('lshl_add', T, [acc, carry], 'bit')
                self.emit('and', f'x{k - S}', [T + '.lo'], 'mask')
                if k == S + V - 2:
                    T2 = 'T1' if T == 'T0' else 'T0'
                    self.emit('shr64', T2, [T], RB)
                    self.emit('mov', f'x{V - 1}', [T2 + '.lo'])
                else:
                    carry = f'c{k % 2}'
                    self.emit('shr64', carry, [T], RB)
        if form == 'sqrshift':
            # x <- 2^bit x (lazy: < 4N), limbs from the top down so that each
            # reads its lower neighbour before that one is shifted
            for j in range(V - 1, 0, -1):
                self.emit('shr_var', 'tsh', [f'x{j - 1}', 'shl'])
                self.emit('lshl_or', f'x{j}', [f'x{j}', 'bit', 'tsh'])
                self.emit('and', f'x{j}', [f'x{j}'], 'mask')
            self.emit('shl_var', 'x0', ['x0', 'bit'])
            self.emit('and', 'x0', ['x0'], 'mask')
        if form == 'mr':
            self.mul_small()
        return self.ops

    def mul_small(self):
        """x <- a' x mod N for the integer a' = 1 + bit (a - 1) < 2^8 (mul_small
        in mont.hpp): q = floor(a' x / N) estimated in double precision from
        the limbs (Horner, V - 1 roundings) and invn = (1 - 2^-40) / N, so q is
        the true quotient or one less, then one signed-carry pass forms
        a' x_j - q n_j; the result lies in [0, 2N)."""
        V, RB = self.V, self.RB
        self.emit('cvt_f64', 'h', [f'x{V - 1}'])
        for j in range(V - 2, -1, -1):
            self.emit('cvt_f64', 'dj', [f'x{j}'])
            self.emit('fma_f64', 'h', ['h', 'p2rb', 'dj'])
        self.emit('cvt_f64', 'da', ['ap'])
        self.emit('mul_f64', 'h', ['h', 'da'])
        self.emit('mul_f64', 'h', ['h', 'invn'])
        self.emit('cvt_u32_f64', 'q', ['h'])
        self.emit('neg', 'nq', ['q'])
        carry = None
        for j in range(V):
            if carry is None:
                self.emit('mad_w', 'T0', [f'x{j}', 'ap'], 'init')
            else:
                self.emit('mad_w', 'T0', [f'x{j}', 'ap', carry])
            self.emit('mad_i64', 'T0', [f'n{j}', 'nq', 'T0'])
            self.emit('and', f'x{j}', ['T0.lo'], 'mask')
            if j < V - 1:
                carry = f'c{j % 2}'
                self.emit('ashr64', carry, ['T0'], RB)

    def prologue(self):
        """Per-bit operand preparation (before the scalar loop control)."""
        V = self.V
        ops = []
        if self.form == 'mr':
            ops.append(('mad_u24', 'ap', ['bit', 'am1'], None))
            for i in range(V):
                ops.append(('shl1', f'xd{i}', [f'x{i}'], None))
        elif self.form in ('rdbl', 'sqrshift'):
            ops.append(('sub_rb', 'shl', ['bit'], None))
            if self.form == 'rdbl':
                ops.append(('bfe_mask', 'mlast', ['shl'], None))
            for i in range(V):
                ops.append(('shl1', f'xd{i}', [f'x{i}'], None))
        else:
            ops.append(('add1', 'bit1', ['bit'], None))
            for i in range(V):
                ops.append(('shlbit1', f'xd{i}', [f'x{i}'], None))
        return ops


def _f64(v):
    return struct.unpack('<d', struct.pack('<Q', v))[0] if isinstance(v, int) else v


def _signed(v, bits):
    v &= (1 << bits) - 1
    return v - (1 << bits) if v >> (bits - 1) else v


def interpret(S, top0, form, ops, x, n, ninv, bit, a=2, invn=0.0):
    RB = radix(S)
    V = S - 1 if top0 else S
    mask = (1 << RB) - 1
    r = {'ninv': ninv, 'bit': bit, 'am1': a - 1, 'invn': invn, 'p2rb': float(1 << RB)}
    for i in range(V):
        r[f'x{i}'] = x[i]
        r[f'n{i}'] = n[i]

    def rd(v):
        return r[v[:-3]] & 0xffffffff if v.endswith('.lo') else r[v]
    for kind, dst, srcs, extra in ops:
        s = [rd(v) for v in srcs]
        if kind == 'mad':
            assert s[0] < 2 ** 32 and s[1] < 2 ** 32
            v = s[0] * s[1] + (0 if extra == 'init' else s[2])
            assert v <= M64, 'multiply-add overflow'
        elif kind == 'lshl_add':
            v = (s[0] << (bit if extra == 'bit' else extra)) + s[1]
            assert v <= M64, 'add overflow'
        elif kind == 'shr64':
            v = s[0] >> ((RB - bit) if extra == 'shl' else extra)
        elif kind == 'mul':
            v = (s[0] * s[1]) & 0xffffffff
        elif kind == 'and':
            v = s[0] & ((mask >> bit) if extra == 'mlast' else mask)
        elif kind == 'mov':
            v = s[0] & 0xffffffff
        elif kind == 'sub_rb':
            v = RB - s[0]
        elif kind == 'bfe_mask':
            v = (1 << s[0]) - 1
        elif kind == 'shl1':
            v = s[0] << 1
            assert v < 2 ** 32
        elif kind == 'shlbit1':
            v = s[0] << s[1] if len(s) > 1 else s[0] << (bit + 1)
            assert v < 2 ** 32
        elif kind == 'add1':
            v = s[0] + 1
        elif kind == 'shr_var':
            v = s[0] >> s[1]
        elif kind == 'lshl_or':
            v = ((s[0] << s[1]) | s[2]) & 0xffffffff
        elif kind == 'shl_var':
            v = (s[0] << s[1]) & 0xffffffff
        elif kind == 'shr1':
            v = s[0] >> 1
        elif kind == 'mad_u24':
            assert s[0] < 2 ** 24 and s[1] < 2 ** 24
            v = s[0] * s[1] + 1
        elif kind == 'cvt_f64':
            v = float(s[0])
        elif kind == 'fma_f64':   # one rounding
            v = float(Fraction(s[0]) * Fraction(s[1]) + Fraction(s[2]))
        elif kind == 'mul_f64':
            v = s[0] * s[1]
        elif kind == 'cvt_u32_f64':   # truncation, saturating
            v = 0 if not s[0] > 0 else min(int(s[0]), 0xffffffff)
        elif kind == 'neg':
            v = (-s[0]) & 0xffffffff
        elif kind == 'mad_w':   # unsigned multiply-add modulo 2^64
            v = (s[0] * s[1] + (0 if extra == 'init' else s[2])) & M64
        elif kind == 'mad_i64':   # signed 32 x 32 + signed 64
            v = (_signed(s[0], 32) * _signed(s[1], 32) + _signed(s[2], 64)) & M64
        elif kind == 'ashr64':
            v = (_signed(s[0], 64) >> extra) & M64
        else:
            raise ValueError(kind)
        r[dst] = v
    return [r[f'x{i}'] for i in range(V)]


def scaled_f64(v, L, RB):
    """mont.hpp's scaled(): sum_j ldexp(v_j, RB j - L) in order, as doubles."""
    d = 0.0
    for j, vj in enumerate(v):
        d += math.ldexp(float(vj), RB * j - L)
    return d


def verify(S, top0, form, trials=200, seed=1):
    g = Gen(S, top0, form)
    ops = g.prologue() + g.build()
    RB, V = g.RB, g.V
    mask = (1 << RB) - 1
    rng = random.Random(seed * 1000 + S)
    R = 1 << (RB * S)
    maxbits = RB * (S - 1) - 1 if top0 else RB * S - 4
    for t in range(trials):
        bits = rng.randint(max(2, maxbits - 40), maxbits) if t % 3 else rng.randint(2, maxbits)
        N = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        if N < 3:
            N = 3
        lazy = 4 if form == 'sqrshift' else 2   # values in [0, lazy N)
        x = lazy * N - 1 - t if t < 4 else rng.randrange(0, lazy * N)
        bit = t & 1 if t < 4 else rng.randint(0, 1)
        ninv = (-pow(N, -1, 1 << 32)) & 0xffffffff
        xl = [(x >> (RB * i)) & mask for i in range(V)]
        nl = [(N >> (RB * i)) & mask for i in range(V)]
        a, invn, mult = 2, 0.0, 1 << bit
        if form == 'mr':
            a = rng.choice([3, 5, 7, 11, 13, 53, 251, 255])
            L = N.bit_length()
            invn = math.ldexp(1.0 / scaled_f64(nl, L, RB), -L) * (1.0 - 2.0 ** -40)
            mult = a if bit else 1
        got = interpret(S, top0, form, ops, xl, nl, ninv, bit, a, invn)
        assert all(v <= mask for v in got), f'S={S}: limb range'
        g_val = sum(v << (RB * i) for i, v in enumerate(got))
        assert g_val % N == (x * x * mult * pow(R, -1, N)) % N, f'S={S}: wrong value'
        assert g_val < lazy * N, f'S={S}: lazy bound'
    return ops


def asm_text(S, top0, form, ops):
    RB = radix(S)
    maskhex = '0x%x' % ((1 << RB) - 1)
    fixed = {'T0': 'v[0:1]', 'T1': 'v[2:3]', 'T0.lo': 'v0', 'T1.lo': 'v2'}

    def o(v):
        return fixed.get(v, '%[' + v + ']')
    lines = []
    for kind, dst, srcs, extra in ops:
        s = [o(v) for v in srcs]
        d = o(dst)
        if kind == 'mad':
            c = '0' if extra == 'init' else s[2]
            lines.append(f'v_mad_u64_u32 {d}, %[sc], {s[0]}, {s[1]}, {c}')
        elif kind == 'lshl_add':
            sh = '%[bit]' if extra == 'bit' else str(extra)
            lines.append(f'v_lshl_add_u64 {d}, {s[0]}, {sh}, {s[1]}')
        elif kind == 'shr64':
            sh = o('shl') if extra == 'shl' else str(extra)
            lines.append(f'v_lshrrev_b64 {d}, {sh}, {s[0]}')
        elif kind == 'mul':
            lines.append(f'v_mul_lo_u32 {d}, {s[0]}, {s[1]}')
        elif kind == 'and':
            if extra == 'mlast':   # two VGPR sources: VOP3 encoding
                lines.append(f'v_and_b32_e64 {d}, {o("mlast")}, {s[0]}')
            else:
                lines.append(f'v_and_b32_e32 {d}, {maskhex}, {s[0]}')
        elif kind == 'mov':
            lines.append(f'v_mov_b32_e64 {d}, {s[0]}')
        elif kind == 'sub_rb':
            lines.append(f'v_sub_u32_e64 {d}, {RB}, {s[0]}')
        elif kind == 'bfe_mask':
            lines.append(f'v_bfe_u32 {d}, -1, 0, {s[0]}')
        elif kind == 'shl1':
            lines.append(f'v_lshlrev_b32_e64 {d}, 1, {s[0]}')
        elif kind == 'shlbit1':   # shift amount in a VGPR: VOP3 encoding
            lines.append(f'v_lshlrev_b32_e64 {d}, {o("bit1")}, {s[0]}')
        elif kind == 'add1':
            lines.append(f'v_add_u32_e64 {d}, {s[0]}, 1')
        elif kind == 'shr_var':
            lines.append(f'v_lshrrev_b32_e64 {d}, {s[1]}, {s[0]}')
        elif kind == 'lshl_or':
            lines.append(f'v_lshl_or_b32 {d}, {s[0]}, {s[1]}, {s[2]}')
        elif kind == 'shl_var':
            lines.append(f'v_lshlrev_b32_e64 {d}, {s[1]}, {s[0]}')
        elif kind == 'mad_u24':
            lines.append(f'v_mad_u32_u24 {d}, {s[0]}, {s[1]}, 1')
        elif kind == 'cvt_f64':
            lines.append(f'v_cvt_f64_u32_e64 {d}, {s[0]}')
        elif kind == 'fma_f64':
            lines.append(f'v_fma_f64 {d}, {s[0]}, {s[1]}, {s[2]}')
        elif kind == 'mul_f64':
            lines.append(f'v_mul_f64 {d}, {s[0]}, {s[1]}')
        elif kind == 'cvt_u32_f64':
            lines.append(f'v_cvt_u32_f64_e64 {d}, {s[0]}')
        elif kind == 'neg':
            lines.append(f'v_sub_u32_e64 {d}, 0, {s[0]}')
        elif kind == 'mad_w':
            c = '0' if extra == 'init' else s[2]
            lines.append(f'v_mad_u64_u32 {d}, %[sc], {s[0]}, {s[1]}, {c}')
        elif kind == 'mad_i64':
            lines.append(f'v_mad_i64_i32 {d}, %[sc], {s[0]}, {s[1]}, {s[2]}')
        elif kind == 'ashr64':
            lines.append(f'v_ashrrev_i64 {d}, {extra}, {s[0]}')
        elif kind == 'shr1':
            lines.append(f'v_lshrrev_b32_e64 {d}, 1, {s[0]}')
    return lines


def function_text(S, top0, small_base=False):
    """C++ text of AsmSquarings<S, top0> (base 2) or, with small_base,
    AsmSmallBase<S> (squaring, then the multiply by a small base on set bits),
    or None where no generated form applies."""
    if S > MAX_S or S < 4:  # below 4 limbs (candidates under 113 bits) the C++ loop stays
        return None
    forms = ('mr',) if small_base else ('rdbl', 'fold', 'sqrshift')
    form = next((f for f in forms if bound_ok(S, top0, f)), None)
    if form is None:
        return None
    g = Gen(S, top0, form)
    pro, body = g.prologue(), g.build()
    verify(S, top0, form, trials=100 if small_base else 200)
    V = g.V
    nvalu_before_salu = 1 + len(pro)
    pad = ['v_nop_e64'] * max(0, 8 - nvalu_before_salu - 4)
    wide = ('acc', 'c0', 'c1', 'h', 'dj', 'da')
    inputs = ('ninv', 'bit', 'am1', 'invn', 'p2rb')
    names32, names64 = [], []
    for kind, dst, srcs, extra in pro + body:
        for v in [dst] + list(srcs):
            b = v.split('.')[0]
            if (b.startswith('T') or (b[0] in 'xn' and b[1:].isdigit()) or b in inputs):
                continue
            if b in wide:
                if b not in names64:
                    names64.append(b)
            elif b not in names32:
                names32.append(b)
    lines = (['s_mov_b32 %[sb], %[bhi]', '.p2align 6', '1:',
              'v_bfe_u32 %[bit], %[w], %[sb], 1'] + asm_text(S, top0, form, pro) + pad +
             ['s_sub_u32 %[sb], %[sb], 1', 's_cmp_ge_i32 %[sb], %[blo]'] +
             asm_text(S, top0, form, body) + ['s_cbranch_scc1 1b'])
    outs = [f'[x{i}] "+v"(x[{i}])' for i in range(V)]
    outs += [f'[{v}] "=&v"({v})' for v in names32]
    outs += [f'[{v}] "=&v"({v})' for v in names64]
    outs += ['[bit] "=&v"(bit)', '[sc] "=&s"(sc)', '[sb] "=&s"(sb)']
    ins = [f'[n{i}] "v"(n[{i}])' for i in range(V)] + ['[ninv] "v"(ninv)', '[w] "v"(w)', '[bhi] "s"(b_hi)', '[blo] "s"(b_lo)']
    extra_args = ''
    if small_base:
        ins += ['[am1] "v"(am1)', '[invn] "v"(invn)', '[p2rb] "s"(p2rb)']
        extra_args = ',\n                                             uint32_t am1, double invn'
    body_txt = '\n'.join(f'      "{l}\\n"' for l in lines)
    decl = ''.join(f'    {t} {", ".join(n)};\n' for t, n in (('uint32_t', names32), ('uint64_t', names64)) if n)
    if small_base:
        decl += f'    const double p2rb = {float(1 << g.RB)!r};\n'
    head = (f'struct AsmSmallBase<{S}>' if small_base else
            f'struct AsmSquarings<{S}, {"true" if top0 else "false"}>')
    what = 'small base' if small_base else ('top limb zero' if top0 else 'all limbs')
    return f"""// S = {S}, {what}, form {form}: {len(body)} instructions per bit
template <>
{head} {{
  static constexpr bool kAvailable = true;
  __device__ __forceinline__ static void run(uint32_t (&x)[{S}], const uint32_t (&n)[{S}], uint32_t ninv,
                                             uint32_t w, int b_hi, int b_lo{extra_args}) {{
{decl}    uint32_t bit, sb;
    uint64_t sc;
    asm volatile(
{body_txt}
        : {', '.join(outs)}
        : {', '.join(ins)}
        : "v0", "v1", "v2", "v3", "scc");
    (void)sc;
    (void)bit;
    (void)sb;
  }}
}};
"""


def main():
    out, widths = sys.argv[1], [int(s) for s in sys.argv[2].split(',') if s]
    parts = ['// Generated by gen_sqr.py; do not edit.', '#pragma once', '']
    for S in widths:
        for top0, small in ((False, False), (True, False), (False, True)):
            t = function_text(S, top0, small)
            if t:
                parts.append(t)
    text = '\n'.join(parts) + '\n'
    try:
        with open(out) as f:
            if f.read() == text:  # unchanged: keep the timestamp (no rebuild)
                return
    except OSError:
        pass
    with open(out, 'w') as f:
        f.write(text)


if __name__ == '__main__':
    main()
