"""Build step: re-encode the compiler's gfx950 assembly so that every 8-byte
instruction starts on an 8-byte boundary, then check the result against the
compiler's own object.

On MI355X an 8-byte instruction (VOP3, DPP, memory, or a 4-byte form with a
32-bit literal) that does not start on an 8-byte boundary issues about one
cycle late.  Compiled code mixes 4- and 8-byte encodings freely, so in a
multiply-add loop roughly half of the 8-byte instructions sit at 4 mod 8.
This pass, over the device assembly of one translation unit, puts
`.p2align 3` before every basic-block label (.LBB*, the targets of branches;
other labels, such as the PC-relative anchors of long branches, stay where
they are) and, before an 8-byte instruction that would start at 4 mod 8,
widens the last 4-byte instruction since the previous 8-byte one that has an
8-byte form doing the same thing (VOP1/VOP2/VOPC without a literal -> VOP3;
`s_nop N` -> `s_nop N-1; s_nop 0`), or else inserts `s_nop 0`.  It only adds
wait states, never removes one.  A function whose aligned form does not
assemble (a short branch pushed past its 16-bit range) stays as compiled.

    align_isa.py build compiler.s aligned.o
    align_isa.py verify compiler.o aligned.o

`verify` disassembles both objects and checks, function by function, that
after dropping `s_nop` and the _e32/_e64 suffix they hold the same
instructions in the same order, that every branch lands on the same
instruction, and that in front of every instruction there are at least the
wait states the compiler's `s_nop`s put there (counted in its assembly when
given, since alignment fill looks the same in an object); and that no
8-byte instruction of the aligned object sits at 4 mod 8.
It exits non-zero otherwise, which fails the build."""
import re
import subprocess
import sys

OBJDUMP = '/opt/rocm/llvm/bin/llvm-objdump'
CLANG = '/opt/rocm/llvm/bin/clang'
ARCH = 'gfx950'

# 4-byte SOPK / SOPP forms (their immediates are part of the 4 bytes)
_SOP4 = ('s_movk', 's_cmpk', 's_addk', 's_mulk', 's_cmovk', 's_setreg', 's_getreg',
         's_cbranch', 's_branch', 's_nop', 's_waitcnt', 's_endpgm', 's_barrier', 's_setprio',
         's_sleep', 's_trap', 's_icache_inv', 's_sethalt', 's_ttracedata', 's_set_gpr_idx_mode',
         's_set_gpr_idx_off')
_SMEM = ('s_load', 's_buffer', 's_store', 's_dcache', 's_memtime', 's_memrealtime', 's_atomic',
         's_scratch')
# 4-byte VOP1 forms the disassembler prints without a suffix
_VOP1_BARE = ('v_readfirstlane_b32', 'v_accvgpr_mov_b32', 'v_nop')
# VOP2 forms whose encoding always carries a literal
_NO_E64 = ('v_madmk', 'v_madak', 'v_fmamk', 'v_fmaak')


def _is_literal(tok):
    """True for an integer operand outside the inline constants (-16..64),
    or an expression (label arithmetic), which the assembler encodes as a
    32-bit literal."""
    t = tok.strip()
    if '.L' in t or '(' in t:
        return True
    m = re.fullmatch(r'(-?)(0x[0-9a-fA-F]+|\d+)', t)
    if not m:
        return False  # registers, inline floats, modifiers
    v = int(m.group(2), 0) * (-1 if m.group(1) else 1)
    return not (-16 <= v <= 64)


def _operands(ins):
    parts = ins.split(None, 1)
    return [] if len(parts) < 2 else parts[1].split(',')


def _size(ins):
    op = ins.split()[0]
    ops = _operands(ins)
    if op.startswith('s_'):
        if op.startswith(_SMEM):
            return 8
        if op.startswith(_SOP4):
            return 4
        return 8 if any(_is_literal(o) for o in ops) else 4
    if op.endswith('_e32') or op in _VOP1_BARE:
        return 8 if any(_is_literal(o) for o in ops) else 4
    return 8


def _convertible(ins):
    """A 4-byte instruction that has an 8-byte form doing the same thing:
    VOP1/VOP2/VOPC without a literal (VOP3), or s_nop N >= 1 (s_nop N-1 and
    s_nop 0: the same N+1 wait states)."""
    op = ins.split()[0]
    if op == 's_nop':
        return int(_operands(ins)[0], 0) >= 1
    return (op.startswith('v_') and op.endswith('_e32') and not op.startswith(_NO_E64) and
            not any(_is_literal(o) for o in _operands(ins)))


def _widen(ins):
    op = ins.split()[0]
    if op == 's_nop':
        return [f's_nop {int(_operands(ins)[0], 0) - 1}', 's_nop 0']
    return [op[:-4] + '_e64' + ins[len(op):]]


def transform(src, dst, skip=()):
    """Aligns every function of src except those named in skip.  Within a run
    of 4-byte instructions before an 8-byte one at 4 mod 8, the last
    convertible one that can take the extra 4 bytes is widened (no extra
    instruction); failing that an s_nop 0 goes in front.  Returns counts and the line ranges of the
    functions, for mapping assembler errors back."""
    lines = open(src).read().split('\n')
    funcs = {m.group(1) for m in (re.match(r'\s*\.type\s+([^,\s]+)\s*,\s*@function', l)
                                  for l in lines) if m}
    out = []          # list of output lines; instructions as [text] lists
    spans = []        # (name, first out line, last out line)
    cur = None
    off = 0
    run = []          # (index in out, offset, is s_nop) of the run's convertible 4-byte instructions
    stats = {'widened': 0, 'pad': 0, 'label_align': 0}
    for line in lines:
        t = line.split(';')[0].strip()  # the text before any comment
        if t.endswith(':') and t[:-1] in funcs:
            cur = t[:-1] if t[:-1] not in skip else None
            spans.append([t[:-1], len(out), None])
            off, run = 0, []
            out.append(line)
            continue
        if t.startswith('.Lfunc_end'):
            if spans and spans[-1][2] is None:
                spans[-1][2] = len(out)
            cur = None
        if cur is None or not t or t.startswith('//'):
            out.append(line)
            continue
        if t.startswith('.p2align'):
            if int(t.split()[1].split(',')[0]) >= 3:
                off, run = 0, []
            out.append(line)
            continue
        if t.endswith(':'):
            if t.startswith('.LBB'):
                out.append('\t.p2align 3')
                stats['label_align'] += 1
                off, run = 0, []
            out.append(line)
            continue
        if t.startswith('.'):
            out.append(line)
            continue
        n = _size(t)
        if n == 8:
            if off % 8 == 4:
                # a VOP form widened where it starts at 0 mod 8 becomes an
                # aligned 8-byte instruction; an s_nop splits into two
                # 4-byte ones wherever it is
                k = next((i for i, o, nop in reversed(run) if nop or o % 8 == 0), None)
                if k is not None:
                    out[k] = '\n'.join('\t' + x for x in _widen(out[k].strip()))
                    stats['widened'] += 1
                else:
                    out.append('\ts_nop 0')
                    stats['pad'] += 1
                off += 4
            run = []
        elif _convertible(t):
            run.append((len(out), off, t.startswith('s_nop')))
        out.append('\t' + t)
        off += n
    open(dst, 'w').write('\n'.join(out))
    return stats, spans


def _assemble(src, obj):
    r = subprocess.run([CLANG, '-target', 'amdgcn-amd-amdhsa', f'-mcpu={ARCH}', '-c', src, '-o', obj],
                       capture_output=True, text=True)
    return r.returncode, r.stderr


def build(dev_s, out_o):
    """Aligned object of dev_s in out_o, checked against the compiler's.  A
    function whose aligned form does not assemble (a short branch pushed out
    of its 16-bit range by the added bytes) is kept as the compiler wrote it."""
    orig_o = out_o + '.orig.o'
    rc, err = _assemble(dev_s, orig_o)
    if rc:
        raise SystemExit(f'align_isa: the compiler\'s assembly does not assemble:\n{err[:2000]}')
    al_s = out_o + '.al.s'
    skip = set()
    for _ in range(8):
        stats, spans = transform(dev_s, al_s, skip)
        rc, err = _assemble(al_s, out_o)
        if rc == 0:
            break
        bad = set()
        # every function of which the assembler reported a line (the output
        # keeps the input's line structure up to the inserted lines, so map
        # through the spans of the aligned file)
        for m in re.finditer(r':(\d+):\d+: error', err):
            ln = int(m.group(1)) - 1
            out_lines = open(al_s).read().split('\n')
            names = [sp[0] for sp in _spans_of(out_lines) if sp[1] <= ln <= sp[2]]
            if not names:
                raise SystemExit(f'align_isa: assembler error outside a function:\n{err[:2000]}')
            bad.update(names)
        if not bad - skip:
            raise SystemExit(f'align_isa: cannot assemble:\n{err[:2000]}')
        skip |= bad
    else:
        raise SystemExit('align_isa: gave up after 8 passes')
    n = verify(orig_o, out_o, skip, dev_s)
    return stats, n, skip


def _spans_of(lines):
    funcs = {m.group(1) for m in (re.match(r'\s*\.type\s+([^,\s]+)\s*,\s*@function', l)
                                  for l in lines) if m}
    spans = []
    for i, line in enumerate(lines):
        t = line.split(';')[0].strip()
        if t.endswith(':') and t[:-1] in funcs:
            spans.append([t[:-1], i, len(lines)])
        elif t.startswith('.Lfunc_end') and spans and spans[-1][2] == len(lines):
            spans[-1][2] = i
    return spans


_INS = re.compile(r'^\s+(\S.*?)\s*//\s*([0-9A-Fa-f]+):\s*((?:[0-9A-Fa-f]{8}\s*)+)(?:<(\S+)\+0x([0-9a-fA-F]+)>)?\s*$')
_FN = re.compile(r'^([0-9a-fA-F]{16}) <(\S+)>:')


def _disasm(obj):
    txt = subprocess.run([OBJDUMP, '-d', '--mcpu=gfx950', obj], check=True,
                         capture_output=True, text=True).stdout
    fns, cur = {}, None
    for line in txt.split('\n'):
        m = _FN.match(line)
        if m:
            cur = fns.setdefault(m.group(2), {'base': int(m.group(1), 16), 'ins': []})
            continue
        m = _INS.match(line)
        if m and cur is not None:
            text = re.sub(r'\s+', ' ', m.group(1).strip())
            addr = int(m.group(2), 16)
            nbytes = 4 * len(m.group(3).split())
            target = int(m.group(5), 16) if m.group(5) else None
            cur['ins'].append((addr, nbytes, text, target))
    return fns


def _canon(fn):
    """Non-nop instructions with suffixes dropped, branch targets as indices,
    and the wait states (s_nop N counts N+1) in front of each."""
    base = fn['base']
    seq, waits, addr_to_idx = [], [], {}
    pending = 0
    for addr, n, text, target in fn['ins']:
        op = text.split()[0]
        if op == 's_nop':
            pending += int(text.split()[1], 0) + 1
            continue
        addr_to_idx[addr - base] = len(seq)
        seq.append((re.sub(r'_e(32|64)\b', '', op), text.split(None, 1)[1:] if target is None else None,
                    target))
        waits.append(pending)
        pending = 0
    sorted_addrs = sorted(addr_to_idx)

    def index_at(off):  # first real instruction at or after off (nops fall through)
        for a in sorted_addrs:
            if a >= off:
                return addr_to_idx[a]
        return len(seq)
    # long branches: s_getpc_b64 s[k:k+1] leaves the address of the next
    # instruction, s_add_u32 / s_addc_u32 add a PC-relative offset, s_setpc
    # jumps; the offset becomes the index of the instruction it reaches
    real = [(addr, text) for addr, n, text, target in fn['ins'] if text.split()[0] != 's_nop']
    pc = {}
    canon = []
    for (op, operands, target), (addr, text) in zip(seq, real):
        ops = operands[0] if operands else ''
        if target is not None:
            canon.append((op, ('->', index_at(target))))
            continue
        m = re.fullmatch(r's\[(\d+):\d+\]', ops.strip()) if op == 's_getpc_b64' else None
        if m:
            pc[int(m.group(1))] = addr + 4 - base
        m = re.fullmatch(r's(\d+), s(\d+), (0x[0-9a-fA-F]+|-?\d+)', ops.strip())
        if m and op == 's_add_u32' and int(m.group(1)) in pc:
            lit = int(m.group(3), 0) & 0xFFFFFFFF
            lit -= (1 << 32) if lit >> 31 else 0
            canon.append((op, ('pc->', index_at(pc[int(m.group(1))] + lit))))
            continue
        if m and op == 's_addc_u32' and int(m.group(1)) - 1 in pc:
            canon.append((op, ('pc-hi', m.group(1))))
            continue
        canon.append((op, ops))
    return canon, waits


def _source_waits(src):
    """Per function of the compiler's assembly, the wait states its explicit
    `s_nop`s put in front of each other instruction.  In the object the fill
    of an alignment directive (`s_nop 0`s) looks the same, and its amount
    depends on the code before it, so the required wait states come from the
    source."""
    lines = open(src).read().split('\n')
    funcs = {m.group(1) for m in (re.match(r'\s*\.type\s+([^,\s]+)\s*,\s*@function', l)
                                  for l in lines) if m}
    out, cur, pending = {}, None, 0
    for line in lines:
        t = line.split(';')[0].strip()
        if t.endswith(':') and t[:-1] in funcs:
            cur, pending = out.setdefault(t[:-1], []), 0
            continue
        if t.startswith('.Lfunc_end'):
            cur = None
        if cur is None or not t or t.startswith(('.', '//')) or t.endswith(':'):
            continue
        if t.split()[0] == 's_nop':
            pending += int(_operands(t)[0], 0) + 1
            continue
        cur.append(pending)
        pending = 0
    return out


def verify(orig, new, skip=(), src=None):
    a, b = _disasm(orig), _disasm(new)
    if set(a) != set(b):
        raise SystemExit(f'align_isa: functions differ: {sorted(set(a) ^ set(b))[:5]}')
    need = _source_waits(src) if src else {}
    misaligned = 0
    for name in a:
        ca, wa = _canon(a[name])
        cb, wb = _canon(b[name])
        if len(ca) != len(cb):
            raise SystemExit(f'align_isa: {name}: {len(ca)} vs {len(cb)} instructions')
        req = need.get(name)
        if req is not None and len(req) != len(ca):
            raise SystemExit(f'align_isa: {name}: {len(req)} instructions in the source, '
                             f'{len(ca)} in the object')
        for i, (x, y) in enumerate(zip(ca, cb)):
            if x != y:
                raise SystemExit(f'align_isa: {name}: instruction {i}: {x} vs {y}')
            w = req[i] if req is not None else wa[i]
            if wb[i] < w:
                raise SystemExit(f'align_isa: {name}: instruction {i}: wait states {w} -> {wb[i]}')
        if name in skip:
            continue
        base = b[name]['base']
        misaligned += sum(1 for addr, n, _, _ in b[name]['ins'] if n == 8 and (addr - base) % 8)
    if misaligned:
        raise SystemExit(f'align_isa: {misaligned} 8-byte instructions still at 4 mod 8')
    return len(a)


def main():
    if len(sys.argv) == 4 and sys.argv[1] == 'build':
        stats, n, skip = build(sys.argv[2], sys.argv[3])
        print('align_isa: ' + ', '.join(f'{k} {v}' for k, v in stats.items()) +
              f'; {n} functions match the compiler\'s, {len(skip)} kept as compiled' +
              (f' ({", ".join(sorted(skip))})' if skip else ''))
    elif len(sys.argv) == 4 and sys.argv[1] == 'verify':
        print(f'align_isa: {verify(sys.argv[2], sys.argv[3])} functions match')
    else:
        raise SystemExit(__doc__)


if __name__ == '__main__':
    main()
