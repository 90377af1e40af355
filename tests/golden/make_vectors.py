"""Writes tests/golden/vectors.json: golden vectors for the PRP path.

The reference holds no test vectors of its own (it has no test suite), so the
expected values come from implementations independent of both the product and
the C oracle:
  * residues a^(N-1) mod N: CPython's built-in ``pow`` (its own bignum code);
  * primality: ``sympy.isprime``;
  * strong / reference verdicts: the chain of miller_rabin.h:216-248 written in
    Python on top of ``pow``;
  * when /opt/conda/bin/python3.9 has gmpy2 (GMP's own mpz_powm and
    is_strong_prp / is_fermat_prp), every vector is cross-checked there too and
    the script fails on any disagreement.
Candidates of the sieved sets are generated here by a Python restatement of the
generator (xoshiro256** multipliers, sieve of m*K + X by primes < 10^6).

Run:  python3 tests/golden/make_vectors.py
"""
from __future__ import annotations

import json
import os
import random
import subprocess
import sys

import sympy

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "vectors.json")


def t_of(N: int) -> int:
    nm1 = N - 1
    return (nm1 & -nm1).bit_length() - 1


def ref_verdict(N: int, a: int) -> bool:
    """miller_rabin.h:221-245 as written (t squarings)."""
    nm1 = N - 1
    t = t_of(N)
    x = pow(a, nm1 >> t, N)
    if x == 1 or x == nm1:
        return True
    if t == 1:
        return False
    count = t
    while True:
        x = x * x % N
        if x == nm1:
            return True
        count -= 1
        if count == 0 or x == 1:
            return False


def sprp(N: int, a: int) -> bool:
    nm1 = N - 1
    t = t_of(N)
    x = pow(a, nm1 >> t, N)
    if x in (1, nm1):
        return True
    for _ in range(t - 1):
        x = x * x % N
        if x == nm1:
            return True
        if x == 1:
            return False
    return False


def record(N: int) -> dict:
    return {
        "n": hex(N),
        "res2": hex(pow(2, N - 1, N)),
        "fermat2": pow(2, N - 1, N) == 1,
        "sprp": {str(a): sprp(N, a) for a in (2, 3, 5, 7)},
        "ref": {str(a): ref_verdict(N, a) for a in (2, 3, 5, 7)},
        "prime": bool(sympy.isprime(N)),
    }


# ---- generator restatement (xoshiro256** + sieve) --------------------------

M64 = (1 << 64) - 1


def splitmix(state: list[int]) -> int:
    state[0] = (state[0] + 0x9E3779B97F4A7C15) & M64
    z = state[0]
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


class Xoshiro:
    def __init__(self, seed: int):
        st = [seed & M64]
        self.s = [splitmix(st) for _ in range(4)]

    @staticmethod
    def rotl(v, k):
        return ((v << k) | (v >> (64 - k))) & M64

    def next(self) -> int:
        s = self.s
        r = (self.rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self.rotl(s[3], 45)
        return r

    def below(self, r: int) -> int:
        reject = ((1 << 64) - r) % r
        while True:
            v = self.next()
            if v >= reject:
                return v % r


def multipliers(seed, m_lo, m_hi, d, n_m):
    from math import gcd
    g = Xoshiro(seed)
    out: list[int] = []
    while len(out) < n_m:
        m = m_lo + g.below(m_hi - m_lo)
        if gcd(m, d) != 1 or m in out:
            continue
        out.append(m)
    return sorted(out)


def sieved(P, d, ms, x_max, per_m, limit=1_000_000):
    K = sympy.primorial(P, nth=False) // d
    assert K * d == sympy.primorial(P, nth=False)
    primes = list(sympy.primerange(2, limit))
    out = []
    for mi, m in enumerate(ms):
        mK = m * K
        comp = bytearray(x_max + 1)
        for p in primes:
            x = (-mK) % p
            if x == 0:
                x = p
            comp[x::p] = b"\x01" * len(range(x, x_max + 1, p))
        got = 0
        for x in range(1, x_max + 1):
            if comp[x] or got >= per_m:
                continue
            out.append((mi, x, mK + x))
            got += 1
    return out


# ---- gmpy2 cross-check ------------------------------------------------------

GMPY_CHECK = r"""
import json, sys, gmpy2
vec = json.load(sys.stdin)
bad = 0
for rec in vec:
    N = int(rec["n"], 16)
    if int(gmpy2.powmod(2, N - 1, N)) != int(rec["res2"], 16): bad += 1
    if N > 3:
        for a in ("2", "3", "5", "7"):
            if int(a) % N and bool(gmpy2.is_strong_prp(N, int(a))) != rec["sprp"][a]: bad += 1
print(bad)
"""


def gmpy_cross_check(records) -> str:
    py39 = "/opt/conda/bin/python3.9"
    if not os.path.exists(py39):
        return "skipped (no /opt/conda/bin/python3.9)"
    r = subprocess.run([py39, "-c", GMPY_CHECK], input=json.dumps(records), text=True,
                       capture_output=True)
    if r.returncode != 0:
        return "skipped (gmpy2 unavailable: %s)" % r.stderr.strip().splitlines()[-1:]
    bad = int(r.stdout.strip())
    if bad:
        sys.exit(f"gmpy2 disagrees on {bad} checks")
    return "gmpy2 %s agrees on all records" % "2.0.8"


def main():
    rnd = random.Random(0x5EED)
    kat_n = [
        # primes
        3, 5, 7, 31, 8191, 131071, 524287, 2**31 - 1, 2**61 - 1, 2**89 - 1, 2**107 - 1,
        2**127 - 1, 2**521 - 1, 2**607 - 1, 2**1279 - 1,
        # base-2 Fermat pseudoprimes / Carmichael numbers
        341, 561, 645, 1105, 1387, 1729, 1905, 2465, 2701, 2821, 4369, 4371, 5461, 6601,
        7957, 8321, 10585, 15841, 29341, 41041, 46657, 52633, 62745, 63973, 75361,
        # strong pseudoprimes: base 2; bases 2,3; 2,3,5; 2,3,5,7
        2047, 3277, 4033, 4681, 1373653, 25326001, 3215031751, 2152302898747,
        # small composites and N - 1 = power of two shapes
        9, 15, 21, 25, 27, 45, 91, 4097, 65537, 2**32 + 1, 2**64 + 1, 2**128 + 1,
        # edge widths: 2^(32s-2) - 1 shapes (largest N each kernel limb count holds)
    ] + [2**(32 * s - 2) - 1 for s in range(1, 9)] + [2**(32 * s - 3) + 1 for s in range(1, 9)]
    kat = [record(N) for N in kat_n]

    tiny = [hex(pow(2, N - 1, N)) for N in range(3, 4096, 2)]

    widths = []
    for s in list(range(1, 41)) + [48, 56, 64]:
        lo = 32 * s - 33 if s > 1 else 2
        for bits in sorted({max(2, lo + 1), rnd.randint(max(2, lo + 1), 32 * s - 2), 32 * s - 2}):
            N = rnd.getrandbits(bits) | 1 | (1 << (bits - 1))
            widths.append(record(N))

    sieve_sets = {}
    for name, (P, d, seed, x_max, per_m, n_m, limbs) in {
        "p337_d2310": (337, 2310, 0x5EED0001, 4000, 40, 2, 16),
        "p907_d30030": (907, 30030, 0x5EED0003, 3000, 12, 2, 64),
    }.items():
        ms = multipliers(seed, 10**7, 10**7 + 10**4, d, n_m)
        rows = sieved(P, d, ms, x_max, per_m)
        sieve_sets[name] = {
            "P": P, "d": d, "seed": seed, "x_max": x_max, "per_m": per_m, "limbs": limbs,
            "m": ms,
            "cands": [dict(record(N), m_index=mi, x=x) for mi, x, N in rows],
        }

    xo = Xoshiro(0x5EED0001)
    prng = {"seed": 0x5EED0001, "first": [hex(xo.next()) for _ in range(8)]}

    allrec = kat + widths + [c for s in sieve_sets.values() for c in s["cands"]]
    check = gmpy_cross_check(allrec)
    vec = {
        "about": "golden vectors for the PRP path; expected values from CPython pow, "
                 "sympy.isprime and a Python restatement of miller_rabin.h:216-248",
        "gmpy2_cross_check": check,
        "prng": prng,
        "kat": kat,
        "tiny_res2": {"first": 3, "step": 2, "res": tiny},
        "widths": widths,
        "sieved": sieve_sets,
    }
    with open(OUT, "w") as f:
        json.dump(vec, f, indent=0)
    print(f"wrote {OUT}: {len(kat)} kat, {len(tiny)} tiny, {len(widths)} widths, "
          f"{sum(len(s['cands']) for s in sieve_sets.values())} sieved; {check}")


if __name__ == "__main__":
    main()
