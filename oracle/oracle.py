"""CPU oracle for the batched PRP path.  TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker or as the timed CPU baseline; the product never imports
it.  Wraps liboracle.so (oracle.c, GMP) and adds pure-Python restatements for
small inputs (CPython's own bignum ``pow``, independent of GMP).

Reference functions followed (file:line in sethtroisi/prime-gap-gpu):
  ref_verdict      miller_rabin.h:216-248  (strong test as written, incl. t squarings)
  sprp             textbook strong test (the reference's intent)
  residue          the CPU baseline path: mpz_powm(2, N-1, N)
  gen_*            gap_common.cpp:60-65 (K = P#/d), gpu_testing.cu:171-172 (N = m*K + x),
                   gap_search_gpu.cpp:619-689 (sieve of m*K + X by small primes)
  next_prime_offset overflow.cpp:142-160 (mpz_nextprime / previous prime around m*K)
  ms_*             the m-interval search step (msieve.c): is_coprime_and_valid_m,
                   coprime_X, run_sieve_thread's composites bitmap, finalize +
                   fill_batch + process_finished_batch + remove_vector
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run make in {_HERE})")
    lib = ctypes.CDLL(LIB_PATH)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    lib.orc_verdicts.argtypes = [u32p, ctypes.c_size_t, ctypes.c_uint32, u32p, ctypes.c_uint32,
                                 ctypes.c_uint32, u32p, u32p]
    lib.orc_probab_prime.argtypes = [u32p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int32)]
    lib.orc_xoshiro_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint32, u64p]
    lib.orc_splitmix_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint32, u64p]
    lib.orc_xoshiro_from_state.argtypes = [u64p, ctypes.c_uint32, u64p]
    lib.orc_gen_multipliers.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                        ctypes.c_uint32, ctypes.c_uint32, u64p]
    lib.orc_gen_sieved.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u64p, ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_uint32, u32p, u32p, u32p,
                                   ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_gen_sieved_side.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, u64p,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, u32p,
                                        ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_small_factor.argtypes = [u32p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                     u32p]
    lib.orc_time_powm2.argtypes = [u32p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_int,
                                   ctypes.c_int, u32p]
    lib.orc_time_powm2.restype = ctypes.c_double
    lib.orc_next_prime_offset.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                          ctypes.c_uint32, u32p]
    lib.orc_build_mk.argtypes = [u32p, ctypes.c_uint32, u64p, ctypes.c_size_t, ctypes.c_uint32,
                                 ctypes.c_int, ctypes.c_uint32, u32p,
                                 ctypes.POINTER(ctypes.c_double)]
    szp = ctypes.POINTER(ctypes.c_size_t)
    lib.orc_ms_active.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                  ctypes.c_uint32, u32p, szp]
    lib.orc_ms_coprime_x.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u32p, ctypes.c_size_t, szp]
    lib.orc_ms_sieve.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                 ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64,
                                 ctypes.c_uint32, ctypes.c_int, u32p,
                                 ctypes.POINTER(ctypes.c_double)]
    lib.orc_ms_step.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32, u32p,
                                u32p, szp, u32p, u32p, ctypes.c_uint32, u32p, szp, u32p, szp]
    return lib


lib = _load()

MODE_REFERENCE, MODE_STRONG, MODE_FERMAT = 0, 1, 2


def _u32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _u64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def verdicts(limbs: np.ndarray, mode: int, bases=(2,), residues: bool = False):
    a = np.ascontiguousarray(limbs, dtype=np.uint32)
    n, S = a.shape
    b = np.asarray(bases, dtype=np.uint32)
    out = np.zeros(n, dtype=np.uint32)
    res = np.zeros((n, S), dtype=np.uint32) if residues else None
    rc = lib.orc_verdicts(_u32(a), n, S, _u32(b), len(b), mode, _u32(out),
                          _u32(res) if residues else None)
    if rc != 0:
        raise ValueError(f"orc_verdicts failed: {rc}")
    return (out, res) if residues else out


def probab_prime(limbs: np.ndarray, reps: int = 25) -> np.ndarray:
    a = np.ascontiguousarray(limbs, dtype=np.uint32)
    out = np.zeros(a.shape[0], dtype=np.int32)
    lib.orc_probab_prime(_u32(a), a.shape[0], a.shape[1], reps,
                         out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out


def splitmix_stream(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint64)
    lib.orc_splitmix_stream(ctypes.c_uint64(seed), count, _u64(out))
    return out


def xoshiro_from_state(state, count: int) -> np.ndarray:
    st = np.ascontiguousarray(state, dtype=np.uint64)
    assert st.shape == (4,)
    out = np.zeros(count, dtype=np.uint64)
    lib.orc_xoshiro_from_state(_u64(st), count, _u64(out))
    return out


def xoshiro_stream(seed: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint64)
    lib.orc_xoshiro_stream(seed, count, _u64(out))
    return out


def gen_multipliers(seed: int, m_lo: int, m_hi: int, d: int, n_m: int) -> np.ndarray:
    out = np.zeros(n_m, dtype=np.uint64)
    rc = lib.orc_gen_multipliers(seed, m_lo, m_hi, d, n_m, _u64(out))
    if rc != 0:
        raise ValueError("orc_gen_multipliers failed")
    return out


def gen_sieved(P, d, m, x_max, per_m, limbs, sieve_limit=1_000_000, sign=1):
    """N = m*K + sign*X survivors (sign -1: the prev-prime side)."""
    m = np.ascontiguousarray(m, dtype=np.uint64)
    cap = len(m) * per_m
    out = np.zeros((cap, limbs), dtype=np.uint32)
    om = np.zeros(cap, dtype=np.uint32)
    ox = np.zeros(cap, dtype=np.uint32)
    w = ctypes.c_size_t(0)
    rc = lib.orc_gen_sieved_side(P, d, sign, _u64(m), len(m), x_max, per_m, sieve_limit, limbs,
                                 _u32(out), _u32(om), _u32(ox), ctypes.byref(w))
    if rc != 0:
        raise ValueError(f"orc_gen_sieved failed: {rc}")
    n = w.value
    return out[:n], om[:n], ox[:n]


def small_factor(limbs: np.ndarray, limit: int) -> np.ndarray:
    a = np.ascontiguousarray(limbs, dtype=np.uint32)
    out = np.zeros(a.shape[0], dtype=np.uint32)
    lib.orc_small_factor(_u32(a), a.shape[0], a.shape[1], limit, _u32(out))
    return out


def time_powm2(limbs: np.ndarray, threads: int, warmup: bool = True):
    """(seconds, fermat_count) for one threaded pass of mpz_powm(2, N-1, N)."""
    a = np.ascontiguousarray(limbs, dtype=np.uint32)
    cnt = np.zeros(1, dtype=np.uint32)
    sec = lib.orc_time_powm2(_u32(a), a.shape[0], a.shape[1], threads, int(warmup), _u32(cnt))
    return sec, int(cnt[0])


def build_mk(K_words, m: np.ndarray, x: int, sign: int, S: int):
    """(candidate-major limbs [n, S] of m[i]*K + sign*x, seconds on one thread):
    fill_batch's construction (gpu_testing.cu:171-172) with GMP."""
    k = np.ascontiguousarray(K_words, dtype=np.uint32)
    mm = np.ascontiguousarray(m, dtype=np.uint64)
    out = np.zeros((len(mm), S), dtype=np.uint32)
    sec = ctypes.c_double(0.0)
    rc = lib.orc_build_mk(_u32(k), len(k), _u64(mm), len(mm), x, sign, S, _u32(out),
                          ctypes.byref(sec))
    if rc != 0:
        raise ValueError(f"orc_build_mk failed: {rc}")
    return out, sec.value


def next_prime_offset(K_words, m: int, sign: int, limit: int) -> int:
    """X of GMP's next prime above m*K (sign 1) or previous below (sign -1), 0 if > limit."""
    k = np.ascontiguousarray(K_words, dtype=np.uint32)
    out = np.zeros(1, dtype=np.uint32)
    rc = lib.orc_next_prime_offset(_u32(k), len(k), m, sign, limit, _u32(out))
    if rc != 0:
        raise ValueError(f"orc_next_prime_offset failed: {rc}")
    return int(out[0])


# ---- m-interval search step (msieve.c) ------------------------------------

def bitmap_words(m_inc: int) -> int:
    return m_inc // 64 + 2


def ms_active(P: int, d: int, m_start: int, m_inc: int) -> np.ndarray:
    n = ctypes.c_size_t(0)
    rc = lib.orc_ms_active(P, d, m_start, m_inc, None, ctypes.byref(n))
    if rc != 0:
        raise ValueError(f"orc_ms_active failed: {rc}")
    out = np.zeros(max(1, n.value), dtype=np.uint32)
    lib.orc_ms_active(P, d, m_start, m_inc, _u32(out), ctypes.byref(n))
    return out[:n.value]


def ms_coprime_x(P: int, d: int) -> np.ndarray:
    n = ctypes.c_size_t(0)
    lib.orc_ms_coprime_x(P, d, None, 0, ctypes.byref(n))
    out = np.zeros(max(1, n.value), dtype=np.uint32)
    lib.orc_ms_coprime_x(P, d, _u32(out), n.value, ctypes.byref(n))
    return out[:n.value]


def ms_sieve(K_words: np.ndarray, P: int, d: int, m_start: int, m_inc: int, max_prime: int,
             X: int, as_written: bool = False, timed: bool = False):
    """Composites bitmap (bit t <-> m_i = 2t + 1), m_inc/64 + 2 words; with
    timed, also the seconds of the per-X part (one thread)."""
    k = np.ascontiguousarray(K_words, dtype=np.uint32)
    out = np.zeros(bitmap_words(m_inc), dtype=np.uint32)
    sec = ctypes.c_double(0)
    rc = lib.orc_ms_sieve(_u32(k), len(k), P, d, m_start, m_inc, max_prime, X, int(as_written),
                          _u32(out), ctypes.byref(sec))
    if rc != 0:
        raise ValueError(f"orc_ms_sieve failed: {rc}")
    return (out, sec.value) if timed else out


def ms_step(K_words, m_start: int, X: int, composites, active: np.ndarray, found: np.ndarray,
            bases=(2,)):
    """One X step; returns (tests, found_mi) and updates active (returned
    trimmed) and found (in place)."""
    k = np.ascontiguousarray(K_words, dtype=np.uint32)
    act = np.ascontiguousarray(active, dtype=np.uint32).copy()
    na = ctypes.c_size_t(len(act))
    b = np.asarray(bases, dtype=np.uint32)
    tests = np.zeros(max(1, len(act)), dtype=np.uint32)
    fmi = np.zeros(max(1, len(act)), dtype=np.uint32)
    nt, nf = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert found.dtype == np.uint32 and found.flags.c_contiguous
    rc = lib.orc_ms_step(_u32(k), len(k), m_start, X, _u32(np.ascontiguousarray(composites)),
                         _u32(act), ctypes.byref(na), _u32(found), _u32(b), len(b), _u32(tests),
                         ctypes.byref(nt), _u32(fmi), ctypes.byref(nf))
    if rc != 0:
        raise ValueError(f"orc_ms_step failed: {rc}")
    return tests[:nt.value], fmi[:nf.value], act[:na.value]


# ---- pure-Python restatements (CPython bignum; small inputs) ---------------

def py_residue(N: int, a: int = 2) -> int:
    return pow(a, N - 1, N)


def py_ref_verdict(N: int, a: int) -> bool:
    """miller_rabin.h:221-245 as written."""
    nm1 = N - 1
    t = (nm1 & -nm1).bit_length() - 1
    d = nm1 >> t
    x = pow(a, d, N)
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


def py_sprp(N: int, a: int) -> bool:
    nm1 = N - 1
    t = (nm1 & -nm1).bit_length() - 1
    d = nm1 >> t
    x = pow(a, d, N)
    if x == 1 or x == nm1:
        return True
    for _ in range(t - 1):
        x = x * x % N
        if x == nm1:
            return True
        if x == 1:
            return False
    return False
