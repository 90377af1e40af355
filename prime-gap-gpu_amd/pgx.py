"""Python binding of the gfx950 PRP tester's C ABI (include/pgx.h).

Host-side mirror of the reference's runner interface, test_runner_t
(miller_rabin.h:295-409): ``Runner(n, rounds)`` owns the pinned staging,
device buffers and stream; ``Runner.run_test(active, tests, results)`` has the
reference's semantics (tests = Python ints, results[i] = 1 when every base
passed, entries from ``active`` on left untouched).  ``Runner.run`` exposes the
raw verdict words and residues on numpy limb arrays, and ``Runner.test_device``
launches on device-resident, limb-interleaved buffers (torch tensors or raw
pointers) without any host copy.

The library is libpgx.so next to this file, built by ``make`` in this
directory.  There is no CPU fallback: importing this module without the built
library raises ImportError, and creating a runner without a GPU raises
PgxError.

One HIP runtime per process: PyTorch ships its own libamdhip64 (same soname
as /opt/rocm's, but its own libraries request it by another name), so torch is
imported before libpgx.so is loaded.  libpgx.so's libamdhip64.so.7 dependency
then binds to torch's copy and device pointers, streams and events are shared.
Loading libpgx.so first would bring in a second runtime that hides the GPU
from torch.
"""
from __future__ import annotations

import collections
import ctypes
import os

import numpy as np

try:  # bind libpgx.so to the HIP runtime torch loads (see above)
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - torch is part of the image
    torch = None

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpgx.so")

PGX_OK = 0
PGX_E_ARG = -1
PGX_E_HIP = -2
PGX_E_CAPACITY = -3
PGX_E_INPUT = -4
PGX_E_STATE = -5
PGX_E_NODEV = -6

MODE_REFERENCE = 0
MODE_STRONG = 1
MODE_FERMAT = 2
WANT_RESIDUE = 4
NOALLOC_MAX_N = 1 << 20  # PGX_NOALLOC_MAX_N

FLAG_FERMAT = 1
FLAG_SPRP = 2
FLAG_QUIRK = 4
VERDICT_INVALID = 0xFFFFFFFF
MAX_BITS = 2068
MAX_BASE = 255

EXPORTED = (
    "pgx_runner_create",
    "pgx_runner_destroy",
    "pgx_submit",
    "pgx_wait",
    "pgx_run",
    "pgx_test_device",
    "pgx_last_error",
    "pgx_gen_sieved",
    "pgx_gen_multipliers",
    "pgx_gen_sieved_side",
    "pgx_k_bits",
    "pgx_valu_rate",
    "pgx_test_persistent",
    "pgx_test_indexed",
    "pgx_compact",
    "pgx_autotune",
    "pgx_plan_create",
    "pgx_plan_launch",
    "pgx_plan_destroy",
    "pgx_stream_create_cus",
    "pgx_stream_destroy",
    "pgx_runner_confine",
    "pgx_serve_create",
    "pgx_serve_submit",
    "pgx_serve_poll",
    "pgx_serve_stats",
    "pgx_serve_destroy",
    "pgx_serve_last_error",
    "pgx_submit_mk",
    "pgx_build_candidates",
    "pgx_search_create",
    "pgx_search_destroy",
    "pgx_search_sieve",
    "pgx_search_test",
    "pgx_search_step",
    "pgx_search_tests",
    "pgx_search_found",
    "pgx_search_bitmap",
    "pgx_search_last_error",
)

SEARCH_AS_WRITTEN = 1
SEARCH_ACTIVE, SEARCH_COMPOSITES, SEARCH_FOUND = 0, 1, 2


class Bucket(ctypes.Structure):
    """pgx_bucket_t: one group of candidates tested by one kernel width."""
    _fields_ = [("d_cand", ctypes.c_void_p), ("ld", ctypes.c_uint32), ("limbs", ctypes.c_uint32),
                ("n", ctypes.c_uint32), ("max_bits", ctypes.c_uint32),
                ("d_verdict", ctypes.c_void_p), ("d_residues", ctypes.c_void_p),
                ("res_ld", ctypes.c_uint32)]


class ServeConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_uint64), ("chunk", ctypes.c_uint32),
                ("max_requests", ctypes.c_uint32), ("pool_columns", ctypes.c_uint32),
                ("max_side", ctypes.c_uint32), ("stage_columns", ctypes.c_uint32),
                ("group_min_bits", ctypes.c_uint32), ("group_max_items", ctypes.c_uint32),
                ("group_lanes", ctypes.c_uint32), ("wide_cu_percent", ctypes.c_uint32),
                ("slots_per_class", ctypes.c_uint32)]


class ServeRequest(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("arrival_ns", ctypes.c_uint64),
                ("limbs", ctypes.c_uint32), ("max_bits", ctypes.c_uint32),
                ("n_next", ctypes.c_uint32), ("n_prev", ctypes.c_uint32),
                ("next", ctypes.c_void_p), ("prev", ctypes.c_void_p)]


class ServeResult(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("next_first", ctypes.c_int32),
                ("prev_first", ctypes.c_int32), ("tests", ctypes.c_uint32),
                ("latency_ns", ctypes.c_uint64)]


class ServeStats(ctypes.Structure):
    _fields_ = [("bytes_in_use", ctypes.c_uint64), ("tests", ctypes.c_uint64),
                ("active", ctypes.c_uint32), ("waiting", ctypes.c_uint32),
                ("in_flight", ctypes.c_uint32), ("finished_unread", ctypes.c_uint32),
                ("iterations", ctypes.c_uint64), ("gaps", ctypes.c_uint64),
                ("iter_ms", ctypes.c_double), ("gap_ms", ctypes.c_double)]


class SearchConfig(ctypes.Structure):
    _fields_ = [("P", ctypes.c_uint32), ("d", ctypes.c_uint32), ("m_start", ctypes.c_uint64),
                ("m_inc", ctypes.c_uint32), ("max_prime", ctypes.c_uint32),
                ("flags", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    """pgx_tuning_t: persistent launch shape."""
    _fields_ = [("blocks_per_cu", ctypes.c_uint32), ("waves_per_block", ctypes.c_uint32),
                ("chunk", ctypes.c_uint32), ("lanes_per_number", ctypes.c_uint32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class PgxError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"pgx error {code}: {message}")
        self.code = code


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run make in {_HERE})")
    lib = ctypes.CDLL(LIB_PATH)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    u64p = ctypes.POINTER(ctypes.c_uint64)
    vp = ctypes.c_void_p
    lib.pgx_runner_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.c_size_t,
                                      ctypes.c_uint32, u32p, ctypes.c_uint32]
    lib.pgx_runner_destroy.argtypes = [vp]
    lib.pgx_submit.argtypes = [vp, u32p, ctypes.c_size_t, ctypes.c_uint32]
    lib.pgx_wait.argtypes = [vp, u32p, u32p]
    lib.pgx_run.argtypes = [vp, u32p, ctypes.c_size_t, ctypes.c_uint32, u32p, u32p]
    lib.pgx_test_device.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_size_t,
                                    ctypes.c_uint32, ctypes.c_uint32, vp, vp, ctypes.c_uint32, vp]
    lib.pgx_test_device_noalloc.argtypes = [ctypes.c_int, vp, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint32,
                                            u32p, ctypes.c_uint32, vp, vp, ctypes.c_uint32, vp]
    lib.pgx_last_error.argtypes = [vp]
    lib.pgx_last_error.restype = ctypes.c_char_p
    lib.pgx_gen_sieved.argtypes = [ctypes.c_uint32, ctypes.c_uint32, u64p, ctypes.c_uint32,
                                   ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                   ctypes.c_uint32, u32p, u32p, u32p,
                                   ctypes.POINTER(ctypes.c_size_t)]
    lib.pgx_gen_sieved_side.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int32, u64p,
                                        ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                        ctypes.c_uint32, ctypes.c_uint32, u32p, u32p, u32p,
                                        ctypes.POINTER(ctypes.c_size_t)]
    lib.pgx_gen_multipliers.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                        ctypes.c_uint32, ctypes.c_uint32, u64p]
    lib.pgx_k_bits.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
    lib.pgx_k_bits.restype = ctypes.c_uint32
    lib.pgx_valu_rate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                  ctypes.POINTER(ctypes.c_double)]
    bp, tp = ctypes.POINTER(Bucket), ctypes.POINTER(Tuning)
    lib.pgx_test_persistent.argtypes = [vp, bp, tp, ctypes.c_uint32, vp]
    lib.pgx_test_indexed.argtypes = [vp, bp, vp, vp, tp, ctypes.c_uint32, vp]
    lib.pgx_compact.argtypes = [vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp, vp, vp]
    lib.pgx_autotune.argtypes = [vp, bp, ctypes.c_uint32, tp, ctypes.POINTER(ctypes.c_double)]
    lib.pgx_plan_create.argtypes = [ctypes.POINTER(vp), vp, bp, ctypes.c_uint32, tp,
                                    ctypes.c_uint32]
    lib.pgx_plan_launch.argtypes = [vp, vp]
    lib.pgx_plan_destroy.argtypes = [vp]
    lib.pgx_stream_create_cus.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(vp)]
    lib.pgx_stream_destroy.argtypes = [vp, vp]
    lib.pgx_runner_confine.argtypes = [vp, ctypes.c_uint32, ctypes.c_uint32]
    lib.pgx_serve_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(ServeConfig)]
    lib.pgx_serve_submit.argtypes = [vp, ctypes.POINTER(ServeRequest)]
    lib.pgx_serve_poll.argtypes = [vp, ctypes.POINTER(ServeResult), ctypes.c_uint32,
                                   ctypes.POINTER(ctypes.c_uint32)]
    lib.pgx_serve_stats.argtypes = [vp, ctypes.POINTER(ServeStats)]
    lib.pgx_serve_destroy.argtypes = [vp]
    szp = ctypes.POINTER(ctypes.c_size_t)
    lib.pgx_search_create.argtypes = [ctypes.POINTER(vp), vp, u32p, ctypes.c_uint32,
                                      ctypes.POINTER(SearchConfig)]
    lib.pgx_search_destroy.argtypes = [vp]
    lib.pgx_search_sieve.argtypes = [vp, ctypes.c_uint32, u32p]
    lib.pgx_search_test.argtypes = [vp, u32p]
    lib.pgx_search_step.argtypes = [vp, ctypes.c_uint32, u32p, u32p]
    lib.pgx_search_tests.argtypes = [vp, u32p, ctypes.c_size_t, szp]
    lib.pgx_search_found.argtypes = [vp, u32p, u32p, ctypes.c_size_t, szp]
    lib.pgx_search_bitmap.argtypes = [vp, ctypes.c_uint32, u32p, ctypes.c_size_t, szp]
    lib.pgx_search_last_error.argtypes = [vp]
    lib.pgx_search_last_error.restype = ctypes.c_char_p
    lib.pgx_submit_mk.argtypes = [vp, u32p, ctypes.c_uint32, u64p, ctypes.c_size_t,
                                  ctypes.c_uint32, ctypes.c_int32, ctypes.c_uint32]
    lib.pgx_build_candidates.argtypes = [vp, vp, ctypes.c_uint32, vp, vp, ctypes.c_uint32,
                                         ctypes.c_int32, ctypes.c_uint32, vp, ctypes.c_uint32,
                                         ctypes.c_uint32, vp, vp]
    lib.pgx_serve_last_error.argtypes = [vp]
    lib.pgx_serve_last_error.restype = ctypes.c_char_p
    return lib


lib = _load()


def _u32(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def _u64(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def first_primes(count: int) -> list[int]:
    """The reference's generate_primes(count) (miller_rabin.h:272-289)."""
    out: list[int] = []
    c = 2
    while len(out) < count:
        if all(c % p for p in out if p * p <= c):
            out.append(c)
        c += 1
    return out


# ---- limb packing (from_mpz / to_mpz, miller_rabin.h:39-54) ---------------

def to_limbs(values, limbs: int) -> np.ndarray:
    """Candidate-major little-endian u32 limbs; raises if a value does not fit
    (the reference's from_mpz exits instead, miller_rabin.h:46-49)."""
    values = list(values)
    out = np.zeros((len(values), limbs), dtype=np.uint32)
    for i, v in enumerate(values):
        if v < 0 or v.bit_length() > 32 * limbs:
            raise PgxError(PGX_E_INPUT, f"candidate {i} does not fit in {limbs} limbs")
        b = v.to_bytes(4 * limbs, "little")
        out[i] = np.frombuffer(b, dtype="<u4")
    return out


def from_limbs(arr: np.ndarray) -> list[int]:
    arr = np.ascontiguousarray(arr, dtype="<u4")
    return [int.from_bytes(row.tobytes(), "little") for row in arr]


# ---- candidate builder (fill_batch / init_K roles) -----------------------

def primorial_over(P: int, d: int) -> int:
    """K = P# / d (gap_common.cpp:60-65)."""
    K = 1
    for q in range(2, P + 1):
        if all(q % r for r in range(2, int(q ** 0.5) + 1)):
            K *= q
    if K % d:
        raise ValueError(f"{d} does not divide {P}#")
    return K // d


def k_bits(P: int, d: int) -> int:
    return int(lib.pgx_k_bits(P, d))


def gen_multipliers(seed: int, m_lo: int, m_hi: int, d: int, n_m: int) -> np.ndarray:
    out = np.zeros(n_m, dtype=np.uint64)
    rc = lib.pgx_gen_multipliers(seed, m_lo, m_hi, d, n_m, _u64(out))
    if rc != PGX_OK:
        raise PgxError(rc, "pgx_gen_multipliers failed (not enough multipliers coprime to d?)")
    return out


def gen_sieved(P: int, d: int, m: np.ndarray, x_max: int, per_m: int, limbs: int,
               sieve_limit: int = 1_000_000, sign: int = 1):
    """Sieved candidates N = m*K + sign*X: returns (limbs[count, limbs], m_index, X)."""
    m = np.ascontiguousarray(m, dtype=np.uint64)
    cap = len(m) * per_m
    out = np.zeros((cap, limbs), dtype=np.uint32)
    om = np.zeros(cap, dtype=np.uint32)
    ox = np.zeros(cap, dtype=np.uint32)
    w = ctypes.c_size_t(0)
    rc = lib.pgx_gen_sieved_side(P, d, sign, _u64(m), len(m), x_max, per_m, sieve_limit, limbs,
                                 _u32(out), _u32(om), _u32(ox), ctypes.byref(w))
    if rc != PGX_OK:
        raise PgxError(rc, "pgx_gen_sieved failed (d must divide P#, N must fit in limbs)")
    n = w.value
    return out[:n], om[:n], ox[:n]


def gen_requests(P: int, d: int, n_req: int, seed: int, limbs: int, lo: int = 20, hi: int = 500,
                 x_max: int = 30000, sieve_limit: int = 1_000_000):
    """Next/prev-prime search requests for n_req multipliers m (coprime to d):
    each side holds between lo and hi sieved candidates (uniform), ascending X;
    next side N = m*K + X, prev side N = m*K - X.  Returns a list of
    (m, next_limbs, prev_limbs)."""
    m = gen_multipliers(seed, 10**7, 10**7 + 10**4, d, n_req)
    nxt, mi_n, _ = gen_sieved(P, d, m, x_max, hi, limbs, sieve_limit, sign=1)
    prv, mi_p, _ = gen_sieved(P, d, m, x_max, hi, limbs, sieve_limit, sign=-1)
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_req):
        cn, cp = rng.integers(lo, hi + 1, size=2)
        out.append((int(m[k]), nxt[mi_n == k][:cn], prv[mi_p == k][:cp]))
    return out


SERVE_LOAD = os.path.join(_HERE, "serve_load")


def write_serve_templates(path: str, templates) -> None:
    """Request templates for serve_load: templates = [(max_bits, next_limbs,
    prev_limbs)], each side candidate-major limbs of one width, ascending X."""
    import struct
    with open(path, "wb") as f:
        f.write(b"PGXS")
        f.write(struct.pack("<I", len(templates)))
        for max_bits, nxt, prv in templates:
            nxt = np.ascontiguousarray(nxt, dtype="<u4")
            prv = np.ascontiguousarray(prv, dtype="<u4")
            limbs = nxt.shape[1] if nxt.ndim == 2 else prv.shape[1]
            f.write(struct.pack("<IIII", limbs, int(max_bits), len(nxt), len(prv)))
            f.write(nxt.tobytes())
            f.write(prv.tobytes())


def run_serve_load(templates_path: str, rate: float, duration_s: float, seed: int,
                   timeout: float = 300, **cfg) -> dict:
    """Runs serve_load (the C++ Poisson load generator over pgx_serve_*) and
    reads its results: per request (in id order) the template index, the first
    PRP index of each side, tests and latency in ns, plus its summary line.
    cfg: pgx_serve_config_t fields (hbm_gb, chunk, group_lanes, ...)."""
    import subprocess
    import tempfile
    if not os.path.exists(SERVE_LOAD):
        raise ImportError(f"{SERVE_LOAD} is not built (run make in {_HERE})")
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "out.bin")
        args = [SERVE_LOAD, templates_path, out, repr(float(rate)), repr(float(duration_s)),
                str(int(seed))] + [f"{k}={int(v)}" for k, v in cfg.items()]
        r = subprocess.run(args, capture_output=True, text=True, errors="replace",
                           timeout=timeout)
        if r.returncode != 0:
            raise PgxError(PGX_E_HIP, f"serve_load exited {r.returncode}: "
                                      f"{(r.stdout + r.stderr)[-2000:]}")
        raw = open(out, "rb").read()
    n = int(np.frombuffer(raw, dtype="<u4", count=1)[0])
    rec = np.frombuffer(raw, dtype=np.dtype([("template", "<u4"), ("next_first", "<i4"),
                                             ("prev_first", "<i4"), ("tests", "<u4"),
                                             ("latency_ns", "<u8")]), count=n, offset=4)
    return {"template": rec["template"].copy(), "next_first": rec["next_first"].copy(),
            "prev_first": rec["prev_first"].copy(), "tests": rec["tests"].copy(),
            "latency_ns": rec["latency_ns"].copy(), "line": r.stdout.strip().splitlines()[-1]}


VALU_OPS = {0: "v_mad_u64_u32", 1: "v_mad_u64_u32_dependent", 2: "v_add_u32", 3: "v_mul_lo_u32",
            4: "v_fma_f64"}


def valu_rate(op: int, device: int = 0, blocks: int = 8192, iters: int = 4096) -> float:
    """Lane-instructions per second of VALU op `op` (see pgx_valu_rate)."""
    out = ctypes.c_double(0)
    rc = lib.pgx_valu_rate(device, op, blocks, iters, ctypes.byref(out))
    if rc != PGX_OK:
        raise PgxError(rc, f"pgx_valu_rate({op}) failed")
    return out.value


def test_device_noalloc(cand_ptr: int, ld: int, limbs: int, n: int, max_bits: int, mode: int,
                        bases, verdict_ptr: int, residue_ptr: int = 0, res_ld: int = 0,
                        stream: int = 0, device: int = 0) -> None:
    """pgx_test_device_noalloc: the several-base batch on device buffers (raw
    pointers) with no runner and no allocation."""
    b = np.asarray(list(bases), dtype=np.uint32)
    flags = mode | (WANT_RESIDUE if residue_ptr else 0)
    rc = lib.pgx_test_device_noalloc(device, ctypes.c_void_p(cand_ptr), ld, limbs, n, max_bits,
                                     flags, _u32(b), len(b), ctypes.c_void_p(verdict_ptr),
                                     ctypes.c_void_p(residue_ptr or None), res_ld,
                                     ctypes.c_void_p(stream or None))
    if rc != PGX_OK:
        raise PgxError(rc, lib.pgx_last_error(None).decode())


# ---- runner ---------------------------------------------------------------

class Runner:
    """One runner per host thread, like test_runner_t (miller_rabin.h:295-409)."""

    def __init__(self, n: int, rounds: int = 1, limbs: int = 16, bases=None, device: int = 0):
        self.n = int(n)
        self.limbs = int(limbs)
        if bases is None:
            bases = first_primes(rounds)
        self.bases = [int(b) for b in bases]
        self.rounds = len(self.bases)
        arr = np.asarray(self.bases, dtype=np.uint32)
        h = ctypes.c_void_p()
        rc = lib.pgx_runner_create(ctypes.byref(h), device, self.n, self.limbs, _u32(arr),
                                   len(self.bases))
        if rc != PGX_OK:
            raise PgxError(rc, lib.pgx_last_error(None).decode())
        self._h = h
        self._pending = collections.deque()

    def close(self):
        if getattr(self, "_h", None):
            lib.pgx_runner_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != PGX_OK:
            raise PgxError(rc, lib.pgx_last_error(self._h).decode())

    def run(self, limbs_arr: np.ndarray, mode: int = MODE_REFERENCE, residues: bool = False):
        """Verdict words (and residues as [n, limbs] u32) for candidate-major limbs."""
        a = np.ascontiguousarray(limbs_arr, dtype=np.uint32)
        if a.ndim != 2 or a.shape[1] != self.limbs:
            raise PgxError(PGX_E_ARG, f"expected [n, {self.limbs}] limbs, got {a.shape}")
        n = a.shape[0]
        v = np.zeros(n, dtype=np.uint32)
        r = np.zeros((n, self.limbs), dtype=np.uint32) if residues else None
        flags = mode | (WANT_RESIDUE if residues else 0)
        self._check(lib.pgx_run(self._h, _u32(a), n, flags, _u32(v),
                                _u32(r) if residues else None))
        return (v, r) if residues else v

    def submit(self, limbs_arr: np.ndarray, mode: int = MODE_REFERENCE, residues: bool = False):
        a = np.ascontiguousarray(limbs_arr, dtype=np.uint32)
        flags = mode | (WANT_RESIDUE if residues else 0)
        self._check(lib.pgx_submit(self._h, _u32(a), a.shape[0], flags))
        self._pending.append((a.shape[0], residues, a))  # only once the library accepted it

    def wait(self):
        """Results of the oldest submitted batch (up to two may be in flight)."""
        if not self._pending:
            raise PgxError(PGX_E_STATE, "no batch in flight")
        n, residues, _ = self._pending[0]
        v = np.zeros(n, dtype=np.uint32)
        r = np.zeros((n, self.limbs), dtype=np.uint32) if residues else None
        rc = lib.pgx_wait(self._h, _u32(v), _u32(r) if residues else None)
        self._pending.popleft()  # the library retired the batch, failed or not
        self._check(rc)
        return (v, r) if residues else v

    def submit_mk(self, K: int, m: np.ndarray, x: int, sign: int = 1,
                  mode: int = MODE_REFERENCE, residues: bool = False):
        """Build N_i = m[i]*K + sign*x on the device (fill_batch,
        gpu_testing.cu:171-172) and launch the test; ``wait`` returns as for
        ``submit``."""
        kw = max(1, (int(K).bit_length() + 31) // 32)
        karr = to_limbs([int(K)], kw)[0]
        marr = np.ascontiguousarray(m, dtype=np.uint64)
        flags = mode | (WANT_RESIDUE if residues else 0)
        self._check(lib.pgx_submit_mk(self._h, _u32(karr), kw, _u64(marr), len(marr), int(x),
                                      int(sign), flags))
        self._pending.append((len(marr), residues, marr))

    def run_mk(self, K: int, m: np.ndarray, x: int, sign: int = 1, mode: int = MODE_REFERENCE,
               residues: bool = False):
        self.submit_mk(K, m, x, sign, mode, residues)
        return self.wait()

    def build_candidates(self, K_ptr: int, k_words: int, m_ptr: int, x_ptr: int, x0: int,
                         sign: int, n: int, out_ptr: int, ld: int, out_words: int,
                         status_ptr: int, stream: int = 0) -> None:
        """Device buffers only: column i of out = m[i]*K + sign*(x[i] or x0)."""
        self._check(lib.pgx_build_candidates(self._h, ctypes.c_void_p(K_ptr), k_words,
                                             ctypes.c_void_p(m_ptr),
                                             ctypes.c_void_p(x_ptr or None), x0, sign, n,
                                             ctypes.c_void_p(out_ptr), ld, out_words,
                                             ctypes.c_void_p(status_ptr),
                                             ctypes.c_void_p(stream or None)))

    def run_test(self, active: int, tests, results) -> None:
        """Reference semantics of run_test (miller_rabin.h:364-408): tests is the
        full batch of ints; the first min(active, len(tests)) are tested and
        results[i] = (passed == rounds) for those."""
        if len(tests) == 0:
            return
        if len(tests) > self.n:
            raise PgxError(PGX_E_CAPACITY, f"tests({len(tests)}) > n({self.n})")
        to_run = min(active, len(tests))
        v = self.run(to_limbs(tests[:to_run], self.limbs), MODE_REFERENCE)
        for i in range(to_run):
            results[i] = int(v[i] == self.rounds)

    def test_persistent(self, bucket: Bucket, mode: int, residues: bool = False,
                        tuning: Tuning | None = None, stream: int = 0) -> None:
        flags = mode | (WANT_RESIDUE if residues else 0)
        self._check(lib.pgx_test_persistent(self._h, ctypes.byref(bucket),
                                            ctypes.byref(tuning) if tuning else None, flags,
                                            ctypes.c_void_p(stream or None)))

    def test_indexed(self, bucket: Bucket, index_ptr: int, count_ptr: int, mode: int,
                     tuning: Tuning | None = None, stream: int = 0) -> None:
        self._check(lib.pgx_test_indexed(self._h, ctypes.byref(bucket), ctypes.c_void_p(index_ptr),
                                         ctypes.c_void_p(count_ptr),
                                         ctypes.byref(tuning) if tuning else None, mode,
                                         ctypes.c_void_p(stream or None)))

    def compact(self, verdict_ptr: int, n: int, mask: int, index_ptr: int, count_ptr: int,
                stream: int = 0) -> None:
        self._check(lib.pgx_compact(self._h, ctypes.c_void_p(verdict_ptr), n, mask,
                                    ctypes.c_void_p(index_ptr), ctypes.c_void_p(count_ptr),
                                    ctypes.c_void_p(stream or None)))

    def cu_stream(self, cu_first: int, cu_count: int) -> int:
        """pgx_stream_create_cus: a stream (raw hipStream_t as an int) whose
        kernels run on compute units [cu_first, cu_first + cu_count) only;
        release it with destroy_stream."""
        out = ctypes.c_void_p()
        self._check(lib.pgx_stream_create_cus(self._h, cu_first, cu_count, ctypes.byref(out)))
        return int(out.value)

    def destroy_stream(self, stream: int) -> None:
        self._check(lib.pgx_stream_destroy(self._h, ctypes.c_void_p(stream)))

    def confine(self, part: int, parts: int) -> None:
        """pgx_runner_confine: the runner's kernels on the part-th of `parts`
        equal compute-unit ranges (parts = 1: every CU)."""
        self._check(lib.pgx_runner_confine(self._h, part, parts))

    def autotune(self, bucket: Bucket, mode: int) -> tuple[Tuning, float]:
        t = Tuning()
        rate = ctypes.c_double(0)
        self._check(lib.pgx_autotune(self._h, ctypes.byref(bucket), mode, ctypes.byref(t),
                                     ctypes.byref(rate)))
        return t, rate.value

    def test_device(self, cand_ptr: int, ld: int, limbs: int, n: int, max_bits: int,
                    mode: int, verdict_ptr: int, residue_ptr: int = 0, res_ld: int = 0,
                    stream: int = 0) -> None:
        """Launch on device-resident limb-interleaved buffers (raw pointers)."""
        flags = mode | (WANT_RESIDUE if residue_ptr else 0)
        self._check(lib.pgx_test_device(self._h, ctypes.c_void_p(cand_ptr), ld, limbs, n,
                                        max_bits, flags, ctypes.c_void_p(verdict_ptr),
                                        ctypes.c_void_p(residue_ptr or None), res_ld,
                                        ctypes.c_void_p(stream or None)))


class Search:
    """pgx_search_t: the m-interval search (sieve at X, test, bookkeeping) on
    the device, one X at a time like the reference's sieve and GPU threads
    (gap_search_gpu.cpp:462-804, gpu_testing.cu:104-354)."""

    def __init__(self, runner: Runner, K: int, P: int, d: int, m_start: int, m_inc: int,
                 max_prime: int, as_written: bool = False):
        self.runner = runner
        kw = max(1, (int(K).bit_length() + 31) // 32)
        karr = to_limbs([int(K)], kw)[0]
        cfg = SearchConfig(P, d, m_start, m_inc, max_prime, SEARCH_AS_WRITTEN if as_written else 0)
        h = ctypes.c_void_p()
        rc = lib.pgx_search_create(ctypes.byref(h), runner._h, _u32(karr), kw, ctypes.byref(cfg))
        if rc != PGX_OK:
            raise PgxError(rc, lib.pgx_last_error(runner._h).decode())
        self._h = h
        self.m_inc = m_inc

    def close(self):
        if getattr(self, "_h", None):
            lib.pgx_search_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != PGX_OK:
            raise PgxError(rc, lib.pgx_search_last_error(self._h).decode())

    def sieve(self, X: int) -> int:
        n = ctypes.c_uint32(0)
        self._check(lib.pgx_search_sieve(self._h, X, ctypes.byref(n)))
        return n.value

    def test(self) -> int:
        n = ctypes.c_uint32(0)
        self._check(lib.pgx_search_test(self._h, ctypes.byref(n)))
        return n.value

    def step(self, X: int) -> tuple[int, int]:
        nt, nf = ctypes.c_uint32(0), ctypes.c_uint32(0)
        self._check(lib.pgx_search_step(self._h, X, ctypes.byref(nt), ctypes.byref(nf)))
        return nt.value, nf.value

    def tests(self) -> np.ndarray:
        n = ctypes.c_size_t(0)
        self._check(lib.pgx_search_tests(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(max(1, n.value), dtype=np.uint32)
        self._check(lib.pgx_search_tests(self._h, _u32(out), len(out), ctypes.byref(n)))
        return out[:n.value]

    def found(self) -> tuple[np.ndarray, np.ndarray]:
        n = ctypes.c_size_t(0)
        self._check(lib.pgx_search_found(self._h, None, None, 0, ctypes.byref(n)))
        mi = np.zeros(max(1, n.value), dtype=np.uint32)
        x = np.zeros(max(1, n.value), dtype=np.uint32)
        self._check(lib.pgx_search_found(self._h, _u32(mi), _u32(x), len(mi), ctypes.byref(n)))
        return mi[:n.value], x[:n.value]

    def bitmap(self, which: int) -> np.ndarray:
        n = ctypes.c_size_t(0)
        self._check(lib.pgx_search_bitmap(self._h, which, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.uint32)
        self._check(lib.pgx_search_bitmap(self._h, which, _u32(out), len(out), ctypes.byref(n)))
        return out

    def active(self) -> np.ndarray:
        """Active m_i, ascending (expanded from the bitmap)."""
        words = self.bitmap(SEARCH_ACTIVE)
        bits = np.unpackbits(words.view(np.uint8), bitorder="little")[: self.m_inc // 2]
        return (2 * np.nonzero(bits)[0] + 1).astype(np.uint32)


class Plan:
    """pgx_plan_t: every bucket captured into one hipGraph launch."""

    def __init__(self, runner: Runner, buckets, mode: int, residues: bool = False,
                 tunings=None):
        self.runner = runner
        arr = (Bucket * len(buckets))(*buckets)
        tarr = (Tuning * len(buckets))(*tunings) if tunings else None
        h = ctypes.c_void_p()
        flags = mode | (WANT_RESIDUE if residues else 0)
        runner._check(lib.pgx_plan_create(ctypes.byref(h), runner._h, arr, len(buckets), tarr,
                                          flags))
        self._p = h
        self._keep = (arr, tarr)

    def launch(self, stream: int = 0) -> None:
        self.runner._check(lib.pgx_plan_launch(self._p, ctypes.c_void_p(stream or None)))

    def close(self):
        if getattr(self, "_p", None):
            lib.pgx_plan_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()


class DeviceBatch:
    """Candidates resident in HBM (torch-owned), limb-interleaved, with verdict
    and residue buffers: the device side of a pgx_bucket_t."""

    def __init__(self, cands: np.ndarray, max_bits: int | None = None, ld: int | None = None,
                 residues: bool = True):
        import torch as _torch
        cands = np.ascontiguousarray(cands, dtype=np.uint32)
        self.n, self.limbs = cands.shape
        self.ld = ld or self.n
        if max_bits is None:
            nz = [int.bit_length(int(x)) for x in from_limbs(cands)] if self.n else [2]
            max_bits = max(max(nz), 2)
        self.max_bits = int(max_bits)
        host = np.zeros((self.limbs, self.ld), dtype=np.uint32)
        host[:, :self.n] = cands.T
        self.cand = _torch.from_numpy(host.view(np.int32)).cuda()
        self.verdict = _torch.zeros(max(self.n, 1), dtype=_torch.int32, device="cuda")
        self.res = (_torch.zeros((self.limbs, self.ld), dtype=_torch.int32, device="cuda")
                    if residues else None)
        # the fills above run on torch's stream; launches on other streams
        # (the runner's own, non-blocking) must not overtake them
        _torch.cuda.synchronize()

    def bucket(self) -> Bucket:
        return Bucket(self.cand.data_ptr(), self.ld, self.limbs, self.n, self.max_bits,
                      self.verdict.data_ptr(), self.res.data_ptr() if self.res is not None else 0,
                      self.ld)

    def verdicts(self) -> np.ndarray:
        return self.verdict.cpu().numpy().view(np.uint32)[:self.n].copy()

    def residues(self) -> np.ndarray:
        return self.res.cpu().numpy().view(np.uint32)[:, :self.n].T.copy()


class Server:
    """pgx_server_t: next/prev-prime searches with continuous batching."""

    def __init__(self, hbm_budget: int = 64 << 30, chunk: int = 32, max_requests: int = 8192,
                 pool_columns: int = 1 << 21, max_side: int = 512, stage_columns: int = 1 << 17,
                 group_min_bits: int = 0, group_max_items: int = 0, group_lanes: int = 4,
                 wide_cu_percent: int = 0, slots_per_class: int = 1, device: int = 0):
        cfg = ServeConfig(hbm_budget, chunk, max_requests, pool_columns, max_side, stage_columns,
                          group_min_bits, group_max_items, group_lanes, wide_cu_percent,
                          slots_per_class)
        h = ctypes.c_void_p()
        rc = lib.pgx_serve_create(ctypes.byref(h), device, ctypes.byref(cfg))
        if rc != PGX_OK:
            raise PgxError(rc, lib.pgx_last_error(None).decode())
        self._s = h
        self._out = (ServeResult * 4096)()

    def _check(self, rc):
        if rc != PGX_OK:
            raise PgxError(rc, lib.pgx_serve_last_error(self._s).decode())

    def submit(self, rid: int, nxt: np.ndarray, prv: np.ndarray, max_bits: int,
               arrival_ns: int = 0) -> None:
        """nxt / prv: candidate-major limbs (same width), ascending X."""
        nxt = np.ascontiguousarray(nxt, dtype=np.uint32)
        prv = np.ascontiguousarray(prv, dtype=np.uint32)
        limbs = nxt.shape[1] if nxt.size else prv.shape[1]
        req = ServeRequest(rid, arrival_ns, limbs, max_bits, len(nxt), len(prv),
                           nxt.ctypes.data if len(nxt) else None,
                           prv.ctypes.data if len(prv) else None)
        self._check(lib.pgx_serve_submit(self._s, ctypes.byref(req)))

    def poll(self) -> list[tuple[int, int, int, int, int]]:
        """(id, next_first, prev_first, tests, latency_ns) of completed requests."""
        got = []
        while True:
            n = ctypes.c_uint32(0)
            self._check(lib.pgx_serve_poll(self._s, self._out, len(self._out), ctypes.byref(n)))
            got += [(r.id, r.next_first, r.prev_first, r.tests, r.latency_ns)
                    for r in self._out[:n.value]]
            if n.value < len(self._out):
                return got

    def stats(self) -> dict:
        st = ServeStats()
        self._check(lib.pgx_serve_stats(self._s, ctypes.byref(st)))
        return {k: (float if k.endswith("_ms") else int)(getattr(st, k)) for k, _ in st._fields_}

    def close(self):
        if getattr(self, "_s", None):
            lib.pgx_serve_destroy(self._s)
            self._s = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
