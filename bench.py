"""Benchmark of the batched PRP hot path on one MI355X.

Workload (the GPU plain-kernel configuration): P=337, d=2310, 512-bit limb
buffers (16 x u32, candidates of 462 bits), 2^20 sieved candidates
N = m*P#/d + X from 64 multipliers m drawn from [1e7, 1e7+1e4) coprime to d,
X widened to [1, 450000] so that each m contributes 16384 survivors of a sieve
by every prime < 1e6.  One step = one pass over all 2^20 candidates as 16
launches of 65536 candidates each (base-2 Fermat, residues written), inputs
already resident in HBM in limb-interleaved layout.

Printed JSON line: tests/s (whole job), the dominant kernel's roofline against
MI355X's full-rate VALU issue (256 CUs x 64 lanes x 2.4 GHz = 3.93e13
v_mad_u64_u32 per second; integer-VALU bound, no HBM or MFMA bound applies:
~130 B of HBM traffic per test against ~2e5 multiply-adds), beside the rate a
pure-mad probe reaches on the same device, and the GMP CPU baseline
(mpz_powm(2, N-1, N)) timed on a bounded sample of the same candidates on the
host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

# One hardware queue per stream: the serving engine runs each width class on
# its own stream, and a class whose stream shares a hardware queue with a
# wider class waits behind that class's kernels (set before HIP starts).
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "prime-gap-gpu_amd"))

METRIC = "PRP tests/sec @512/1024/2048-bit vs plain; % INT32 VALU peak; p99 m-latency"
BATCH = 65536
N_M = 64
PER_M = 16384
P, D, LIMBS = 337, 2310, 16
X_MAX = 450000
SEED = 0x5EED0002
# MI355X full-rate VALU issue: 256 CUs x 4 SIMDs x 16 lanes per clock at the
# 2.4 GHz peak engine clock, one v_mad_u64_u32 (32x32->64 multiply-add) per
# lane per issue.  AMD publishes no INT32 VALU figure; this matches the
# published fp64 vector FMA rate (78.6 TFLOP/s = 2 x 3.93e13).
PEAK_SPEC = 256 * 64 * 2.4e9


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def work_per_test(bits: int) -> int:
    """Algorithmic 32x32->64 multiply-adds per base-2 test: (L-1)(2s^2+s),
    s = ceil(L/32) (one Montgomery product per exponent bit of N-1)."""
    s = (bits + 31) // 32
    return (bits - 1) * (2 * s * s + s)


# Kernel limb counts the library instantiates (one-lane kernels).
KERNEL_LIMBS = list(range(1, 49)) + [56, 64, 74]


def kernel_limbs(max_bits: int) -> int:
    """Radix limbs of the kernel the library picks: the smallest instantiated S
    with RB*S - 4 >= max_bits (RB = 29 up to 31 limbs, 28 above)."""
    return next(S for S in KERNEL_LIMBS if (29 if S <= 31 else 28) * S - 4 >= max_bits)


def gen(P: int, d: int, words: int, n_m: int, per_m: int, seed: int) -> np.ndarray:
    import pgx
    m = pgx.gen_multipliers(seed, 10**7, 10**7 + 10**4, d, n_m)
    L, _, _ = pgx.gen_sieved(P, d, m, X_MAX, per_m, words)
    return L


def median_bits(L: np.ndarray) -> int:
    import pgx
    return int(np.median([int(x).bit_length() for x in pgx.from_limbs(L[:1024])]))


def time_ms(stream, fn, reps: int) -> float:
    """Average ms of fn() over reps, HIP events on `stream`."""
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def host_threads() -> tuple[int, dict]:
    """Threads for the CPU legs: every core this process may run on
    (sched_getaffinity), capped by OMP_NUM_THREADS where the environment sets
    it (the CPU share a shared host grants one GPU's jobs); and the counts
    behind that choice."""
    aff = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    cap = int(omp) if omp.isdigit() and int(omp) > 0 else aff
    return min(aff, cap), {"affinity_cores": aff, "cpu_count": os.cpu_count(),
                           "omp_num_threads": int(omp) if omp.isdigit() else None}


def cpu_rate(cands: np.ndarray, cpu_seconds: float = 3.0) -> dict:
    """GMP mpz_powm(2, N-1, N) tests/s on host_threads() threads over a bounded
    sample of about cpu_seconds of CPU work (the oracle, timed as the reported
    CPU baseline)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads, _ = host_threads()
    sec1, _ = oracle.time_powm2(cands[:32], threads=1, warmup=True)
    per = sec1 / min(32, len(cands))
    n = int(min(len(cands), max(256, cpu_seconds / per)))
    sec, _ = oracle.time_powm2(cands[:n], threads=threads, warmup=False)
    return {"tests_per_s": n / sec, "threads": threads, "sample": n,
            "us_per_test_one_thread": per * 1e6}


MIXED = [(211, 16), (337, 16), (503, 24), (701, 32)]  # (P, buffer words), d = 210

# HBM traffic of the headline kernel, from two rocprofv3 --pmc passes of this
# script with --copy-probe (profiles/k_prp_plain_s17_p337_d2310_2e20.md):
# KiB per dispatch of FETCH_SIZE and WRITE_SIZE, each scaled by known bytes /
# counter reading of the 256 MiB device copy in the same run (copy read
# 262144 KiB, FETCH_SIZE read 131087.19; copy wrote 262144 KiB, WRITE_SIZE 262144).
TRAFFIC_KIB = {
    "k_prp_plain<17,true>": {"fetch": 2202.42, "fetch_scale": 262144 / 131087.19,
                             "write": 4352.0, "write_scale": 1.0},
}


def bench_mixed(stream, peak: float, reps: int) -> dict:
    """2^21 candidates, P in {211, 337, 503, 701}, d=210: one hipGraph launch
    over four autotuned persistent buckets vs the plain kernel (one launch per
    65536) on the same HBM-resident data.  Beside them, the plain kernel at one
    width for the whole mix, as the reference runs a job (one compile-time
    BITS, gpu_testing.cu:43-62; a CGBN instance of that size for every
    candidate): the widest bucket's kernel over the same per-P slices, and over
    the candidates interleaved P by P in 32-word buffers (an unbucketed
    stream: every wavefront then holds a 987-bit candidate)."""
    import torch

    import oracle
    import pgx
    t0 = time.time()
    data = [gen(P, 210, w, 32, 16384, SEED + P) for P, w in MIXED]
    batches = [pgx.DeviceBatch(L) for L in data]
    log(f"mixed: generated {sum(len(L) for L in data)} candidates in {time.time() - t0:.1f}s")
    runner = pgx.Runner(1, limbs=16)
    tunings, per_bucket = [], {}
    for (P, w), b, L in zip(MIXED, batches, data):
        t, rate = runner.autotune(b.bucket(), pgx.MODE_FERMAT)
        tunings.append(t)
        per_bucket[f"P{P}"] = {"bits": median_bits(L), "kernel_limbs": kernel_limbs(b.max_bits),
                               "tuned": t.as_dict(), "tests_per_s": rate}
        log(f"mixed: P={P} tuned {t.as_dict()} {rate / 1e6:.2f} M/s")
    plan = pgx.Plan(runner, [b.bucket() for b in batches], pgx.MODE_FERMAT, residues=True,
                    tunings=tunings)
    plan.launch(stream.cuda_stream)
    ms_graph = time_ms(stream, lambda: plan.launch(stream.cuda_stream), reps)
    ok = True
    for b, L in zip(batches, data):
        ov, ores = oracle.verdicts(L[:1024], oracle.MODE_FERMAT, residues=True)
        ok &= bool(np.array_equal(b.verdicts()[:1024], ov) and
                   np.array_equal(b.residues()[:1024], ores))

    def plain():
        for b in batches:
            for c0 in range(0, b.n, BATCH):
                k = min(BATCH, b.n - c0)
                runner.test_device(b.cand.data_ptr() + 4 * c0, b.ld, b.limbs, k, b.max_bits,
                                   pgx.MODE_FERMAT, b.verdict.data_ptr() + 4 * c0,
                                   b.res.data_ptr() + 4 * c0, b.ld, stream.cuda_stream)
    plain()
    ms_plain = time_ms(stream, plain, max(1, reps // 2))
    total = sum(len(L) for L in data)
    widest = max(b.max_bits for b in batches)

    def one_width():
        for b in batches:
            for c0 in range(0, b.n, BATCH):
                k = min(BATCH, b.n - c0)
                runner.test_device(b.cand.data_ptr() + 4 * c0, b.ld, b.limbs, k, widest,
                                   pgx.MODE_FERMAT, b.verdict.data_ptr() + 4 * c0,
                                   b.res.data_ptr() + 4 * c0, b.ld, stream.cuda_stream)
    one_width()
    ms_one_width = time_ms(stream, one_width, max(1, reps // 2))
    ok_one_width = all(
        np.array_equal(b.verdicts()[:1024], oracle.verdicts(L[:1024], oracle.MODE_FERMAT))
        for b, L in zip(batches, data))
    # interleaved P by P (equal counts per P), zero-padded to 32 words
    per = min(len(L) for L in data)
    mixed = np.zeros((per, len(data), 32), dtype=np.uint32)
    for q, L in enumerate(data):
        mixed[:, q, :L.shape[1]] = L[:per]
    mixed = mixed.reshape(per * len(data), 32)
    plan.close()
    ub = pgx.DeviceBatch(mixed, max_bits=widest)

    def unbucketed():
        for c0 in range(0, ub.n, BATCH):
            k = min(BATCH, ub.n - c0)
            runner.test_device(ub.cand.data_ptr() + 4 * c0, ub.ld, ub.limbs, k, widest,
                               pgx.MODE_FERMAT, ub.verdict.data_ptr() + 4 * c0,
                               ub.res.data_ptr() + 4 * c0, ub.ld, stream.cuda_stream)
    unbucketed()
    ms_unbucketed = time_ms(stream, unbucketed, max(1, reps // 2))
    ok_one_width &= bool(np.array_equal(
        ub.verdicts()[:1024], oracle.verdicts(mixed[:1024], oracle.MODE_FERMAT)))
    macs = sum(work_per_test(median_bits(L)) * len(L) for L in data)
    runner.close()
    cpu = [cpu_rate(L, 1.5) for L in data]
    cpu_s = sum(len(L) / c["tests_per_s"] for L, c in zip(data, cpu))
    return {
        "workload": "P in {211,337,503,701}, d=210, 2^21 sieved candidates in 16/16/24/32-word "
                    "buffers, 4 buckets by kernel limb count, one hipGraph launch",
        "candidates": total,
        "graph_ms": ms_graph, "graph_tests_per_s": total / (ms_graph * 1e-3),
        "plain_ms": ms_plain, "plain_tests_per_s": total / (ms_plain * 1e-3),
        "speedup_vs_plain": ms_plain / ms_graph,
        "one_width_plain": {
            "kernel_limbs": kernel_limbs(widest),
            "per_p_slices_ms": ms_one_width,
            "per_p_slices_tests_per_s": total / (ms_one_width * 1e-3),
            "unbucketed_candidates": per * len(data),
            "unbucketed_ms": ms_unbucketed,
            "unbucketed_tests_per_s": per * len(data) / (ms_unbucketed * 1e-3),
            "graph_speedup_vs_per_p_slices": ms_one_width / ms_graph,
            "graph_speedup_vs_unbucketed": (ms_unbucketed / (per * len(data))) /
                                           (ms_graph / total),
            "parity_sample_ok": bool(ok_one_width),
        },
        "valu_frac": macs / (ms_graph * 1e-3) / peak,
        "buckets": per_bucket,
        "parity_sample_ok": ok,
        "cpu_tests_per_s": total / cpu_s, "cpu_threads": cpu[0]["threads"],
    }


def bench_persistent(cands: np.ndarray, stream, peak: float, bits: int, reps: int) -> dict:
    """The headline's 2^20 candidates in one limb-interleaved buffer, tested by
    one autotuned persistent launch per step instead of 16 plain launches."""
    import pgx
    a = pgx.DeviceBatch(cands)
    run = pgx.Runner(1, limbs=LIMBS)
    t, _ = run.autotune(a.bucket(), pgx.MODE_FERMAT)

    def step():
        run.test_persistent(a.bucket(), pgx.MODE_FERMAT, residues=True, tuning=t,
                            stream=stream.cuda_stream)
    step()
    ms = time_ms(stream, step, reps)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ov = oracle.verdicts(cands[:1024], oracle.MODE_FERMAT)
    ok = bool(np.array_equal(a.verdicts()[:1024], ov))
    run.close()
    n = len(cands)
    return {
        "workload": "the headline's 2^20 candidates, one persistent launch per step",
        "tuned": t.as_dict(), "ms_per_step": ms, "tests_per_s": n / (ms * 1e-3),
        "valu_frac": work_per_test(bits) * n / (ms * 1e-3) / peak, "parity_sample_ok": ok,
    }


def bench_plain_streams(cands: np.ndarray, max_bits: int, reps: int, nstreams: int = 3) -> dict:
    """The headline's plain launches (65,536 candidates each) spread round-robin
    over `nstreams` streams, as the reference runs three runner instances
    concurrently (gpu_testing.cu:203-354, one stream each): neighbouring
    launches overlap, so a SIMD holds two wavefronts at the seams."""
    import torch
    import pgx
    nb = len(cands) // BATCH
    host = np.ascontiguousarray(cands.reshape(nb, BATCH, LIMBS).transpose(0, 2, 1))
    dev = torch.from_numpy(host.view(np.int32)).cuda()
    verdict = torch.zeros((nb, BATCH), dtype=torch.int32, device="cuda")
    res = torch.zeros((nb, LIMBS, BATCH), dtype=torch.int32, device="cuda")
    main = torch.cuda.Stream()
    streams = [torch.cuda.Stream() for _ in range(nstreams)]
    runner = pgx.Runner(1, limbs=LIMBS)

    def step():
        go = torch.cuda.Event()
        go.record(main)
        for s in streams:
            s.wait_event(go)
        for b in range(nb):
            s = streams[b % nstreams]
            runner.test_device(dev[b].data_ptr(), BATCH, LIMBS, BATCH, max_bits, pgx.MODE_FERMAT,
                               verdict[b].data_ptr(), res[b].data_ptr(), BATCH, s.cuda_stream)
        for s in streams:
            done = torch.cuda.Event()
            done.record(s)
            main.wait_event(done)
    step()
    ms = time_ms(main, step, reps)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ov = oracle.verdicts(cands[-1024:], oracle.MODE_FERMAT)
    ok = bool(np.array_equal(verdict[-1].cpu().numpy().view(np.uint32)[-1024:], ov))
    runner.close()
    return {"workload": f"the headline's 16 plain launches round-robin over {nstreams} streams",
            "ms_per_step": ms, "tests_per_s": len(cands) / (ms * 1e-3),
            "parity_sample_ok": ok}


def bench_host_boundary(cands: np.ndarray, reps: int = 3) -> dict:
    """The headline batches handed over at the C ABI from host memory, as the
    drop-in does: pgx_run with candidate limbs (memcpy into pinned staging,
    H2D, device transpose, kernel, D2H of the verdicts) and pgx_submit_mk with
    the multipliers only (candidates built on the device).  PCIe and host work
    included.  Then the construction alone: the device builder against GMP."""
    import pgx
    K = pgx.primorial_over(P, D)
    n = len(cands)
    mult = np.asarray([int(v) // K for v in pgx.from_limbs(cands[:BATCH])], dtype=np.uint64)
    x = int(pgx.from_limbs(cands[:1])[0]) - int(mult[0]) * K
    out = {}
    with pgx.Runner(BATCH, limbs=LIMBS) as run:
        run.run(cands[:BATCH], pgx.MODE_REFERENCE)
        t0 = time.perf_counter()
        for _ in range(reps):
            for c0 in range(0, n, BATCH):
                run.run(cands[c0:c0 + BATCH], pgx.MODE_REFERENCE)
        out["pgx_run_tests_per_s"] = reps * n / (time.perf_counter() - t0)
        run.run_mk(K, mult, x, 1, pgx.MODE_REFERENCE)
        t0 = time.perf_counter()
        for _ in range(reps * (n // BATCH)):
            run.run_mk(K, mult, x, 1, pgx.MODE_REFERENCE)
        out["pgx_submit_mk_tests_per_s"] = reps * (n // BATCH) * BATCH / (time.perf_counter() - t0)
    out["workload"] = ("headline candidates in 65536-candidate batches from host memory, "
                       "synchronous per batch (one runner, one stream)")
    # the reference's own structure: three runners, each driven by a host
    # thread of its own (gpu_testing.cu:203-354 starts one run_gpu_thread per
    # runner); the library calls release the GIL, so one thread's host copy
    # overlaps another's kernel
    from concurrent.futures import ThreadPoolExecutor
    runners = [pgx.Runner(BATCH, limbs=LIMBS) for _ in range(3)]
    try:
        for r in runners:
            r.run(cands[:BATCH], pgx.MODE_REFERENCE)

        def drive(k):
            for c0 in range(k * BATCH, n, 3 * BATCH):
                runners[k].run(cands[c0:c0 + BATCH], pgx.MODE_REFERENCE)

        with ThreadPoolExecutor(3) as ex:
            t0 = time.perf_counter()
            for _ in range(reps):
                list(ex.map(drive, range(3)))
            out["pgx_run_three_threads_tests_per_s"] = reps * n / (time.perf_counter() - t0)
    finally:
        for r in runners:
            r.close()
    # construction alone (fill_batch, gpu_testing.cu:136-195): the device
    # builder on resident buffers against GMP on one host thread
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    kw = pgx.to_limbs([K], (K.bit_length() + 31) // 32)[0]
    d_k = torch.from_numpy(kw.view(np.int32)).cuda()
    d_m = torch.from_numpy(mult.view(np.int64)).cuda()
    d_out = torch.zeros((LIMBS, BATCH), dtype=torch.int32, device="cuda")
    d_st = torch.zeros(1, dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    with pgx.Runner(1, limbs=LIMBS) as run:
        def build():
            run.build_candidates(d_k.data_ptr(), len(kw), d_m.data_ptr(), 0, x, 1, BATCH,
                                 d_out.data_ptr(), BATCH, LIMBS, d_st.data_ptr(),
                                 stream.cuda_stream)
        build()
        ms = time_ms(stream, build, 50)
    host = d_out.cpu().numpy().view(np.uint32).T
    want, cpu_s = oracle.build_mk(kw, mult, x, 1, LIMBS)
    out["device_build_candidates_per_s"] = BATCH / (ms * 1e-3)
    out["device_build_ms_per_65536"] = ms
    out["cpu_build_candidates_per_s_one_thread"] = BATCH / cpu_s
    out["build_equals_gmp"] = bool(np.array_equal(host, want) and int(d_st.cpu()[0]) == 0)
    return out


def bench_dropin(n_x: int = 8, passes: int = 3, keep_dir: str | None = None) -> dict:
    """The drop-in gpu_testing.cpp run by tests/dropin/harness under the
    reference's own dispatch protocol: GPU_BATCHES = 2 host threads, each with
    a runner and a GPUBatch of up to 32768 multipliers (candidates built on the
    device), fill_batch / process_finished_batch on the host, the X barrier
    between steps.  Interval: P=337, d=2310, m in [1e9, 1e9+1e7), composites
    from the device sieve by the primes below 1e8, first n_x coprime X.  PCIe
    and host work included.  Pass 1 warms up (runner creation); the rate is
    over the later passes.  Cross-checked X by X against pgx_search_step over
    the same interval.  With keep_dir the harness's input and output files
    stay there (in.bin, out.bin), e.g. for a profiled rerun of the harness."""
    import struct
    import subprocess
    import tempfile

    import pgx
    harness = os.path.join(ROOT, "tests", "dropin", "harness")
    if not os.path.exists(harness):
        return {"skipped": "tests/dropin/harness is not built"}
    P, d, m_start, m_inc, mp = 337, 2310, 10**9, 10_000_000, 100_000_000
    K = pgx.primorial_over(P, d)
    # the reference's X list (SieveData ctor, gap_search_gpu.cpp:236-254): even
    # X < 10 P with no prime factor among the primes <= P that do not divide d
    kp = [q for q in range(3, P + 1) if all(q % r for r in range(2, int(q ** 0.5) + 1))
          and d % q]
    xs = [x for x in range(2, 10 * P, 2) if all(x % q for q in kp)][:n_x]
    with pgx.Runner(1 << 20, limbs=16) as run:
        with pgx.Search(run, K, P, d, m_start, m_inc, mp) as s:
            active = s.active()
            comps = []
            for X in xs:
                s.sieve(X)
                comps.append(s.bitmap(pgx.SEARCH_COMPOSITES))
        with pgx.Search(run, K, P, d, m_start, m_inc, mp) as s:
            want = [s.step(X) for X in xs]
            want_found = s.bitmap(pgx.SEARCH_FOUND)
    with tempfile.TemporaryDirectory() as tmp:
        if keep_dir:
            os.makedirs(keep_dir, exist_ok=True)
        where = keep_dir or tmp
        inp, out = os.path.join(where, "in.bin"), os.path.join(where, "out.bin")
        with open(inp, "wb") as f:
            f.write(b"PGXH")
            f.write(struct.pack("<IIQIIII", P, d, m_start, m_inc, len(xs), len(comps[0]),
                                len(active)))
            f.write(active.astype("<u4").tobytes())
            for X, c in zip(xs, comps):
                f.write(struct.pack("<I", X))
                f.write(np.ascontiguousarray(c, dtype="<u4").tobytes())
        r = subprocess.run([harness, inp, out, str(passes)], capture_output=True, text=True,
                           errors="replace", timeout=600)
        if r.returncode != 0:
            return {"failed": (r.stdout + r.stderr)[-2000:]}
        raw = open(out, "rb").read()
    (nx,) = struct.unpack_from("<I", raw, 0)
    pairs = np.frombuffer(raw, dtype="<u4", count=2 * nx, offset=4).reshape(nx, 2)
    (w,) = struct.unpack_from("<I", raw, 4 + 8 * nx)
    found = np.frombuffer(raw, dtype="<u4", count=w, offset=8 + 8 * nx)
    line = r.stdout.split("seconds per pass:")[1]
    secs = [float(v) for v in line.split(";")[0].split()]
    stages = dict(re.findall(r"(fill|queued|run|results) ([0-9.]+)", line))
    tests = int(pairs[:, 0].sum())
    equal = bool([tuple(int(v) for v in p) for p in pairs] == [tuple(x) for x in want]
                 and np.array_equal(found, want_found[:w]) and not want_found[w:].any())
    timed = secs[1:] if len(secs) > 1 else secs
    return {
        "workload": f"drop-in gpu_testing.cpp under the reference's dispatch (2 GPU threads "
                    f"on CU halves, "
                    f"batches of 32768 built on the device): P={P} d={d} m in [1e9, 1e9+1e7), "
                    f"primes < 1e8, first {n_x} coprime X",
        "tests_per_pass": tests, "found_per_pass": int(pairs[:, 1].sum()),
        "seconds_per_pass": secs,
        "tests_per_s": tests / float(np.mean(timed)),
        "equal_to_device_search": equal,
        "stage_seconds_all_passes": {k: float(v) for k, v in stages.items()},
        "harness_line": r.stdout.strip().splitlines()[-1],
    }


def bench_rounds(cands: np.ndarray, stream, n: int = 1 << 18) -> dict:
    """Several Miller-Rabin bases (test_runner_t(n, rounds) with rounds > 1,
    miller_rabin.h:322-349) on the first n headline candidates, reference mode.
    pgx_run in batches of 65536 (one pass over every base: below two device
    fills) and of n (the first base on every candidate, the others on the
    device-compacted list of those that passed it); pgx_test_device one pass
    over 65536-candidate batches on resident buffers (no PCIe) for scale."""
    import pgx
    import torch
    n = min(n, len(cands))
    out = {"workload": f"first {n} headline candidates, reference mode"}
    verdicts = {}
    for rounds in (1, 2, 7):
        for batch in (BATCH, n):
            if rounds == 1 and batch == n:
                continue
            with pgx.Runner(batch, rounds=rounds, limbs=LIMBS) as run:
                run.run(cands[:batch], pgx.MODE_REFERENCE)
                t0 = time.perf_counter()
                got = [run.run(cands[c0:c0 + batch], pgx.MODE_REFERENCE)
                       for c0 in range(0, n, batch)]
                out[f"pgx_run_rounds{rounds}_batch{batch}_tests_per_s"] = (
                    n / (time.perf_counter() - t0))
                verdicts[(rounds, batch)] = np.concatenate(got)
                if batch == BATCH and rounds > 1:
                    # two batches in flight on the runner: batch k+1 is submitted
                    # before batch k is waited for, so its first base runs beside
                    # batch k's later bases
                    starts = list(range(0, n, batch))
                    run.submit(cands[:batch], pgx.MODE_REFERENCE)  # the second slot's
                    run.submit(cands[:batch], pgx.MODE_REFERENCE)  # buffers, made once
                    run.wait()
                    run.wait()
                    # three timed passes, the median reported: a single pass
                    # is a few milliseconds, and one-off host stalls inside it
                    # (seen once per process, several ms) would dominate it
                    rates = []
                    for _ in range(3):
                        t0 = time.perf_counter()
                        got2 = []
                        run.submit(cands[:batch], pgx.MODE_REFERENCE)
                        for k in range(1, len(starts) + 1):
                            if k < len(starts):
                                run.submit(cands[starts[k]:starts[k] + batch], pgx.MODE_REFERENCE)
                            got2.append(run.wait())
                        rates.append(n / (time.perf_counter() - t0))
                    out[f"pgx_submit_two_in_flight_rounds{rounds}_batch{batch}_tests_per_s"] = (
                        float(np.median(rates)))
                    verdicts[(rounds, batch, 2)] = np.concatenate(got2)
    batches = [pgx.DeviceBatch(cands[c0:c0 + BATCH], residues=False) for c0 in range(0, n, BATCH)]
    for rounds in (1, 7):
        with pgx.Runner(1, rounds=rounds, limbs=LIMBS) as run:
            def one_pass():
                for b in batches:
                    run.test_device(b.cand.data_ptr(), b.ld, LIMBS, b.n, b.max_bits,
                                    pgx.MODE_REFERENCE, b.verdict.data_ptr(),
                                    stream=stream.cuda_stream)
            one_pass()
            ms = time_ms(stream, one_pass, 2)
            out[f"one_pass_device_rounds{rounds}_tests_per_s"] = n / (ms * 1e-3)
        if rounds == 7:
            torch.cuda.synchronize()
            one = np.concatenate([b.verdicts() for b in batches])
            out["rounds7_verdicts_equal"] = bool(
                np.array_equal(one, verdicts[(7, BATCH)]) and np.array_equal(one, verdicts[(7, n)])
                and np.array_equal(one, verdicts[(7, BATCH, 2)]))
    out["passed_first_base"] = int((verdicts[(7, n)] >= 1).sum())
    out["passed_all_7"] = int((verdicts[(7, n)] == 7).sum())
    return out


def bench_1024(stream, peak: float, reps: int) -> dict:
    """The 1024-bit class alone: P=701, d=210 (987-bit N in 32-word buffers),
    2^18 candidates; plain launches of 65536 vs one autotuned launch."""
    import pgx
    L = gen(701, 210, 32, 16, 16384, SEED + 701)
    a = pgx.DeviceBatch(L)
    run = pgx.Runner(1, limbs=32)
    t, _ = run.autotune(a.bucket(), pgx.MODE_FERMAT)

    def plain():
        for c0 in range(0, a.n, BATCH):
            run.test_device(a.cand.data_ptr() + 4 * c0, a.ld, a.limbs, min(BATCH, a.n - c0),
                            a.max_bits, pgx.MODE_FERMAT, a.verdict.data_ptr() + 4 * c0,
                            a.res.data_ptr() + 4 * c0, a.ld, stream.cuda_stream)

    def tuned():
        run.test_persistent(a.bucket(), pgx.MODE_FERMAT, residues=True, tuning=t,
                            stream=stream.cuda_stream)
    plain()
    ms_plain = time_ms(stream, plain, reps)
    tuned()
    ms_tuned = time_ms(stream, tuned, reps)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ov, ores = oracle.verdicts(L[:1024], oracle.MODE_FERMAT, residues=True)
    ok = bool(np.array_equal(a.verdicts()[:1024], ov) and np.array_equal(a.residues()[:1024], ores))
    run.close()
    bits = median_bits(L)
    w = work_per_test(bits) * a.n
    return {
        "workload": "P=701 d=210, 2^18 sieved candidates (~987 bits), 32-word buffers, Fermat",
        "bits": bits, "tuned": t.as_dict(),
        "plain_tests_per_s": a.n / (ms_plain * 1e-3), "plain_valu_frac": w / (ms_plain * 1e-3) / peak,
        "tuned_tests_per_s": a.n / (ms_tuned * 1e-3), "tuned_valu_frac": w / (ms_tuned * 1e-3) / peak,
        "tuned_over_plain": ms_plain / ms_tuned, "parity_sample_ok": ok,
    }


# Width classes timed with the plain kernel alone (--plain-classes): (P, d,
# buffer words, multipliers, survivors per multiplier), the same generators and
# seeds as the 1024-bit, 1260-bit (Fermat -> MR) and 2048-bit legs.
PLAIN_CLASSES = {
    "1024": (701, 210, 32, 16, 16384),    # 2^18 candidates, ~987 bits, k_prp_plain<36>
    "1260": (907, 30030, 64, 32, 8192),   # 2^18 candidates, ~1260 bits, k_prp_plain<46>
    "2048": (1447, 210, 64, 16, 4096),    # 2^16 candidates, ~2020 bits, k_prp_plain<74>
}


def bench_plain_class(name: str, stream, peak: float, reps: int) -> dict:
    """One width class through the plain kernel alone: launches of 65536
    candidates (base-2 Fermat + residues, HBM-resident), HIP events around
    every launch on the launch stream, the roofline of that kernel, a parity
    sample against GMP."""
    import torch

    import pgx
    P, d, words, n_m, per_m = PLAIN_CLASSES[name]
    L = gen(P, d, words, n_m, per_m, SEED + P)
    a = pgx.DeviceBatch(L)
    run = pgx.Runner(1, limbs=words)
    starts = list(range(0, a.n, BATCH))

    def plain(events=None):
        for k, c0 in enumerate(starts):
            if events is not None:
                events[k][0].record(stream)
            run.test_device(a.cand.data_ptr() + 4 * c0, a.ld, a.limbs, min(BATCH, a.n - c0),
                            a.max_bits, pgx.MODE_FERMAT, a.verdict.data_ptr() + 4 * c0,
                            a.res.data_ptr() + 4 * c0, a.ld, stream.cuda_stream)
            if events is not None:
                events[k][1].record(stream)
    plain()
    torch.cuda.synchronize()
    ev = [[[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in starts]
          for _ in range(reps)]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for r in range(reps):
        plain(ev[r])
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    launch_ms = float(np.mean([x[0].elapsed_time(x[1]) for er in ev for x in er]))
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ov, ores = oracle.verdicts(L[:256], oracle.MODE_FERMAT, residues=True)
    ok = bool(np.array_equal(a.verdicts()[:256], ov) and np.array_equal(a.residues()[:256], ores))
    run.close()
    bits = median_bits(L)
    per_launch = min(BATCH, a.n)
    macs = work_per_test(bits) * per_launch
    return {
        "workload": f"P={P} d={d}, {a.n} sieved candidates (~{bits} bits), {words}-word buffers, "
                    f"plain launches of {per_launch}, base-2 Fermat + residues",
        "bits": bits, "kernel": f"k_prp_plain<{kernel_limbs(a.max_bits)},true>",
        "tests_per_s": a.n / (ms * 1e-3), "avg_launch_ms": launch_ms,
        "macs_per_launch": macs, "valu_frac": macs / (launch_ms * 1e-3) / peak,
        "parity_sample_ok": ok,
    }


def bench_2048(stream, peak: float, reps: int) -> dict:
    """The 2048-bit class: P=1447, d=210 (~2020-bit N in 64-word buffers),
    2^16 sieved candidates, Fermat + residues: the plain one-lane launch
    against groups of 4, 8, 16 and 32 lanes per candidate, and the
    autotuner's pick."""
    import pgx
    L = gen(1447, 210, 64, 16, 4096, SEED + 1447)
    a = pgx.DeviceBatch(L)
    run = pgx.Runner(1, limbs=64)

    def plain():
        run.test_device(a.cand.data_ptr(), a.ld, a.limbs, a.n, a.max_bits, pgx.MODE_FERMAT,
                        a.verdict.data_ptr(), a.res.data_ptr(), a.ld, stream.cuda_stream)

    def lanes4():
        run.test_persistent(a.bucket(), pgx.MODE_FERMAT, residues=True,
                            tuning=pgx.Tuning(8, 4, 64, 4), stream=stream.cuda_stream)

    def lanes(g):
        return lambda: run.test_persistent(a.bucket(), pgx.MODE_FERMAT, residues=True,
                                           tuning=pgx.Tuning(8, 4, 64, g),
                                           stream=stream.cuda_stream)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ov, ores = oracle.verdicts(L[:256], oracle.MODE_FERMAT, residues=True)
    out = {"workload": "P=1447 d=210, 2^16 sieved candidates (~2020 bits), 64-word buffers, "
                       "Fermat + residues", "bits": median_bits(L)}
    w = work_per_test(out["bits"]) * a.n
    ok = True
    for name, fn in (("plain", plain), ("lanes4", lanes4), ("lanes8", lanes(8)),
                     ("lanes16", lanes(16)), ("lanes32", lanes(32))):
        fn()
        ms = time_ms(stream, fn, reps)
        ok = ok and bool(np.array_equal(a.verdicts()[:256], ov) and
                         np.array_equal(a.residues()[:256], ores))
        out[f"{name}_tests_per_s"] = a.n / (ms * 1e-3)
        out[f"{name}_valu_frac"] = w / (ms * 1e-3) / peak
    t, rate = run.autotune(a.bucket(), pgx.MODE_FERMAT)
    out["tuned"] = t.as_dict()
    out["tuned_tests_per_s"] = rate
    out["parity_sample_ok"] = ok
    out["cpu"] = cpu_rate(L, 3.0)
    run.close()
    return out


def bench_large(stream, peak: float, reps: int) -> dict:
    """P=907, d=30030 (~1260-bit N in 64-word buffers), 2^18 candidates:
    base-2 Fermat (persistent, autotuned), on-device compaction of the PRPs,
    then strong tests to bases 3 and 5 on the survivors only, the two bases
    concurrently on two side streams (the survivors are ~3% of the batch, so
    each base alone fills only a few percent of the SIMDs)."""
    import torch

    import oracle
    import pgx
    L = gen(907, 30030, 64, 32, 8192, SEED + 907)
    n = len(L)
    a = pgx.DeviceBatch(L)
    v3 = torch.zeros(n, dtype=torch.int32, device="cuda")
    v5 = torch.zeros(n, dtype=torch.int32, device="cuda")
    index = torch.zeros(n, dtype=torch.int32, device="cuda")
    count = torch.zeros(64, dtype=torch.int32, device="cuda")
    fermat = pgx.Runner(1, limbs=64)
    r3 = pgx.Runner(1, limbs=64, bases=[3])
    r5 = pgx.Runner(1, limbs=64, bases=[5])
    t, rate = fermat.autotune(a.bucket(), pgx.MODE_FERMAT)
    log(f"large: tuned {t.as_dict()} {rate / 1e6:.3f} M/s")
    s3, s5 = torch.cuda.Stream(), torch.cuda.Stream()

    def fermat_only():
        fermat.test_persistent(a.bucket(), pgx.MODE_FERMAT, residues=True, tuning=t,
                               stream=stream.cuda_stream)

    mr_lanes = [1]

    def pipeline():
        fermat_only()
        fermat.compact(a.verdict.data_ptr(), n, pgx.FLAG_FERMAT, index.data_ptr(),
                       count.data_ptr(), stream.cuda_stream)
        ev = torch.cuda.Event()
        ev.record(stream)
        for run, st, vv in ((r3, s3, v3), (r5, s5, v5)):
            st.wait_event(ev)
            mb = pgx.Bucket(a.cand.data_ptr(), a.ld, a.limbs, n, a.max_bits, vv.data_ptr(), 0,
                            a.ld)
            run.test_indexed(mb, index.data_ptr(), count.data_ptr(), pgx.MODE_STRONG,
                             tuning=pgx.Tuning(8, 4, 64, mr_lanes[0]), stream=st.cuda_stream)
        for st in (s3, s5):  # join the side streams back for the timing events
            ev = torch.cuda.Event()
            ev.record(st)
            stream.wait_event(ev)

    # chunked: Fermat over 65536-candidate chunks on the main stream (plain
    # grids: a shape tuned on the whole bucket can leave a chunk with fewer
    # work chunks than wavefronts), each chunk's survivors compacted and
    # tested with bases 3 and 5 on the side streams while the next chunk's
    # Fermat runs
    chunks = [(c0, min(BATCH, n - c0)) for c0 in range(0, n, BATCH)]
    t_chunk = pgx.Tuning(0, 4, 64, 1)
    c_index = torch.zeros(n, dtype=torch.int32, device="cuda")
    c_count = torch.zeros(64 * len(chunks), dtype=torch.int32, device="cuda")

    def pipeline_chunked():
        for k, (c0, cn) in enumerate(chunks):
            cb = pgx.Bucket(a.cand.data_ptr() + 4 * c0, a.ld, a.limbs, cn, a.max_bits,
                            a.verdict.data_ptr() + 4 * c0, a.res.data_ptr() + 4 * c0, a.ld)
            fermat.test_persistent(cb, pgx.MODE_FERMAT, residues=True, tuning=t_chunk,
                                   stream=stream.cuda_stream)
            idx = c_index.data_ptr() + 4 * c0
            cnt = c_count.data_ptr() + 4 * 64 * k
            fermat.compact(a.verdict.data_ptr() + 4 * c0, cn, pgx.FLAG_FERMAT, idx, cnt,
                           stream.cuda_stream)
            ev = torch.cuda.Event()
            ev.record(stream)
            for run, st, vv in ((r3, s3, v3), (r5, s5, v5)):
                st.wait_event(ev)
                mb = pgx.Bucket(a.cand.data_ptr() + 4 * c0, a.ld, a.limbs, cn, a.max_bits,
                                vv.data_ptr() + 4 * c0, 0, a.ld)
                run.test_indexed(mb, idx, cnt, pgx.MODE_STRONG,
                                 tuning=pgx.Tuning(8, 4, 64, mr_lanes[0]), stream=st.cuda_stream)
        for st in (s3, s5):
            ev = torch.cuda.Event()
            ev.record(st)
            stream.wait_event(ev)

    # steady state: batch after batch of the same 2^18 candidates, two buffer
    # sets; the Miller-Rabin stage of batch k (side streams) runs beside the
    # Fermat launch of batch k+1, and batch k+2 reuses batch k's set once its
    # Miller-Rabin stage is done.  A 1260-bit exponentiation takes ~8 ms at
    # one wavefront per SIMD, a quarter of the Fermat launch, so the Fermat
    # wavefronts that share SIMDs with survivors finish up to one such
    # latency late whatever the stream layout (same queue, queues of their
    # own, or the Miller-Rabin streams on a share of the compute units via
    # pgx_stream_create_cus).
    sets = []
    for j in range(2):
        sets.append({"b": a if j == 0 else pgx.DeviceBatch(L),
                     "idx": torch.zeros(n, dtype=torch.int32, device="cuda"),
                     "cnt": torch.zeros(64, dtype=torch.int32, device="cuda"),
                     "v3": torch.zeros(n, dtype=torch.int32, device="cuda"),
                     "v5": torch.zeros(n, dtype=torch.int32, device="cuda"), "done": []})

    def pipeline_steady(batches, lanes, main, m3, m5, tf):
        for k in range(batches):
            S = sets[k % 2]
            for ev in S["done"]:
                main.wait_event(ev)
            S["done"] = []
            b = S["b"]
            fermat.test_persistent(b.bucket(), pgx.MODE_FERMAT, residues=True, tuning=tf,
                                   stream=main.cuda_stream)
            fermat.compact(b.verdict.data_ptr(), n, pgx.FLAG_FERMAT, S["idx"].data_ptr(),
                           S["cnt"].data_ptr(), main.cuda_stream)
            ev = torch.cuda.Event()
            ev.record(main)
            for run, st, vv in ((r3, m3, S["v3"]), (r5, m5, S["v5"])):
                st.wait_event(ev)
                mb = pgx.Bucket(b.cand.data_ptr(), b.ld, b.limbs, n, b.max_bits, vv.data_ptr(), 0,
                                b.ld)
                run.test_indexed(mb, S["idx"].data_ptr(), S["cnt"].data_ptr(), pgx.MODE_STRONG,
                                 tuning=pgx.Tuning(8, 4, 64, lanes), stream=st.cuda_stream)
                done = torch.cuda.Event()
                done.record(st)
                S["done"].append(done)
        for S in sets:
            for ev in S["done"]:
                main.wait_event(ev)
            S["done"] = []

    steady_batches = 8
    steady = {}
    steady_sets = []

    def run_steady(key, lanes, main, m3, m5, tf):
        pipeline_steady(2, lanes, main, m3, m5, tf)
        steady[key] = time_ms(main, lambda: pipeline_steady(steady_batches, lanes, main, m3, m5,
                                                            tf), 1) / steady_batches
        torch.cuda.synchronize()
        steady_sets.extend((S["v3"].cpu().numpy().view(np.uint32).copy(),
                            S["v5"].cpu().numpy().view(np.uint32).copy()) for S in sets)

    for lanes in (1, 16):
        run_steady(f"lanes{lanes}", lanes, stream, s3, s5, t)
    n_cu = torch.cuda.get_device_properties(0).multi_processor_count
    # every stage on a stream with a hardware queue of its own (a CU mask of
    # every compute unit)
    raw = [fermat.cu_stream(0, n_cu), r3.cu_stream(0, n_cu), r5.cu_stream(0, n_cu)]
    main, m3, m5 = (torch.cuda.ExternalStream(x) for x in raw)
    run_steady("lanes16_own_queues", 16, main, m3, m5, t)
    for x, run in zip(raw, (fermat, r3, r5)):
        run.destroy_stream(x)
    # the Miller-Rabin streams on 12% of the compute units, Fermat on the rest
    mr_cus = max(1, n_cu * 12 // 100)
    raw = [fermat.cu_stream(mr_cus, n_cu - mr_cus), r3.cu_stream(0, mr_cus),
           r5.cu_stream(0, mr_cus)]
    main, m3, m5 = (torch.cuda.ExternalStream(x) for x in raw)
    run_steady("lanes1_mr_on_12pct_cus", 1, main, m3, m5, t)
    for x, run in zip(raw, (fermat, r3, r5)):
        run.destroy_stream(x)

    pipeline()
    ms_fermat = time_ms(stream, fermat_only, reps)
    by_lanes = {}
    for lanes in (1, 2, 4, 8, 16, 32):  # lanes per candidate of the Miller-Rabin stage
        mr_lanes[0] = lanes
        pipeline()
        by_lanes[lanes] = time_ms(stream, pipeline, reps)
    mr_lanes[0] = min(by_lanes, key=by_lanes.get)
    v3.fill_(-1)
    v5.fill_(-1)
    pipeline_chunked()
    ms_chunked = time_ms(stream, pipeline_chunked, reps)
    torch.cuda.synchronize()
    f3c = v3.cpu().numpy().view(np.uint32).copy()
    f5c = v5.cpu().numpy().view(np.uint32).copy()
    v3.fill_(-1)
    v5.fill_(-1)
    ms_all = time_ms(stream, pipeline, reps)
    k = int(count[0].item())
    v = a.verdicts()
    f3 = v3.cpu().numpy().view(np.uint32)
    f5 = v5.cpu().numpy().view(np.uint32)
    surv = np.nonzero(v & 1)[0]
    prp = surv[(f3[surv] == 1) & (f5[surv] == 1)]
    ov = oracle.verdicts(L[:1024], oracle.MODE_FERMAT)
    s_sample = surv[:256]
    want = oracle.verdicts(L[s_sample], oracle.MODE_STRONG, (3, 5))
    ok = bool(np.array_equal(v[:1024], ov) and k == len(surv) and
              np.array_equal((f3[s_sample] == 1).astype(np.uint32) +
                             ((f3[s_sample] == 1) & (f5[s_sample] == 1)).astype(np.uint32), want)
              and np.array_equal(f3c[surv], f3[surv]) and np.array_equal(f5c[surv], f5[surv])
              and all(np.array_equal(s3v[surv], f3[surv]) and np.array_equal(s5v[surv], f5[surv])
                      for s3v, s5v in steady_sets))
    bits = median_bits(L)
    for r in (fermat, r3, r5):
        r.close()
    cpu = cpu_rate(L, 3.0)
    # CPU pipeline: Fermat on all, two more exponentiations per survivor
    cpu_pipe = n / ((n + 2 * k) / cpu["tests_per_s"])
    return {
        "workload": "P=907 d=30030, 2^18 sieved candidates (~1260 bits) in 64-word buffers: "
                    "Fermat base 2 -> device compaction -> strong bases 3 and 5 on survivors, "
                    "the two bases concurrently on two streams",
        "candidates": n, "bits": bits, "kernel_limbs": kernel_limbs(a.max_bits),
        "tuned": t.as_dict(),
        "fermat_ms": ms_fermat, "fermat_tests_per_s": n / (ms_fermat * 1e-3),
        "pipeline_ms": ms_all, "pipeline_tests_per_s": n / (ms_all * 1e-3),
        "pipeline_chunked_ms": ms_chunked,
        "pipeline_chunked_tests_per_s": n / (ms_chunked * 1e-3),
        "pipeline_chunked_over_fermat": ms_chunked / ms_fermat,
        "pipeline_steady_ms_per_batch": steady,
        "pipeline_steady_best": min(steady, key=steady.get),
        "pipeline_steady_over_fermat": min(steady.values()) / ms_fermat,
        "pipeline_steady_batches": steady_batches,
        "mr_lanes_per_candidate": mr_lanes[0],
        "pipeline_ms_by_mr_lanes": {str(k): v for k, v in by_lanes.items()},
        "survivors": k, "prp_all_bases": int(len(prp)),
        "fermat_valu_frac": work_per_test(bits) * n / (ms_fermat * 1e-3) / peak,
        "parity_sample_ok": ok,
        "cpu_fermat_tests_per_s": cpu["tests_per_s"], "cpu_pipeline_tests_per_s": cpu_pipe,
        "cpu_threads": cpu["threads"],
    }


def bench_search(tune_sample: np.ndarray, n_x: int = 8) -> dict:
    """The m-interval search on the device: P=337, d=2310, m in [1e9, 1e9+2e7),
    sieve of every X by the primes below 1e8, tests list, base-2 reference
    PRP tests, found bookkeeping, for the first n_x X of the reference's
    coprime X list.  The runner's launch shape for 462-bit reference-mode
    tests is autotuned first on `tune_sample` (candidates of that width), as a
    user of the library would.  CPU baseline leg: the oracle's restatement of
    the reference's sieve (one thread) at the first X."""
    import oracle
    import pgx
    P, d, m_start, m_inc, mp = 337, 2310, 10**9, 20_000_000, 100_000_000
    K = pgx.primorial_over(P, d)
    kw = pgx.to_limbs([K], (K.bit_length() + 31) // 32)[0]
    xs = [int(x) for x in oracle.ms_coprime_x(P, d)[:n_x]]
    run = pgx.Runner(1 << 20, limbs=16)
    tb = pgx.DeviceBatch(tune_sample)
    tuned, _ = run.autotune(tb.bucket(), pgx.MODE_REFERENCE)
    del tb
    t0 = time.perf_counter()
    s = pgx.Search(run, K, P, d, m_start, m_inc, mp)
    t_setup = time.perf_counter() - t0
    n_active = len(s.active())
    # parity of the sieve at the first X (the oracle's C sieve, one thread)
    comp, cpu_sieve_s = oracle.ms_sieve(kw, P, d, m_start, m_inc, mp, xs[0], timed=True)
    n0 = s.sieve(xs[0])
    ok = bool(np.array_equal(s.bitmap(pgx.SEARCH_COMPOSITES), comp))
    t_list = s.tests()
    sample = pgx.to_limbs([(m_start + int(mi)) * K + xs[0] for mi in t_list[:4096]], 16)
    cpu = cpu_rate(sample, 2.0)
    s.test()
    s.close()
    s = pgx.Search(run, K, P, d, m_start, m_inc, mp)
    sieve_s, test_s, tests, found = [], [], 0, 0
    for X in xs:
        t0 = time.perf_counter()
        nt = s.sieve(X)
        t1 = time.perf_counter()
        nf = s.test()
        t2 = time.perf_counter()
        sieve_s.append(t1 - t0)
        test_s.append(t2 - t1)
        tests += nt
        found += nf
    s.close()
    run.close()
    wall = sum(sieve_s) + sum(test_s)
    return {
        "workload": f"m-interval search: P={P} d={d} m in [1e9, 1e9+2e7) ({n_active} active m), "
                    f"sieve primes < 1e8, first {n_x} coprime X, base-2 PRP tests",
        "tuned": tuned.as_dict(),
        "setup_s": t_setup, "active_m": n_active, "tests": tests, "found": found,
        "tests_first_x": n0,
        "sieve_ms_per_x": 1e3 * float(np.mean(sieve_s)),
        "test_ms_per_x": 1e3 * float(np.mean(test_s)),
        "tests_per_s_incl_sieve": tests / wall,
        "m_x_per_s": n_active * len(xs) / wall,
        "cpu_sieve_ms_one_thread": 1e3 * cpu_sieve_s,  # per-X part, setup excluded
        "sieve_speedup_vs_cpu_thread": cpu_sieve_s / float(np.mean(sieve_s)),
        # CPU step at the first X: the sieve on one thread + the tests on the threads
        "cpu_ms_first_x": 1e3 * (cpu_sieve_s + n0 / cpu["tests_per_s"]),
        "gpu_ms_per_x": 1e3 * wall / len(xs), "cpu_threads": cpu["threads"],
        "parity_composites_first_x": ok,
    }


def bench_serving(rate: float = 2e4, duration_s: float = 1.0, templates_per_p: int = 100,
                  group_min_bits: int = 0, group_max_items: int = 0,
                  group_lanes: int = 4, wide_cu_percent: int = 0,
                  slots_per_class: int = 1) -> dict:
    """Poisson stream of per-m next/prev-prime requests (P in {337, 503, 907},
    d=2310, 20-500 sieved X per side), continuous batching under a 64 GB cap.
    Each request is submitted at its scheduled arrival time; latency runs from
    that time to completion, so lateness of this load generator counts against the engine."""
    import oracle
    import pgx
    sizes = [(337, 16), (503, 24), (907, 48)]
    t0 = time.time()
    tmpl = []
    for P, words in sizes:
        bits = pgx.k_bits(P, 2310) + 25  # m < 2^24, plus the X carry
        for m, nxt, prv in pgx.gen_requests(P, 2310, templates_per_p, SEED + 5 * P, words):
            tmpl.append((P, bits, nxt, prv))
    log(f"serving: {len(tmpl)} request templates in {time.time() - t0:.1f}s")
    # expected first PRP per side for a sample of templates (GMP oracle)
    rng = np.random.default_rng(SEED)
    check = set(rng.choice(len(tmpl), size=24, replace=False).tolist())

    def first(limbs):
        v = oracle.verdicts(limbs, oracle.MODE_REFERENCE, (2,))
        hit = np.nonzero(v == 1)[0]
        return int(hit[0]) if len(hit) else -1
    want = {k: (first(tmpl[k][2]), first(tmpl[k][3])) for k in check}

    n_req = int(rate * duration_s)
    gaps = rng.exponential(1.0 / rate, size=n_req)
    pick = rng.integers(0, len(tmpl), size=n_req)
    results, lat, lat_by_p = {}, [], {}
    with pgx.Server(hbm_budget=64 << 30, chunk=32, group_min_bits=group_min_bits,
                    group_max_items=group_max_items, group_lanes=group_lanes,
                    wide_cu_percent=wide_cu_percent, slots_per_class=slots_per_class) as srv:
        for k in range(len(sizes)):  # warm each width class (allocation, code load)
            P, bits, nxt, prv = tmpl[k * templates_per_p]
            srv.submit(10**9 + k, nxt, prv, bits)
        warm = 0
        t_warm = time.time()
        while warm < len(sizes):
            warm += len(srv.poll())
            if time.time() - t_warm > 60:
                raise RuntimeError(f"serving warm-up did not finish: {srv.stats()}")
        import gc
        gc.collect()
        gc.disable()  # a collector pause of this loop would count as engine latency
        start = time.monotonic_ns()
        arrivals = start + np.cumsum(gaps * 1e9).astype(np.int64)
        i = 0
        last = time.monotonic_ns()
        while len(results) < n_req:
            now = time.monotonic_ns()
            if now - last > 30e9:
                raise RuntimeError(f"serving made no progress for 30 s: {srv.stats()}")
            while i < n_req and arrivals[i] <= now:
                P, bits, nxt, prv = tmpl[pick[i]]
                srv.submit(i, nxt, prv, bits, arrival_ns=int(arrivals[i]))
                i += 1
            for rid, nf, pf, tests, lns in srv.poll():
                last = now
                if rid < n_req:
                    results[rid] = (nf, pf, tests)
                    lat.append(lns)
                    lat_by_p.setdefault(tmpl[pick[rid]][0], []).append(lns)
        wall = (time.monotonic_ns() - start) * 1e-9
        gc.enable()
        tests = sum(r[2] for r in results.values())
        st = srv.stats()  # warm-up iterations included
    lat = np.array(lat) * 1e-6
    # CPU leg: GMP tests of each side in ascending X until its first PRP, one
    # thread, over a sample of request templates
    t0 = time.perf_counter()
    cpu_tests = 0
    for k in sorted(check):
        for side in (2, 3):
            f = want[k][side - 2]
            upto = len(tmpl[k][side]) if f < 0 else f + 1
            oracle.verdicts(tmpl[k][side][:upto], oracle.MODE_REFERENCE, (2,))
            cpu_tests += upto
    cpu_req_s = len(check) / (time.perf_counter() - t0)
    threads, _ = host_threads()
    consistent = all(results[r][:2] == want[pick[r]] for r in results if pick[r] in want)
    by_t = {}
    for r, v in results.items():
        by_t.setdefault(int(pick[r]), set()).add(v[:2])
    consistent &= all(len(v) == 1 for v in by_t.values())
    return {
        "workload": "Poisson per-m requests at 2e4/s for 1 s, P in {337,503,907}, d=2310, "
                    "20-500 sieved X per side, next+prev search with early exit, 64 GB cap, "
                    f"{len(tmpl)} request templates",
        "requests": n_req, "wall_s": wall,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat.max()),
        "p50_p99_ms_by_P": {str(P): [float(np.percentile(np.array(v) * 1e-6, 50)),
                                     float(np.percentile(np.array(v) * 1e-6, 99))]
                            for P, v in sorted(lat_by_p.items())},
        "tests": tests, "tests_per_s": tests / wall,
        "results_match_oracle_and_repeat": bool(consistent),
        "cpu_requests_per_s_one_thread": cpu_req_s,
        "cpu_requests_per_s": cpu_req_s * threads, "cpu_threads": threads,
        "gpu_requests_per_s": n_req / wall,
        "group_min_bits": group_min_bits, "group_max_items": group_max_items,
        "group_lanes": group_lanes,
        # device time per iteration and the device's idle time between an
        # iteration and the next on the same stream (host scheduling turnaround,
        # or no work): the part a device-side loop could take away
        "iterations": st["iterations"],
        "iteration_ms_mean": st["iter_ms"] / max(1, st["iterations"]),
        "gap_ms_mean": st["gap_ms"] / max(1, st["gaps"]),
        "gap_fraction": st["gap_ms"] / max(1e-9, st["gap_ms"] + st["iter_ms"]),
    }


def bench_serving_capacity(cfg: dict, rates=(2e4, 2.5e4, 3e4, 3.5e4, 4e4, 6e4, 8e4),
                           p99_limit_ms: float = 20.0) -> dict:
    """Offered load from 2e4 requests/s upward, each load for up to 1 s and at
    most 40,000 requests: served rate, p50 / p99 and tests/s per load.  The
    capacity is the highest offered load whose p99 stays within the limit; the
    sweep stops after the first load over it.  Arrivals are generated by this
    Python loop, so at the top loads its own lateness counts as latency."""
    points = []
    for r in rates:
        s = bench_serving(rate=r, duration_s=min(1.0, 40000 / r), **cfg)
        points.append({"offered_per_s": r, "served_per_s": s["gpu_requests_per_s"],
                       "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"],
                       "tests_per_s": s["tests_per_s"],
                       "results_ok": s["results_match_oracle_and_repeat"],
                       "iteration_ms_mean": s["iteration_ms_mean"],
                       "gap_ms_mean": s["gap_ms_mean"], "gap_fraction": s["gap_fraction"]})
        log(f"serving at {r:.0f}/s offered: served {s['gpu_requests_per_s']:.0f}/s, "
            f"p50 {s['p50_ms']:.2f} ms, p99 {s['p99_ms']:.2f} ms")
        if s["p99_ms"] > p99_limit_ms:
            break
    ok = [p["offered_per_s"] for p in points if p["p99_ms"] <= p99_limit_ms]
    return {"p99_limit_ms": p99_limit_ms, "points": points,
            "max_offered_per_s_within_limit": max(ok) if ok else None}


def cpu_baseline(cands: np.ndarray, target_cpu_s: float = 15.0) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the GMP path, timed as the reported CPU baseline

    threads, counts = host_threads()
    calib = cands[:512]
    sec1, _ = oracle.time_powm2(calib, threads=1, warmup=True)
    per_test = sec1 / len(calib)
    n = int(min(len(cands), max(4096, target_cpu_s / per_test)))
    sample = cands[:n]
    sec, fermat = oracle.time_powm2(sample, threads=threads, warmup=True)
    return {
        "value": n / sec,
        "unit": "tests/s",
        "cores": threads,
        "kind": "port",
        "sample": f"first {n} of the same candidates, GMP mpz_powm(2, N-1, N), {threads} "
                  f"threads static shards ({counts['affinity_cores']} cores in this process's "
                  f"affinity, {counts['cpu_count']} on the host, OMP_NUM_THREADS="
                  f"{counts['omp_num_threads']}), {sec:.2f} s wall ({per_test * 1e6:.1f} us/test "
                  f"single-thread); {fermat} Fermat PRPs",
        "single_thread_tests_per_s": 1.0 / per_test,
        "host_cores": counts,
    }


def summary(plain_512: float, frac_512, extra: dict) -> dict:
    """The metric's parts side by side: tests/s at 512/1024/2048 bits (plain
    launches and the batched/persistent pattern), the INT32 VALU fraction, the
    serving p99 latency, seven Miller-Rabin bases and the host hand-over
    through pgx_run (None where that leg did not run)."""
    def get(cfg, key):
        return extra.get(cfg, {}).get(key)
    return {
        "tests_per_s_512_plain": plain_512,
        "tests_per_s_512_persistent": get("same_workload_persistent", "tests_per_s"),
        "tests_per_s_1024_plain": get("class_1024", "plain_tests_per_s"),
        "tests_per_s_1024_tuned": get("class_1024", "tuned_tests_per_s"),
        "tests_per_s_2048_plain": get("class_2048", "plain_tests_per_s"),
        "tests_per_s_2048_tuned": get("class_2048", "tuned_tests_per_s"),
        "valu_frac_512": frac_512,
        "valu_frac_1024": get("class_1024", "plain_valu_frac"),
        "valu_frac_2048": get("class_2048", "plain_valu_frac"),
        "valu_frac_1024_plain_class": get("plain_1024", "valu_frac"),
        "valu_frac_1260_plain_class": get("plain_1260", "valu_frac"),
        "valu_frac_2048_plain_class": get("plain_2048", "valu_frac"),
        "serving_p50_ms": get("serving", "p50_ms"),
        "serving_p99_ms": get("serving", "p99_ms"),
        "tests_per_s_512_7_bases_pgx_run": get("rounds", "pgx_run_rounds7_batch65536_tests_per_s"),
        "tests_per_s_512_7_bases_two_in_flight": get(
            "rounds", "pgx_submit_two_in_flight_rounds7_batch65536_tests_per_s"),
        "tests_per_s_512_from_host_three_threads": get("host_boundary",
                                                       "pgx_run_three_threads_tests_per_s"),
        "tests_per_s_512_dropin_three_gpu_threads": get("dropin", "tests_per_s"),
        "tests_per_s_2048_lanes16": get("class_2048", "lanes16_tests_per_s"),
        "tests_per_s_2048_lanes32": get("class_2048", "lanes32_tests_per_s"),
        "mixed_graph_speedup_vs_one_width_unbucketed":
            (extra.get("mixed_graph", {}).get("one_width_plain") or {})
            .get("graph_speedup_vs_unbucketed"),
        "serving_capacity_per_s_p99_within_20ms": (extra.get("serving", {}).get("capacity") or {})
        .get("max_offered_per_s_within_limit"),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline plain-kernel workload")
    ap.add_argument("--extras", default="persistent,streams,host,dropin,rounds,1024,2048,mixed,"
                                        "large,search,serving",
                    help="comma-separated extra configurations to run")
    ap.add_argument("--serve-group", default="1000,32768,16,88",
                    help="bits,items,lanes[,cu%%[,slots]]: serving iterations of classes at "
                         "least `bits` wide with at most `items` candidates use `lanes` lanes per "
                         "candidate (0,0,4 = never); cu%% > 0 runs those classes on that share of "
                         "the CUs and the narrower classes on the rest; `slots` iterations of a "
                         "class may be in flight at once")
    ap.add_argument("--plain-classes", default="1024,1260,2048",
                    help="comma-separated width classes (1024,1260,2048) to time through the "
                         "plain kernel alone, one entry per class in configs")
    ap.add_argument("--dropin-dir", default=None,
                    help="keep the drop-in harness's input and output files in this folder")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path "
              "shards by candidate with no exchange: one runner per device)", file=sys.stderr)
        return 2

    import torch

    import pgx

    t0 = time.time()
    m = pgx.gen_multipliers(SEED, 10**7, 10**7 + 10**4, D, N_M)
    cands, _, _ = pgx.gen_sieved(P, D, m, X_MAX, PER_M, LIMBS)
    total = len(cands)
    assert total == N_M * PER_M == 1 << 20, total
    bits = np.array([int(x).bit_length() for x in pgx.from_limbs(cands[:4096])])
    max_bits = 462 if bits.max() <= 462 else int(bits.max())
    L = int(np.median(bits))
    log(f"generated {total} candidates (L={bits.min()}..{bits.max()}) in {time.time() - t0:.1f}s")

    nb = total // BATCH
    # HBM layout: batch b occupies limb rows [LIMBS] x [BATCH], limb-interleaved
    host = np.ascontiguousarray(cands.reshape(nb, BATCH, LIMBS).transpose(0, 2, 1))
    dev = torch.from_numpy(host.view(np.int32)).cuda()
    verdict = torch.zeros((nb, BATCH), dtype=torch.int32, device="cuda")
    res = torch.zeros((nb, LIMBS, BATCH), dtype=torch.int32, device="cuda")
    # a dedicated stream (default priority): kernels and events share it; a
    # NULL stream handle would mean the runner's own stream to the library
    stream = torch.cuda.Stream()
    runner = pgx.Runner(1, limbs=LIMBS)

    def step(events=None):
        for b in range(nb):
            if events is not None:
                events[b][0].record(stream)
            runner.test_device(dev[b].data_ptr(), BATCH, LIMBS, BATCH, max_bits, pgx.MODE_FERMAT,
                               verdict[b].data_ptr(), res[b].data_ptr(), BATCH,
                               stream.cuda_stream)
            if events is not None:
                events[b][1].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # correctness of the timed configuration on a sample (checker: GMP oracle)
    v_all = verdict.cpu().numpy().view(np.uint32).reshape(-1)
    r0 = res[0].cpu().numpy().view(np.uint32).T
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    ov, ores = oracle.verdicts(cands[:4096], oracle.MODE_FERMAT, residues=True)
    parity = bool(np.array_equal(v_all[:4096], ov) and np.array_equal(r0[:4096], ores))
    log(f"sample parity vs GMP (4096 candidates): {parity}")

    events = [[[torch.cuda.Event(enable_timing=True) for _ in range(2)] for _ in range(nb)]
              for _ in range(args.steps)]
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    start.record(stream)
    for k in range(args.steps):
        step(events[k])
    end.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    total_ms = start.elapsed_time(end)
    launch_ms = [e[0].elapsed_time(e[1]) for ev in events for e in ev]
    avg_launch_ms = float(np.mean(launch_ms))
    tests_per_s = total * args.steps / (total_ms * 1e-3)
    log(f"{args.steps} steps: {total_ms:.2f} ms (wall {wall * 1e3:.2f} ms), "
        f"{tests_per_s / 1e6:.3f} M tests/s, launch avg {avg_launch_ms:.3f} ms")

    # roofline: algorithmic multiply-adds per launch / launch duration vs the
    # measured v_mad_u64_u32 issue rate of this device
    macs_launch = work_per_test(L) * BATCH
    achieved = macs_launch / (avg_launch_ms * 1e-3)
    kname = f"k_prp_plain<{kernel_limbs(max_bits)},true>"
    tk = TRAFFIC_KIB.get(kname)
    traffic = (1024 * (tk["fetch"] * tk["fetch_scale"] + tk["write"] * tk["write_scale"])
               if tk else None)
    # algorithmic bytes per launch: candidates in, verdict words and residues out
    bytes_launch = BATCH * (4 * LIMBS + 4 + 4 * LIMBS)
    if args.copy_probe:  # 256 MiB read + 256 MiB written by one copy kernel
        src = torch.ones(64 << 20, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        with torch.cuda.stream(stream):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    rate = {name: pgx.valu_rate(op) for op, name in pgx.VALU_OPS.items()}
    probe = rate["v_mad_u64_u32"]
    # every valu_frac below is against the full-rate issue of the spec; the
    # probe's own rate (what a pure-mad kernel reaches here) is reported beside
    peak = PEAK_SPEC
    log("VALU rates (lane-ops/s): " + ", ".join(f"{k}={v:.3e}" for k, v in rate.items() if v)
        + f"; spec {PEAK_SPEC:.3e}")

    extra = {}
    todo = set() if args.no_extras else set(args.extras.split(","))
    for name in [c for c in args.plain_classes.split(",") if c]:
        extra[f"plain_{name}"] = bench_plain_class(name, stream, peak, max(2, args.steps // 4))
        pc = extra[f"plain_{name}"]
        log(f"plain {name}-bit class: {pc['tests_per_s'] / 1e6:.3f} M tests/s, launch "
            f"{pc['avg_launch_ms']:.3f} ms, VALU frac {pc['valu_frac']:.3f} ({pc['kernel']})")
    if "persistent" in todo:
        extra["same_workload_persistent"] = bench_persistent(cands, stream, peak, L,
                                                             max(2, args.steps // 4))
        log(f"persistent: {extra['same_workload_persistent']['tests_per_s'] / 1e6:.2f} M tests/s")
    if "streams" in todo:
        extra["plain_three_streams"] = bench_plain_streams(cands, max_bits, max(2, args.steps // 4))
        log(f"plain on 3 streams: {extra['plain_three_streams']['tests_per_s'] / 1e6:.2f} M tests/s")
    if "host" in todo:
        extra["host_boundary"] = bench_host_boundary(cands)
        hb = extra["host_boundary"]
        log(f"host boundary: pgx_run {hb['pgx_run_tests_per_s'] / 1e6:.2f} M/s (three runners "
            f"on three threads {hb['pgx_run_three_threads_tests_per_s'] / 1e6:.2f} M/s), pgx_submit_mk "
            f"{hb['pgx_submit_mk_tests_per_s'] / 1e6:.2f} M/s; construction on the device "
            f"{hb['device_build_candidates_per_s'] / 1e9:.2f} G/s vs GMP one thread "
            f"{hb['cpu_build_candidates_per_s_one_thread'] / 1e6:.2f} M/s "
            f"(equal: {hb['build_equals_gmp']})")
    if "dropin" in todo:
        extra["dropin"] = bench_dropin(keep_dir=args.dropin_dir)
        dr = extra["dropin"]
        log(f"drop-in: {dr.get('tests_per_s', 0) / 1e6:.2f} M tests/s through gpu_testing.cpp "
            f"on 2 GPU threads (equal to the device search: {dr.get('equal_to_device_search')})"
            if "tests_per_s" in dr else f"drop-in: {dr}")
    if "rounds" in todo:
        extra["rounds"] = bench_rounds(cands, stream)
        r = extra["rounds"]
        log("rounds: " + ", ".join(f"{k.replace('_tests_per_s', '')} {v / 1e6:.2f} M/s"
                                    for k, v in r.items() if k.endswith("_tests_per_s"))
            + f" (equal: {r['rounds7_verdicts_equal']})")
    if "1024" in todo:
        extra["class_1024"] = bench_1024(stream, peak, max(2, args.steps // 4))
        log(f"1024-bit: plain {extra['class_1024']['plain_tests_per_s'] / 1e6:.2f} M/s, "
            f"tuned {extra['class_1024']['tuned_tests_per_s'] / 1e6:.2f} M/s")
    if "2048" in todo:
        extra["class_2048"] = bench_2048(stream, peak, max(2, args.steps // 4))
        c = extra["class_2048"]
        log(f"2048-bit: plain {c['plain_tests_per_s'] / 1e6:.3f} M/s, 4/8/16/32 lanes "
            + "/".join(f"{c[f'lanes{g}_tests_per_s'] / 1e6:.3f}" for g in (4, 8, 16, 32))
            + f" M/s, tuned {c['tuned_tests_per_s'] / 1e6:.3f} M/s "
            f"{c['tuned']}, cpu {c['cpu']['tests_per_s']:.0f}/s")
    if "mixed" in todo:
        extra["mixed_graph"] = bench_mixed(stream, peak, max(2, args.steps // 4))
    if "large" in todo:
        extra["large_fermat_mr"] = bench_large(stream, peak, max(2, args.steps // 4))
    if "search" in todo:
        extra["search_interval"] = bench_search(cands[:1 << 18])
        log(f"search: sieve {extra['search_interval']['sieve_ms_per_x']:.2f} ms/X, "
            f"tests {extra['search_interval']['test_ms_per_x']:.2f} ms/X")
    if "serving" in todo:
        sg = [int(v) for v in args.serve_group.split(",")]
        gb, gi, gl = sg[:3]
        wide = sg[3] if len(sg) > 3 else 0
        slots = sg[4] if len(sg) > 4 else 1
        srv = bench_serving(group_min_bits=gb, group_max_items=gi, group_lanes=gl,
                            wide_cu_percent=wide, slots_per_class=slots)
        srv["wide_cu_percent"] = wide
        srv["slots_per_class"] = slots
        # the same stream with one lane per candidate in every class
        one = bench_serving(group_min_bits=0, group_max_items=0, group_lanes=4)
        srv["one_lane_per_candidate"] = {
            k: one[k] for k in ("p50_ms", "p99_ms", "p50_p99_ms_by_P", "tests_per_s",
                                "results_match_oracle_and_repeat")}
        srv["plain_512_tests_per_s"] = tests_per_s
        # the same stream with other group sizes for the wide class
        srv["by_group_lanes"] = {}
        for g in sorted({4, 8, 16, 32} - {gl}):
            o = bench_serving(group_min_bits=gb, group_max_items=gi, group_lanes=g,
                              wide_cu_percent=wide, slots_per_class=slots)
            srv["by_group_lanes"][str(g)] = {
                k: o[k] for k in ("p50_ms", "p99_ms", "p50_p99_ms_by_P", "tests_per_s",
                                  "results_match_oracle_and_repeat")}
            log(f"serving, {g}-lane groups for the wide class: p50 {o['p50_ms']:.2f} ms "
                f"p99 {o['p99_ms']:.2f} ms, by P {o['p50_p99_ms_by_P']}")
        srv["capacity"] = bench_serving_capacity(dict(
            group_min_bits=gb, group_max_items=gi, group_lanes=gl, wide_cu_percent=wide,
            slots_per_class=slots))
        extra["serving"] = srv
        log(f"serving: p50 {srv['p50_ms']:.2f} ms p99 {srv['p99_ms']:.2f} ms, "
            f"{srv['tests_per_s'] / 1e6:.2f} M tests/s, by P {srv['p50_p99_ms_by_P']}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cands)
        log(f"cpu baseline: {cpu['value']:.0f} tests/s on {cpu['cores']} threads")

    line = {
        "metric": METRIC,
        "value": tests_per_s,
        "unit": "tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "gpu_plain_512: P=337 d=2310, 2^20 sieved candidates (64 m, X<=450000, "
                        "primes<1e6 sieved), 16-limb buffers, 16 launches x 65536, base-2 Fermat "
                        "+ residues, plain one-launch-per-batch kernel",
            "bits": L,
            "kernel_limbs": kernel_limbs(max_bits),
            "global_batch": total,
            "launch_batch": BATCH,
        },
        "roofline": {
            "bound": "valu-int32",
            "achieved": achieved / 1e12,
            "peak": peak / 1e12,
            "unit": "TMAC/s",
            "frac": achieved / peak,
            "peak_source": "256 CU x 64 lanes x 2.4 GHz full-rate v_mad_u64_u32 issue",
            "peak_probe": probe / 1e12 if probe else None,
            "frac_probe": achieved / probe if probe else None,
            "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-scaled)",
            "traffic_source": ("constant read from the rocprofv3 --pmc passes of this kernel in "
                               "profiles/k_prp_plain_s17_p337_d2310_2e20.md, not measured in "
                               "this run") if traffic else None,
            "algorithmic_bytes_per_launch": bytes_launch,
            "kernel": kname,
            "macs_per_launch": macs_launch,
            "avg_launch_ms": avg_launch_ms,
            "valu_rates_per_s": rate,
            # the same launches against HBM (8 TB/s published): far from that bound
            "hbm": {"achieved_gbs": bytes_launch / (avg_launch_ms * 1e-3) / 1e9,
                    "peak_gbs": 8000.0,
                    "frac": bytes_launch / (avg_launch_ms * 1e-3) / 8e12},
        },
        "cpu_baseline": cpu,
        "parity_sample_ok": parity,
        "summary": summary(tests_per_s, achieved / peak, extra),
        "configs": extra,
    }
    print(json.dumps(line), flush=True)
    runner.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
