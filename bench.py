"""Benchmark of the batched PRP hot path on one MI355X.

Workload (the GPU plain-kernel configuration): P=337, d=2310, 512-bit limb
buffers (16 x u32, candidates of 462 bits), 2^20 sieved candidates
N = m*P#/d + X from 64 multipliers m drawn from [1e7, 1e7+1e4) coprime to d,
X widened to [1, 450000] so that each m contributes 16384 survivors of a sieve
by every prime < 1e6.  One step = one pass over all 2^20 candidates as 16
launches of 65536 candidates each (base-2 Fermat, residues written), inputs
already resident in HBM in limb-interleaved layout.

Printed JSON line: tests/s (whole job), the dominant kernel's roofline against
the measured v_mad_u64_u32 issue rate (integer-VALU bound; no HBM or MFMA
bound applies: ~130 B of HBM traffic per test against ~2e5 multiply-adds), and
the GMP CPU baseline (mpz_powm(2, N-1, N)) timed on a bounded sample of the
same candidates on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "prime-gap-gpu_amd"))

METRIC = "PRP tests/sec @512/1024/2048-bit vs plain; % INT32 VALU peak; p99 m-latency"
BATCH = 65536
N_M = 64
PER_M = 16384
P, D, LIMBS = 337, 2310, 16
X_MAX = 450000
SEED = 0x5EED0002


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def work_per_test(bits: int) -> int:
    """Algorithmic 32x32->64 multiply-adds per base-2 test: (L-1)(2s^2+s),
    s = ceil(L/32) (one Montgomery product per exponent bit of N-1)."""
    s = (bits + 31) // 32
    return (bits - 1) * (2 * s * s + s)


def kernel_limbs(max_bits: int) -> int:
    """Radix limbs of the kernel the library picks: the smallest S with
    RB*S - 4 >= max_bits (RB = 29 up to 31 limbs, 28 above)."""
    S = 1
    while (29 if S <= 31 else 28) * S - 4 < max_bits:
        S += 1
    return S


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


MIXED = [(211, 16), (337, 16), (503, 24), (701, 32)]  # (P, buffer words), d = 210


def bench_mixed(stream, peak: float, reps: int) -> dict:
    """2^21 candidates, P in {211, 337, 503, 701}, d=210: one hipGraph launch
    over four autotuned persistent buckets vs the plain kernel (one launch per
    65536) on the same HBM-resident data."""
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
    macs = sum(work_per_test(median_bits(L)) * len(L) for L in data)
    plan.close()
    runner.close()
    return {
        "workload": "P in {211,337,503,701}, d=210, 2^21 sieved candidates in 16/16/24/32-word "
                    "buffers, 4 buckets by kernel limb count, one hipGraph launch",
        "candidates": total,
        "graph_ms": ms_graph, "graph_tests_per_s": total / (ms_graph * 1e-3),
        "plain_ms": ms_plain, "plain_tests_per_s": total / (ms_plain * 1e-3),
        "speedup_vs_plain": ms_plain / ms_graph,
        "valu_frac": macs / (ms_graph * 1e-3) / peak,
        "buckets": per_bucket,
        "parity_sample_ok": ok,
    }


def bench_large(stream, peak: float, reps: int) -> dict:
    """P=907, d=30030 (~1260-bit N in 64-word buffers), 2^18 candidates:
    base-2 Fermat (persistent, autotuned), on-device compaction of the PRPs,
    strong tests to bases 3 and 5 on the survivors only.  Pipelined as four
    sub-batches: Fermat + compaction on one stream, base 3 and base 5 each on
    their own stream, so sub-batch k's Miller-Rabin overlaps k+1's Fermat."""
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
    SUB = 4
    step = n // SUB

    def sl(ptr, k):  # column offset k * step of a limb-interleaved buffer
        return ptr + 4 * k * step

    def fermat_only():
        fermat.test_persistent(a.bucket(), pgx.MODE_FERMAT, residues=True,
                               stream=stream.cuda_stream)

    def pipeline():
        for k in range(SUB):
            fb = pgx.Bucket(sl(a.cand.data_ptr(), k), a.ld, a.limbs, step, a.max_bits,
                            sl(a.verdict.data_ptr(), k), sl(a.res.data_ptr(), k), a.ld)
            fermat.test_persistent(fb, pgx.MODE_FERMAT, residues=True, stream=stream.cuda_stream)
            fermat.compact(fb.d_verdict, step, pgx.FLAG_FERMAT, sl(index.data_ptr(), k),
                           count.data_ptr() + 4 * k, stream.cuda_stream)
            ev = torch.cuda.Event()
            ev.record(stream)
            for run, st, vv in ((r3, s3, v3), (r5, s5, v5)):
                st.wait_event(ev)
                mb = pgx.Bucket(fb.d_cand, a.ld, a.limbs, step, a.max_bits,
                                sl(vv.data_ptr(), k), 0, a.ld)
                run.test_indexed(mb, sl(index.data_ptr(), k), count.data_ptr() + 4 * k,
                                 pgx.MODE_STRONG, stream=st.cuda_stream)
        for st in (s3, s5):  # join the side streams back for the timing events
            ev = torch.cuda.Event()
            ev.record(st)
            stream.wait_event(ev)

    pipeline()
    ms_fermat = time_ms(stream, fermat_only, reps)
    v3.fill_(-1)
    v5.fill_(-1)
    ms_all = time_ms(stream, pipeline, reps)
    k = int(count[:SUB].sum().item())
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
                             ((f3[s_sample] == 1) & (f5[s_sample] == 1)).astype(np.uint32), want))
    bits = median_bits(L)
    for r in (fermat, r3, r5):
        r.close()
    return {
        "workload": "P=907 d=30030, 2^18 sieved candidates (~1260 bits) in 64-word buffers: "
                    "Fermat base 2 -> device compaction -> strong bases 3 and 5 on survivors, "
                    "4 sub-batches, Miller-Rabin on two side streams",
        "candidates": n, "bits": bits, "kernel_limbs": kernel_limbs(a.max_bits),
        "tuned": t.as_dict(),
        "fermat_ms": ms_fermat, "fermat_tests_per_s": n / (ms_fermat * 1e-3),
        "pipeline_ms": ms_all, "pipeline_tests_per_s": n / (ms_all * 1e-3),
        "survivors": k, "prp_all_bases": int(len(prp)),
        "fermat_valu_frac": work_per_test(bits) * n / (ms_fermat * 1e-3) / peak,
        "parity_sample_ok": ok,
    }


def bench_serving(rate: float = 2e4, duration_s: float = 1.0, templates_per_p: int = 100) -> dict:
    """Poisson stream of per-m next/prev-prime requests (P in {337, 503, 907},
    d=2310, 20-500 sieved X per side), continuous batching under a 64 GB cap.
    Each request is submitted at its scheduled arrival time; latency runs from
    that time to completion, so driver lateness counts against the engine."""
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
    results, lat = {}, []
    with pgx.Server(hbm_budget=64 << 30, chunk=32) as srv:
        for k in range(len(sizes)):  # warm each width class (allocation, code load)
            P, bits, nxt, prv = tmpl[k * templates_per_p]
            srv.submit(10**9 + k, nxt, prv, bits)
        warm = 0
        t_warm = time.time()
        while warm < len(sizes):
            warm += len(srv.poll())
            if time.time() - t_warm > 60:
                raise RuntimeError(f"serving warm-up did not finish: {srv.stats()}")
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
        wall = (time.monotonic_ns() - start) * 1e-9
        tests = sum(r[2] for r in results.values())
    lat = np.array(lat) * 1e-6
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
        "tests": tests, "tests_per_s": tests / wall,
        "results_match_oracle_and_repeat": bool(consistent),
    }


def cpu_baseline(cands: np.ndarray, target_cpu_s: float = 15.0) -> dict:
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # the GMP path, timed as the reported CPU baseline

    threads = min(16, len(os.sched_getaffinity(0)))
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
                  f"threads static shards, {sec:.2f} s wall ({per_test * 1e6:.1f} us/test "
                  f"single-thread); {fermat} Fermat PRPs",
        "single_thread_tests_per_s": 1.0 / per_test,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline plain-kernel workload")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a kernel of known "
                         "traffic for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the path "
              "shards by candidate with no exchange; see DESIGN.md)", file=sys.stderr)
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
    if args.copy_probe:  # 256 MiB read + 256 MiB written by one copy kernel
        src = torch.ones(64 << 20, dtype=torch.int32, device="cuda")
        dst = torch.empty_like(src)
        with torch.cuda.stream(stream):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    rate = {name: pgx.valu_rate(op) for op, name in pgx.VALU_OPS.items()}
    peak = rate["v_mad_u64_u32"]
    log("VALU rates (lane-ops/s): " + ", ".join(f"{k}={v:.3e}" for k, v in rate.items() if v))

    extra = {}
    if not args.no_extras:
        extra["mixed_graph"] = bench_mixed(stream, peak, max(2, args.steps // 4))
        extra["large_fermat_mr"] = bench_large(stream, peak, max(2, args.steps // 4))
        srv = bench_serving()
        srv["plain_512_tests_per_s"] = tests_per_s
        extra["serving"] = srv
        log(f"serving: p50 {srv['p50_ms']:.2f} ms p99 {srv['p99_ms']:.2f} ms, "
            f"{srv['tests_per_s'] / 1e6:.2f} M tests/s")

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
            "peak": peak / 1e12 if peak else None,
            "unit": "TMAC/s",
            "frac": achieved / peak if peak else None,
            "traffic": None,
            "kernel": f"k_prp_plain<{kernel_limbs(max_bits)},true>",
            "macs_per_launch": macs_launch,
            "avg_launch_ms": avg_launch_ms,
            "valu_rates_per_s": rate,
        },
        "cpu_baseline": cpu,
        "parity_sample_ok": parity,
        "configs": extra,
    }
    print(json.dumps(line), flush=True)
    runner.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
