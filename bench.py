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
    rate = {name: pgx.valu_rate(op) for op, name in pgx.VALU_OPS.items()}
    peak = rate["v_mad_u64_u32"]
    log("VALU rates (lane-ops/s): " + ", ".join(f"{k}={v:.3e}" for k, v in rate.items() if v))

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
    }
    print(json.dumps(line), flush=True)
    runner.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
