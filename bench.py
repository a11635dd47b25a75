#!/usr/bin/env python3
"""Benchmark of the transform path on one MI355X.

Workload (one "step"): the mixed-length batched complex FFT — 4096 transforms
with lengths drawn uniformly (numpy default_rng(0)) from {1024, 4096, 65536,
6720 = 3*5*7*64, 1009 (Rader prime), 10007 (Rader prime), 20014 = 2*10007
(Bluestein)}, re/im iid N(0,1) from seed 0, packed back to back in HBM, one
forward transform of the whole batch per step through one plan. The headline
value is the complex64 (FP32) form's GFLOP/s with the reference's flop count
5*N*log2(N) per transform; the plain complex128 (FP64) form is timed beside it.

Prints ONE JSON line (the bench contract), with a `roofline` object for the
dominant kernel and a `cpu_baseline` object (the C++ restatement of the
reference's CPU algorithm timed on this host).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LENGTHS = [1024, 4096, 65536, 6720, 1009, 10007, 20014]
COUNT = 4096
HBM_PEAK_GBS = 8000.0


def workload_lengths(seed=0):
    return [int(v) for v in np.random.default_rng(seed).choice(LENGTHS, size=COUNT)]


def gaussian_complex(total, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(total) + 1j * rng.standard_normal(total)


def flops_of(lengths):
    return float(sum(5.0 * n * math.log2(n) for n in lengths))


def time_events(fn, steps, warmup, torch):
    """Average milliseconds per call of fn() over `steps` calls (after warmup),
    bracketed by device synchronisation, measured with events on the current
    stream (the stream the library launches on)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=1)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from fft_amd import device as D

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    lengths = workload_lengths(0)
    total = sum(lengths)
    flops = flops_of(lengths)
    x64 = gaussian_complex(total, seed=0)

    results = {}
    for name, tdtype, ndtype in [("fp32", torch.complex64, np.complex64),
                                 ("fp64", torch.complex128, np.complex128)]:
        plan = D.ManyPlan(lengths, tdtype)
        xin = torch.from_numpy(x64.astype(ndtype)).to(dev)
        out = torch.empty_like(xin)
        ms = time_events(lambda: plan.exec(xin, out=out), args.steps, args.warmup, torch)
        elt = 8 if name == "fp32" else 16
        alg_bytes = 2.0 * total * elt
        # per-length groups, for the roofline of the dominant kernel
        groups = {}
        off = 0
        offsets = {}
        for n in lengths:
            offsets.setdefault(n, []).append(off)
            off += n
        for n in sorted(set(lengths)):
            cnt = len(offsets[n])
            gp = D.C2CPlan(n, cnt, tdtype)
            gx = torch.from_numpy(x64[: n * cnt].astype(ndtype)).to(dev)
            gy = torch.empty_like(gx)
            gms = time_events(lambda: gp.exec(gx, out=gy), max(5, args.steps // 2), 2, torch)
            groups[n] = {"count": cnt, "ms": gms, "launches": gp.num_launches(),
                         "alg_bytes": 2.0 * n * cnt * elt, "flops": cnt * 5.0 * n * math.log2(n)}
            del gp, gx, gy
        results[name] = {"ms": ms, "gflops": flops / (ms * 1e6), "gbs": alg_bytes / (ms * 1e6),
                         "alg_bytes": alg_bytes, "groups": groups, "launches": plan.num_launches()}
        del plan, xin, out
        torch.cuda.empty_cache()

    # Dominant kernel: the single-launch group with the largest share of the
    # FP32 step (multi-launch groups are listed in `groups`).
    g32 = results["fp32"]["groups"]
    single = {n: g for n, g in g32.items() if g["launches"] == 1}
    dom_n = max(single, key=lambda n: single[n]["ms"])
    dom = single[dom_n]
    achieved = dom["alg_bytes"] / (dom["ms"] * 1e6)
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "kernel": "k_pass<float,16> single-pass Stockham, length %d x %d transforms" % (dom_n, dom["count"]),
        "alg_bytes_per_launch": dom["alg_bytes"], "ms_per_launch": round(dom["ms"], 5),
    }

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        O.lib()
        # Bounded sample: the first 512 transforms of the workload (about 1/8),
        # serial like the reference (it has no batch API).
        sample = lengths[:512]
        off = 0
        xs = []
        for n in sample:
            xs.append(x64[off: off + n])
            off += n
        for n in set(sample):
            O.fft(xs[sample.index(n)])  # build plans outside the timed region
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            for v in xs:
                O.fft(v)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        cpu = {"value": round(flops_of(sample) / best / 1e9, 3), "unit": "GFLOP/s", "cores": 1,
               "kind": "port",
               "sample": "first 512 of the 4096 mixed-length transforms, complex128, serial "
                         "Plan.FFT per transform (C++ restatement of go-fft's CPU algorithm, "
                         "-O2 -ffp-contract=off), best of 2; host %s, %d logical cores"
                         % (os.uname().machine, os.cpu_count() or 0),
               "seconds": round(best, 3)}

    r32, r64 = results["fp32"], results["fp64"]
    line = {
        "metric": "FFT GFLOP/s (5N*log2N) & HBM GB/s: FP32 batched vs FP64 plain, 1 GPU",
        "value": round(r32["gflops"], 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (f32 butterflies, FP64-generated tables)",
        "data": "synthetic: re/im iid N(0,1) from numpy default_rng(0)",
        "config": {
            "workload": "mixed-length batched complex FFT: 4096 transforms, lengths uniform from "
                        "{1024,4096,65536,6720,1009,10007,20014}, one plan, inputs resident in HBM",
            "transforms": COUNT, "points": total, "flops_per_step": flops,
            "parallelism": "1 GPU",
        },
        "hbm_gbs": round(r32["gbs"], 1),
        "fp64_plain": {"gflops": round(r64["gflops"], 2), "gbs": round(r64["gbs"], 1),
                       "ms_per_step": round(r64["ms"], 4)},
        "fp32_over_fp64": round(r32["gflops"] / r64["gflops"], 3),
        "launches_per_step": {"fp32": r32["launches"], "fp64": r64["launches"]},
        "groups_ms": {"fp32": {str(n): round(g["ms"], 4) for n, g in r32["groups"].items()},
                      "fp64": {str(n): round(g["ms"], 4) for n, g in r64["groups"].items()}},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
