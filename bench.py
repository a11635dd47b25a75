#!/usr/bin/env python3
"""Benchmark of the transform path on one MI355X.

Workload (one "step"): the mixed-length batched complex FFT — 4096 transforms
with lengths drawn uniformly (numpy default_rng(0)) from {1024, 4096, 65536,
6720 = 3*5*7*64, 1009 (Rader prime), 10007 (Rader prime), 20014 = 2*10007
(Bluestein)}, re/im iid N(0,1) from seed 0, packed back to back in HBM, one
forward transform of the whole batch per step through one plan, replayed as
the plan's captured graph (the same exec launched directly is timed too). The
headline value is the complex64 (FP32) form's GFLOP/s with the reference's
flop count 5*N*log2(N) per transform; the plain complex128 (FP64) form is
timed beside it.

Prints ONE JSON line (the bench contract), with a `roofline` object for the
dominant kernel (the Stockham pass kernel k_pass<float,16>) and a
`cpu_baseline` object (the C++ restatement of the reference's CPU algorithm
timed on this host).

Roofline of the pass kernel: the step runs its length groups side by side on
several streams, so its launches overlap. Each length group is therefore also
timed on its own, on one stream, through a plan of exactly that group
replayed as its captured graph (the step's count, at least 20 warm-up
replays); the groups whose launches are all
pass-kernel launches give the kernel's time per step. The library reports
each launch's data bytes (its input and output lines once each), so achieved
= bytes / time. `trace_window` names those isolated dispatches among the run's
k_pass<float, ...> dispatches, for comparison with a kernel trace. `traffic`
(HBM bytes per launch from FETCH_SIZE + WRITE_SIZE, calibrated on a 1 GiB copy
run by the same command) is read from the committed counter summary when
present.

`python bench.py --summarize-trace STATS.csv [SKIP TAKE]` and
`--summarize-pmc FETCH.csv WRITE.csv` print those summaries from rocprofv3's
CSV output.
"""
import argparse
import csv
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
COPY_BYTES = 1 << 30  # one direction of the calibration copy
PASS_KERNEL = "k_pass<float, 16, *>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_pass_f32_mixed4096_traffic.txt")


def workload_lengths(seed=0):
    return [int(v) for v in np.random.default_rng(seed).choice(LENGTHS, size=COUNT)]


def gaussian_complex(total, seed=0):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(total) + 1j * rng.standard_normal(total)


def flops_of(lengths):
    return float(sum(5.0 * n * math.log2(n) for n in lengths))


GROUP_WARMUP = 20  # warm-up replays of each length group alone (roofline)


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


def read_traffic():
    """HBM bytes per pass-kernel launch from the committed counter summary."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def run_dtype(D, torch, dev, name, lengths, x64, steps, warmup):
    tdtype, ndtype, elt = {"fp32": (torch.complex64, np.complex64, 8),
                           "fp64": (torch.complex128, np.complex128, 16)}[name]
    L = D.L
    total = sum(lengths)
    plan = D.ManyPlan(lengths, tdtype)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        xin = torch.from_numpy(x64.astype(ndtype)).to(dev)
        out = torch.empty_like(xin)
        # the step as plain execs, then as one replayed graph of the same exec
        ms_eager = time_events(lambda: plan.exec(xin, out=out, stream=stream), steps, warmup, torch)
        graph = D.Graph(plan, xin, out, stream=stream)
        ms = time_events(lambda: graph.launch(stream=stream), steps, warmup, torch)
    torch.cuda.synchronize()
    work = {k: plan.kernel_work(k) for k in (L.KERNEL_PASS, L.KERNEL_CONV, L.KERNEL_REAL, L.KERNEL_AUX)}
    res = {"ms": ms, "ms_eager": ms_eager, "gflops": flops_of(lengths) / (ms * 1e6),
           "gbs": 2.0 * total * elt / (ms * 1e6), "launches": plan.num_launches(), "work": work, "groups": {},
           "out": out.cpu().numpy()}  # the last replay's output, for the accuracy figures
    del graph, plan, xin, out
    counts = {}
    for n in lengths:
        counts[n] = counts.get(n, 0) + 1
    # Each group alone, replayed as its plan's captured graph like the step,
    # after at least GROUP_WARMUP replays (a fresh plan on fresh buffers runs
    # a few percent slower for its first few dozen calls).
    gwarm = max(warmup, GROUP_WARMUP)
    for n in sorted(counts):
        cnt = counts[n]
        gp = D.C2CPlan(n, cnt, tdtype)
        with torch.cuda.stream(stream):
            gx = torch.from_numpy(x64[: n * cnt].astype(ndtype)).to(dev)
            gy = torch.empty_like(gx)
            gg = D.Graph(gp, gx, gy, stream=stream)
            gms = time_events(lambda: gg.launch(stream=stream), steps, gwarm, torch)
        torch.cuda.synchronize()
        gw = {k: gp.kernel_work(k) for k in (L.KERNEL_PASS, L.KERNEL_CONV, L.KERNEL_REAL, L.KERNEL_AUX)}
        res["groups"][n] = {"count": cnt, "ms": gms, "launches": gp.num_launches(), "work": gw, "warmup": gwarm}
        del gg, gp, gx, gy
    torch.cuda.empty_cache()
    return res


def pass_roofline(r, L, steps, warmup):
    """Achieved bandwidth of the pass kernel over the step's pass-only groups,
    each timed alone on one stream."""
    ms = launches = nbytes = 0
    for g in r["groups"].values():
        w = g["work"]
        pl, pb = w[L.KERNEL_PASS]
        others = sum(w[k][0] for k in w if k != L.KERNEL_PASS)
        if pl and not others:
            ms += g["ms"]
            launches += pl
            nbytes += pb
    step_launches = r["work"][L.KERNEL_PASS][0]
    gwarm = max(warmup, GROUP_WARMUP)
    if launches != step_launches:
        raise RuntimeError("pass-kernel launches of the groups (%d) differ from the step's (%d)"
                           % (launches, step_launches))
    avg_ms = ms / launches
    achieved = nbytes / launches / (avg_ms * 1e6)
    return {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": read_traffic(),
        "kernel": PASS_KERNEL, "launches_per_step": launches,
        "alg_bytes_per_launch": round(nbytes / launches), "avg_launch_ms": round(avg_ms, 5),
        "kernel_ms_per_step": round(ms, 4),
        "measured_on": "each length group alone on one stream, replayed as its captured graph after %d "
                       "warm-up replays (the step overlaps the groups)" % gwarm,
        # the run's k_pass<float, ...> dispatches: (warmup + steps) * launches
        # in the eager step, the same in its graph replays (the captures
        # launch nothing), then the per-group replays
        "trace_window": {"skip": 2 * (steps + warmup) * step_launches, "take": (steps + gwarm) * launches},
    }


def copy_bandwidth(torch, dev, steps, warmup):
    a = torch.empty(COPY_BYTES // 4, dtype=torch.float32, device=dev)
    a.fill_(1.0)
    b = torch.empty_like(a)
    ms = time_events(lambda: b.copy_(a), steps, warmup, torch)
    del a, b
    torch.cuda.empty_cache()
    return {"gbs": round(2.0 * COPY_BYTES / (ms * 1e6), 1), "bytes_each_way": COPY_BYTES, "ms": round(ms, 4)}


def cpu_leg(lengths, x64):
    """The oracle (C++ restatement of the reference's CPU path) over the whole
    step, serial Plan.FFT per transform as the reference has no batch API."""
    from oracle import oracle as O
    O.lib()
    xs = []
    off = 0
    for n in lengths:
        xs.append(x64[off: off + n])
        off += n
    for n in set(lengths):
        O.fft(xs[lengths.index(n)])  # plans built outside the timed region
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        for v in xs:
            O.fft(v)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    ys = [O.fft(v) for v in xs]  # untimed: the outputs the accuracy figures compare against
    return {"value": round(flops_of(lengths) / best / 1e9, 3), "unit": "GFLOP/s", "cores": 1, "kind": "port",
            "sample": "the whole step: all 4096 mixed-length transforms, complex128, one thread, serial "
                      "Plan.FFT per transform (C++ restatement of go-fft's CPU algorithm, -O2 "
                      "-ffp-contract=off), best of 3; host %s with %d logical cores"
                      % (os.uname().machine, os.cpu_count() or 0),
            "seconds": round(best, 3)}, ys


def accuracy(lengths, y32, y64, ref=None):
    """Per length: the worst per-transform relative L2 error and the largest
    absolute error of the FP32 and FP64 step outputs, against the CPU leg's
    oracle outputs on the same inputs when `ref` holds them, else (FP32 only)
    against the FP64 GPU output."""
    per = {}
    off = 0
    for i, n in enumerate(lengths):
        want = ref[i] if ref is not None else y64[off: off + n]
        den = max(float(np.linalg.norm(want)), 1e-300)
        row = per.setdefault(n, {"transforms": 0})
        row["transforms"] += 1
        for name, y in (("c64", y32), ("c128", y64 if ref is not None else None)):
            if y is None:
                continue
            d = y[off: off + n].astype(np.complex128) - want
            e = float(np.linalg.norm(d)) / den
            m = float(np.abs(d).max())
            row[name + "_rel_l2"] = max(row.get(name + "_rel_l2", 0.0), e)
            row[name + "_max_abs"] = max(row.get(name + "_max_abs", 0.0), m)
        off += n
    for n, row in per.items():
        for k in list(row):
            if k != "transforms":
                row[k] = float("%.3g" % row[k])
        row["c64_bar"] = 1e-6
        if ref is not None:
            row["c128_bar"] = float("%.3g" % (1e-14 * math.log2(n)))
    return {"reference": "the oracle's outputs from the CPU leg (FP64 restatement of go-fft, same inputs)"
                         if ref is not None else "the FP64 GPU path's outputs (same inputs)",
            "rel_l2": "worst transform of the length", "max_abs": "largest |error| over the length's transforms "
                                                                  "(inputs re/im N(0,1))",
            "per_length": {str(n): per[n] for n in sorted(per)}}


# ---------------------------------------------------------------------------
# The other workloads (one JSON line each with --workload NAME); the default
# run measures the mixed-length batch above.


def _line(name, ms, flops, nbytes, dtype, config, extra=None):
    d = {"workload": name, "ms": round(ms, 4), "gflops": round(flops / (ms * 1e6), 2),
         "gbs": round(nbytes / (ms * 1e6), 1), "dtype": dtype, "config": config}
    if extra:
        d.update(extra)
    print(json.dumps(d), flush=True)


def _cpu_baseline(fn, what, reps=3):
    """The oracle's multicore form of the reference's CPU path (parChunks of
    parallel.go over every host core), best of `reps`; a reported baseline."""
    from oracle import oracle as O
    cores = os.cpu_count() or 1
    O.set_workers(cores)
    fn(cores)  # warm (plans, page faults)
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        fn(cores)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return {"ms": round(best * 1e3, 3), "cores": cores, "kind": "port",
            "sample": what + " (C++ restatement of go-fft's CPU path with parallel.go's line fan-out), "
                      "best of %d; host %s" % (reps, os.uname().machine)}


def _err(y, ref, what="the FP64 GPU path, same input"):
    """Relative L2 and largest absolute error of y against ref (both on the
    device), in FP64."""
    d = y.reshape(-1).to(ref.dtype) - ref.reshape(-1)
    return {"rel_l2": float("%.3g" % (d.norm().item() / max(ref.norm().item(), 1e-300))),
            "max_abs": float("%.3g" % d.abs().max().item()), "reference": what}


def _sinusoids(torch, dev, n, dtype, seed=0):
    """Sum of 64 sinusoids (f ~ U[0, 0.5) cycles/sample, amplitude ~ U[0, 1],
    phase 0) plus N(0, 0.1^2) noise, built on the GPU in FP64."""
    rng = np.random.default_rng(seed)
    f = rng.uniform(0, 0.5, 64)
    a = rng.uniform(0, 1, 64)
    t = torch.arange(n, dtype=torch.float64, device=dev)
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    for fj, aj in zip(f, a):
        x += aj * torch.sin((2.0 * math.pi * fj) * t)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    x += 0.1 * torch.randn(n, dtype=torch.float64, device=dev, generator=g)
    del t
    return x.to(dtype)


def workload_roundtrip4096(D, torch, dev, steps, warmup):
    n, batch = 4096, 256
    x = torch.from_numpy(gaussian_complex(n * batch, 0)).to(dev)
    p = D.C2CPlan(n, batch, torch.complex128)
    y = torch.empty_like(x)
    z = torch.empty_like(x)

    def step():
        p.exec(x, out=y)
        p.exec(y, out=z, inverse=True)
    ms = time_events(step, steps, warmup, torch)
    _line("roundtrip4096", ms, 2 * batch * 5.0 * n * 12, 4 * n * batch * 16, "c128",
          "FFT + IFFT, N=4096, batch 256, complex128 (64 MiB in total, resident in the 256 MB MALL)",
          {"roundtrip_err": _err(z, x, "the input (IFFT(FFT(x)) = x)")})
    # The same step from HBM: before each timed step a 1 GiB fill (outside the
    # timed events) evicts the data from the memory-side cache.
    flush = torch.empty(1 << 28, dtype=torch.float32, device=dev)
    times = []
    for i in range(warmup + steps):
        flush.fill_(float(i))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        step()
        e1.record()
        torch.cuda.synchronize()
        if i >= warmup:
            times.append(e0.elapsed_time(e1))
    del flush
    _line("roundtrip4096_cold", sum(times) / len(times), 2 * batch * 5.0 * n * 12, 4 * n * batch * 16, "c128",
          "FFT + IFFT, N=4096, batch 256, complex128, each step timed alone after a 1 GiB fill that evicts "
          "its data from the memory-side cache")


def workload_spectral(D, torch, dev, steps, warmup, cpu=True):
    n, seg, ov = 1 << 26, 1024, 512
    win = [0.5 - 0.5 * math.cos(2 * math.pi * k / (seg - 1)) for k in range(seg)]
    x64 = _sinusoids(torch, dev, n, torch.float64)
    f64 = torch.float64
    for name, rt in (("f32", torch.float32), ("f64", torch.float64)):
        x = x64 if rt == f64 else x64.to(rt)

        def vs64(y, run64):
            """FP32 lines: the error against the FP64 plan on the same signal
            (untimed)."""
            return {} if rt == f64 else {"err_vs_fp64": _err(y, run64())}
        eb = 4 if rt == torch.float32 else 8
        sp = D.SpectrogramPlan(n, seg, ov, win, 1.0, rt)
        out = torch.empty(sp.frames * sp.bins, dtype=rt, device=dev)
        ms = time_events(lambda: sp.exec(x, out=out), steps, warmup, torch)
        extra = vs64(out, lambda: D.SpectrogramPlan(n, seg, ov, win, 1.0, f64).exec(x64))
        if cpu and rt == f64:
            from oracle import oracle as O
            xh = x.cpu().numpy()
            extra["cpu_baseline"] = _cpu_baseline(
                lambda c: O.spectrogram(xh, seg, ov, np.asarray(win), 1.0, threads=c),
                "the same Spectrogram, float64")
            del xh
        _line("spectrogram", ms, sp.frames * 2.5 * seg * 10, (n + sp.frames * sp.bins) * eb, name,
              "Spectrogram of 2^26 samples, Hann(1024), overlap 512: %d frames x 513 bins" % sp.frames, extra)
        del sp, out
        pp = D.PSDPlan(n, 1, 1.0, rt)
        out = torch.empty(n // 2 + 1, dtype=rt, device=dev)
        ms = time_events(lambda: pp.exec(x, out=out), steps, warmup, torch)
        extra = vs64(out, lambda: D.PSDPlan(n, 1, 1.0, f64).exec(x64))
        _line("psd", ms, 2.5 * n * 26, (n + n // 2 + 1) * eb, name, "one-sided PSD of 2^26 samples", extra)
        del pp, out
        m = 1 << 24
        rp = D.R2CPlan(m, 1, rt)
        xs = x[:m].contiguous()
        X = rp.forward(xs)
        ms = time_events(lambda: rp.forward(xs, out=X), steps, warmup, torch)
        extra = vs64(X, lambda: D.R2CPlan(m, 1, f64).forward(x64[:m].contiguous()))
        _line("rfft", ms, 2.5 * m * 24, m * eb + (m // 2 + 1) * 2 * eb, name, "RFFT of 2^24 samples", extra)
        back = torch.empty_like(xs)
        ms = time_events(lambda: rp.inverse(X, out=back), steps, warmup, torch)
        # the FP32 inverse of the FP32 spectrum against the FP64 inverse of
        # the same (widened) spectrum
        extra = vs64(back, lambda: D.R2CPlan(m, 1, f64).inverse(X.to(torch.complex128)))
        _line("irfft", ms, 2.5 * m * 24, m * eb + (m // 2 + 1) * 2 * eb, name, "IRFFT to 2^24 samples", extra)
        del rp, X, back, xs, x
        torch.cuda.empty_cache()
    del x64


def workload_nd(D, torch, dev, steps, warmup, cpu=True):
    L = D.L
    c64, c128 = torch.complex64, torch.complex128
    for shape, dts in (((4096, 4096), (c64, c128)), ((256, 256, 256), (c64,))):
        N = int(np.prod(shape))
        g = torch.Generator(device=dev)
        g.manual_seed(0)
        x128 = torch.randn(N, dtype=c128, device=dev, generator=g)
        for ct in dts:
            x = x128 if ct == c128 else x128.to(ct)
            p = D.NDPlan(shape, L.ND_C2C, 1, ct)
            y = torch.empty_like(x)
            ms = time_events(lambda: p.exec(x, out=y), steps, warmup, torch)
            eb = 8 if ct == c64 else 16
            extra = {}
            if ct == c64:  # against the FP64 plan on the same input (untimed)
                extra["err_vs_fp64"] = _err(y, D.NDPlan(shape, L.ND_C2C, 1, c128).exec(x128))
            if cpu and ct == c128 and len(shape) == 2:
                from oracle import oracle as O
                xh = x.cpu().numpy()
                extra["cpu_baseline"] = _cpu_baseline(lambda c: O.fftn(xh, list(shape)),
                                                      "the same FFT2, complex128")
                del xh
            _line("fft2" if len(shape) == 2 else "fftn", ms, 5.0 * N * math.log2(N), 2 * N * eb,
                  "c64" if eb == 8 else "c128",
                  "x".join(map(str, shape)) + " complex normal (torch.randn, re/im N(0, 1/2))", extra)
            del p, x, y
            torch.cuda.empty_cache()
        del x128
    rows = cols = 8192
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    x = torch.rand(rows * cols, dtype=torch.float32, device=dev, generator=g)
    p = D.NDPlan((rows, cols), L.ND_R2C, 1, c64)
    X = p.rfft2(x)
    ms = time_events(lambda: p.rfft2(x, out=X), steps, warmup, torch)
    nb = rows * cols * 4 + rows * (cols // 2 + 1) * 8
    p64 = D.NDPlan((rows, cols), L.ND_R2C, 1, c128)  # FP64 references on the same inputs (untimed)
    _line("rfft2", ms, 2.5 * rows * cols * 26, nb, "f32", "8192x8192 float32 U[0,1)",
          {"err_vs_fp64": _err(X, p64.rfft2(x.double()))})
    back = torch.empty_like(x)
    ms = time_events(lambda: p.irfft2(X, out=back), steps, warmup, torch)
    _line("irfft2", ms, 2.5 * rows * cols * 26, nb, "f32", "8192x8192 -> 8192x4097 bins and back",
          {"err_vs_fp64": _err(back, p64.irfft2(X.to(c128)))})


def workload_blur(D, torch, dev, steps, warmup):
    L = D.L
    batch, rows, cols = 64, 1024, 1024
    rc = cols // 2 + 1
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    x = torch.rand(batch * rows * cols, dtype=torch.float32, device=dev, generator=g)
    fr = torch.fft.fftfreq(rows, device=dev, dtype=torch.float64)[:, None]
    fc = torch.fft.rfftfreq(cols, device=dev, dtype=torch.float64)[None, :]
    H = torch.exp(-2.0 * (math.pi * 3.0) ** 2 * (fr ** 2 + fc ** 2)).to(torch.float32).reshape(-1)
    H = H.repeat(batch)
    p = D.NDPlan((rows, cols), L.ND_R2C, batch, torch.complex64)
    y = p.irfft2(p.rfft2(x) * H) + 0.01 * torch.randn(batch * rows * cols, device=dev, generator=g)
    S = torch.empty(batch * rows * rc, dtype=torch.complex64, device=dev)
    r = torch.empty_like(x)
    gS = torch.empty_like(S)
    gx = torch.empty_like(x)

    def step():
        p.rfft2(x, out=S)
        S.mul_(H)
        p.irfft2(S, out=r)
        r.sub_(y)
        r.mul_(2.0)
        p.vjp(r, gS, direction=1)
        gS.mul_(H)
        p.vjp(gS, gx, direction=0)
    ms = time_events(step, steps, warmup, torch)
    # The same gradient through FP64 plans on the same x, H and y (untimed).
    # The residual subtracts nearly equal values, which scales FP32 rounding
    # up, so the adjoint transforms are also compared on their own: both
    # precisions applied to the FP32 step's residual r.
    p64 = D.NDPlan((rows, cols), L.ND_R2C, batch, torch.complex128)
    H64 = H.double()

    def grad64(res):
        gS64 = p64.vjp(res, torch.empty_like(S, dtype=torch.complex128), direction=1)
        return p64.vjp(gS64 * H64, torch.empty_like(res), direction=0)
    err = _err(gx, grad64(2.0 * (p64.irfft2(p64.rfft2(x.double()) * H64) - y.double())))
    err_adj = _err(gx, grad64(r.double()), "the FP64 adjoint transforms on the FP32 step's residual")
    del p64, H64
    per = 2.5 * rows * cols * 20 * batch
    passb = batch * (rows * cols * 4 + rows * rc * 8)
    _line("blur_grad", ms, 4 * per, 4 * passb, "f32",
          "loss ||IRFFT2(RFFT2(x) H) - y||^2 and its gradient, 64 x 1024x1024 float32, Gaussian H sigma 3 px; "
          "one step = RFFT2, IRFFT2 and the two adjoint transforms plus elementwise work",
          {"err_vs_fp64": err, "adjoint_err_vs_fp64": err_adj})
    # The same loss and gradient in one call (gofft_filter2_loss_grad): the
    # filter multiplies, residual and loss sums run inside, and the spectrum
    # is recomputed per chunk of arrays instead of kept (here one chunk).
    Hc = H[: rows * rc].to(torch.complex64)
    ms_f = time_events(lambda: p.filter_loss_grad(x, Hc, y, chunk=0), steps, warmup, torch)
    _, g_f = p.filter_loss_grad(x, Hc, y, chunk=0)
    _line("blur_grad_fused", ms_f, 4 * per, 4 * passb, "f32",
          "the same loss and gradient through gofft_filter2_loss_grad (one call per step)",
          {"vs_composed_step": _err(g_f, gx, "the composed step's gradient, same input")})


def workload_groups(D, torch, dev, steps, warmup, lengths=None, precisions=("c64", "c128")):
    """Each length group of the mixed-length batch alone: a plan of exactly
    that group's lines (the lengths and inputs of the headline step), one
    line per group and precision. For focused timing and kernel profiles."""
    all_lengths = workload_lengths(0)
    x64 = gaussian_complex(sum(all_lengths), seed=0)
    for n in sorted(set(all_lengths)) if lengths is None else lengths:
        count = sum(1 for v in all_lengths if v == n)
        for prec in precisions:
            ct, eb = (torch.complex64, 8) if prec == "c64" else (torch.complex128, 16)
            p = D.C2CPlan(n, count, ct)
            x = torch.from_numpy(x64[: n * count].astype(np.complex64 if prec == "c64" else np.complex128)).to(dev)
            y = torch.empty_like(x)
            ms = time_events(lambda: p.exec(x, out=y), steps, warmup, torch)
            _line("group", ms, count * 5.0 * n * math.log2(n), 2 * n * count * eb, prec,
                  "%d transforms of length %d (the mixed batch's group), one plan" % (count, n),
                  {"plan": p.describe().strip()})
            del p, x, y


def workload_twiddles(D, torch, dev, steps, warmup):
    """The complex64 pass-twiddle forms side by side: FP64-formed (the
    default), plain FP32 tables and products, and hi + lo float pairs in
    double-float arithmetic, on the transforms whose passes carry twiddles
    (FFT2 4096^2 has none: both of its axes are single 4096-point passes).
    Error against the FP64 GPU path on the same input."""
    L = D.L
    modes = (("fp64_formed", 0), ("fp32", L.FLAG_TWIDDLES_FP32), ("split", L.FLAG_TWIDDLES_SPLIT))
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    cases = []
    x = torch.randn(16 << 20, dtype=torch.complex128, device=dev, generator=g)
    cases.append(("fft 2^20 x 16", lambda f: D.C2CPlan(1 << 20, 16, torch.complex64, f),
                  lambda p, v: p.exec(v), x, 16 * 5.0 * (1 << 20) * 20, 2 * (16 << 20) * 8))
    cases.append(("fft2 4096x4096", lambda f: D.NDPlan((4096, 4096), L.ND_C2C, 1, torch.complex64, f),
                  lambda p, v: p.exec(v), x, 5.0 * (1 << 24) * 24, 2 * (1 << 24) * 8))
    r = torch.rand(1 << 26, dtype=torch.float64, device=dev, generator=g)
    cases.append(("rfft2 8192x8192 U[0,1)", lambda f: D.NDPlan((8192, 8192), L.ND_R2C, 1, torch.complex64, f),
                  lambda p, v: p.rfft2(v), r, 2.5 * (1 << 26) * 26, (1 << 26) * 4 + 8192 * 4097 * 8))
    sig = _sinusoids(torch, dev, 1 << 26, torch.float64)
    cases.append(("rfft 2^24 sinusoids", lambda f: D.R2CPlan(1 << 24, 1, torch.float32, f),
                  lambda p, v: p.forward(v), sig[: 1 << 24].contiguous(), 2.5 * (1 << 24) * 24,
                  (1 << 24) * 4 + ((1 << 23) + 1) * 8))
    cases.append(("psd 2^26 sinusoids", lambda f: D.PSDPlan(1 << 26, 1, 1.0, torch.float32, f),
                  lambda p, v: p.exec(v), sig, 2.5 * (1 << 26) * 26, (1 << 26) * 4 + ((1 << 25) + 1) * 4))
    for name, mk, run, x64, flops, nbytes in cases:
        xin = x64.to(torch.complex64 if x64.is_complex() else torch.float32)
        if name.startswith("fft 2^20"):
            xin = xin[: 16 << 20].contiguous()
        elif name.startswith("fft2"):
            xin = xin[: 1 << 24].contiguous()
        # the FP64 reference: the same transform in complex128 / float64 on
        # the rounded input, so that the error is the FP32 arithmetic's
        xw = xin.to(torch.complex128 if xin.is_complex() else torch.float64)
        if name.startswith("fft 2^20"):
            ref = D.C2CPlan(1 << 20, 16, torch.complex128).exec(xw)
        elif name.startswith("fft2"):
            ref = D.NDPlan((4096, 4096), L.ND_C2C, 1, torch.complex128).exec(xw)
        elif name.startswith("rfft2"):
            ref = D.NDPlan((8192, 8192), L.ND_R2C, 1, torch.complex128).rfft2(xw)
        elif name.startswith("rfft"):
            ref = D.R2CPlan(1 << 24, 1, torch.float64).forward(xw)
        else:
            ref = D.PSDPlan(1 << 26, 1, 1.0, torch.float64).exec(xw)
        del xw
        for mname, flag in modes:
            p = mk(flag)
            y = run(p, xin)
            ms = time_events(lambda: run(p, xin), steps, warmup, torch)
            _line("twiddles", ms, flops, nbytes, "c64", name, {"twiddle_form": mname, "err_vs_fp64": _err(y, ref)})
            del p, y
        del ref, xin
        torch.cuda.empty_cache()


def workload_host(D, torch, dev, steps, warmup):
    """The host-staged path a Go binding calls (Plan.FFT on host slices):
    FFT + IFFT of 256 x 4096 complex128 from and to ordinary pageable host
    memory through gofft_exec_c2c_host (pinned staging, the call's own
    stream), beside the same round trip on device-resident data. Wall-clock
    per round trip (the host calls synchronise themselves); the data rate
    counts the API bytes (input and output of both calls)."""
    import ctypes
    L = D.L
    n, batch = 4096, 256
    x = gaussian_complex(n * batch, 0)
    y = np.empty_like(x)
    z = np.empty_like(x)
    p = D.C2CPlan(n, batch, torch.complex128)
    lib = L.lib()
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)

    def host_step():
        L.check(lib.gofft_exec_c2c_host(p.handle, vp(x), vp(y), L.FORWARD))
        L.check(lib.gofft_exec_c2c_host(p.handle, vp(y), vp(z), L.INVERSE))
    for _ in range(warmup):
        host_step()
    t0 = time.perf_counter()
    for _ in range(steps):
        host_step()
    ms_host = (time.perf_counter() - t0) * 1e3 / steps
    xd = torch.from_numpy(x).to(dev)
    yd = torch.empty_like(xd)
    zd = torch.empty_like(xd)

    def dev_step():
        p.exec(xd, out=yd)
        p.exec(yd, out=zd, inverse=True)
    ms_dev = time_events(dev_step, steps, warmup, torch)
    err = float(np.linalg.norm(z - x) / np.linalg.norm(x))
    nbytes = 4 * n * batch * 16
    _line("host_roundtrip4096", ms_host, 2 * batch * 5.0 * n * 12, nbytes, "c128",
          "FFT + IFFT, N=4096, batch 256, complex128, host slices in and out (gofft_exec_c2c_host: "
          "PCIe both ways, pinned staging)",
          {"device_resident_ms": round(ms_dev, 4), "host_over_device": round(ms_host / ms_dev, 2),
           "roundtrip_err": {"rel_l2": float("%.3g" % err)}})


def workload_dft16(D, torch, dev, steps, warmup):
    """One radix-16 stage both ways (gofft_probe_dft16): VALU butterflies, one
    transform per lane, against the dense DFT-matrix product on the matrix
    cores (a 32 x 32 real matrix, sixteen v_mfma_f32_32x32x2f32 per 32
    transforms), on 2^22 complex64 lines of 16 points held in registers and
    transformed `reps` times between one load and one store. reps = 1 is
    bound by memory; reps = 64 by the arithmetic. Rate: stages per second (a
    stage = one 16-point DFT of one line); the matrix form does 2048 flops per
    stage, the butterflies about 144."""
    import ctypes
    L = D.L
    n = 1 << 22
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    x = torch.randn(n * 16, dtype=torch.complex64, device=dev, generator=g)
    y = torch.empty_like(x)
    lib = L.lib()
    s = torch.cuda.current_stream()
    for reps in (1, 16, 64):
        res = {}
        for method, name in ((0, "valu"), (1, "mfma")):
            def run():
                L.check(lib.gofft_probe_dft16(method, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                              n, reps, ctypes.c_void_p(s.cuda_stream)))
            ms = time_events(run, steps, warmup, torch)
            res[name] = ms
        print(json.dumps({"workload": "dft16_stage", "reps": reps, "lines": n,
                          "valu_ms": round(res["valu"], 4), "mfma_ms": round(res["mfma"], 4),
                          "valu_gstages_s": round(n * reps / (res["valu"] * 1e6), 1),
                          "mfma_gstages_s": round(n * reps / (res["mfma"] * 1e6), 1),
                          "mfma_tflops": round(n * reps * 2048 / (res["mfma"] * 1e9), 1),
                          "mfma_over_valu_time": round(res["mfma"] / res["valu"], 2)}), flush=True)


WORKLOADS = {"host": workload_host, "dft16": workload_dft16, "roundtrip4096": workload_roundtrip4096, "spectral": workload_spectral, "nd": workload_nd,
             "blur": workload_blur, "groups": workload_groups, "twiddles": workload_twiddles}


def summarize_trace(path, skip=None, take=None):
    """Text summary of rocprofv3 --stats kernel statistics (CSV); with SKIP
    and TAKE, also the average duration of k_pass<float, ...> dispatches
    [SKIP, SKIP + TAKE) from the kernel trace beside it."""
    rows = list(csv.DictReader(open(path)))
    print("%-62s %7s %12s %12s %7s" % ("kernel", "calls", "total_ms", "avg_us", "pct"))
    for r in rows:
        name = r["Name"]
        if len(name) > 62:
            name = name[:59] + "..."
        print("%-62s %7s %12.3f %12.2f %7.2f" % (name, r["Calls"], float(r["TotalDurationNs"]) / 1e6,
                                                 float(r["AverageNs"]) / 1e3, float(r["Percentage"])))
    if skip is None:
        return
    trace = os.path.join(os.path.dirname(path), os.path.basename(path).replace("kernel_stats", "kernel_trace"))
    disp = [r for r in csv.DictReader(open(trace)) if "k_pass<float" in r["Kernel_Name"]]
    disp.sort(key=lambda r: int(r["Dispatch_Id"]))
    win = disp[skip: skip + take]
    dur = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3 for r in win]
    print()
    print("k_pass<float, 16, *> dispatches %d..%d of %d (the per-group runs): average %.2f us"
          % (skip, skip + len(win), len(disp), sum(dur) / max(1, len(dur))))


def _pmc_rows(path, counter):
    return [r for r in csv.DictReader(open(path)) if r.get("Counter_Name", counter) == counter]


def _mean(rows):
    return sum(float(r["Counter_Value"]) for r in rows) / len(rows), len(rows)


def _calibration(rows, path):
    """Mean counter value of the 1 GiB device copy's dispatches: the copy
    dispatches that read (write) the most, within 10% of the largest (other
    copies of the run are smaller, some with larger grids)."""
    copies = [r for r in rows if "copy" in r["Kernel_Name"].lower()]
    if not copies:
        raise SystemExit("no copy dispatch in %s (run the workload with --calibrate)" % path)
    top = max(float(r["Counter_Value"]) for r in copies)
    return _mean([r for r in copies if float(r["Counter_Value"]) >= 0.9 * top])


def summarize_pmc(fetch_csv, write_csv):
    """FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the pass kernel, each
    scaled by what the counter reads for the 1 GiB device copy of the same run."""
    out = {}
    for name, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = _pmc_rows(path, name)
        cal, ncal = _calibration(rows, path)
        kern = [r for r in rows if "k_pass<float" in r["Kernel_Name"]]
        if not kern:
            raise SystemExit("no pass-kernel dispatch in %s" % path)
        raw, nk = _mean(kern)
        factor = COPY_BYTES / (cal * 1024.0)
        out[name] = (raw, nk, cal, ncal, factor, raw * 1024.0 * factor)
    print("kernel: k_pass<float, 16, *> (every form of the complex64 pass kernel in the run)")
    for name, (raw, nk, cal, ncal, factor, b) in out.items():
        print("%s: %d pass-kernel dispatches, mean %.1f KiB raw; 1 GiB copy reads %.1f KiB "
              "(%d dispatches) -> factor %.4f -> %.0f bytes per launch" % (name, nk, raw, cal, ncal, factor, b))
    print("traffic_bytes_per_launch: %.0f" % (out["FETCH_SIZE"][5] + out["WRITE_SIZE"][5]))


def _kernel_key(name):
    """A kernel's name without its argument list (template arguments kept)."""
    name = name.strip()
    if name.endswith(")") and "(" in name:
        depth = 0
        for i in range(len(name) - 1, -1, -1):
            depth += {")": 1, "(": -1}.get(name[i], 0)
            if depth == 0:
                return name[:i].replace("void ", "", 1).strip()
    return name


def summarize_kernels(stats_csv, fetch_csv=None, write_csv=None):
    """Per kernel of a profiled run: dispatches and average duration
    (rocprofv3 --kernel-trace --stats), and with the two counter runs the
    HBM bytes per dispatch, FETCH_SIZE and WRITE_SIZE each scaled by what the
    counter reads for the run's 1 GiB device copy (bench.py --calibrate)."""
    rows = list(csv.DictReader(open(stats_csv)))
    traffic = {}
    factors = {}
    for cname, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        if not path:
            continue
        prow = _pmc_rows(path, cname)
        cal, _ = _calibration(prow, path)
        factors[cname] = COPY_BYTES / (cal * 1024.0)
        per = {}
        for r in prow:
            per.setdefault(_kernel_key(r["Kernel_Name"]), []).append(float(r["Counter_Value"]))
        for k, v in per.items():
            traffic.setdefault(k, {})[cname] = sum(v) / len(v) * 1024.0 * factors[cname]
    if factors:
        print("calibration (1 GiB copy): " + ", ".join("%s factor %.4f" % kv for kv in factors.items()))
    print("%-58s %6s %10s %10s %14s %14s" % ("kernel", "calls", "total_ms", "avg_us", "fetch_B/disp", "write_B/disp"))
    for r in rows:
        k = _kernel_key(r["Name"])
        t = traffic.get(k, {})
        name = k if len(k) <= 58 else k[:55] + "..."
        f = "%14.0f" % t["FETCH_SIZE"] if "FETCH_SIZE" in t else "%14s" % "-"
        w = "%14.0f" % t["WRITE_SIZE"] if "WRITE_SIZE" in t else "%14s" % "-"
        print("%-58s %6s %10.3f %10.2f %s %s" % (name, r["Calls"], float(r["TotalDurationNs"]) / 1e6,
                                                float(r["AverageNs"]) / 1e3, f, w))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--workload", choices=["mixed", "all"] + sorted(WORKLOADS), default="mixed",
                    help="mixed: the headline line; the others print one line per transform")
    ap.add_argument("--lengths", help="groups workload: comma-separated lengths (default: all seven)")
    ap.add_argument("--precisions", default="c64,c128", help="groups workload: c64 and/or c128")
    ap.add_argument("--summarize-trace", nargs="+", metavar="STATS_CSV [SKIP TAKE]")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"))
    ap.add_argument("--summarize-kernels", nargs="+", metavar="STATS_CSV [FETCH_CSV WRITE_CSV]")
    ap.add_argument("--calibrate", action="store_true",
                    help="other workloads: also run the 1 GiB device copy that counter summaries scale by")
    args = ap.parse_args()
    if args.summarize_kernels:
        return summarize_kernels(*args.summarize_kernels[:3])
    if args.summarize_trace:
        t = args.summarize_trace
        return summarize_trace(t[0], *(int(v) for v in t[1:3])) if len(t) >= 3 else summarize_trace(t[0])
    if args.summarize_pmc:
        return summarize_pmc(*args.summarize_pmc)
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    from fft_amd import device as D

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.workload != "mixed":
        # Each of these lines times a freshly built plan on freshly allocated
        # buffers, whose first few dozen calls run up to 7% slower than the
        # steady state (measured: 20-call means 0.517, 0.488, 0.482, 0.482 ms
        # for the float64 Spectrogram after 3 warm-up calls), so they warm
        # up at least 20 calls.
        warm = max(args.warmup, 20)
        for name in (sorted(WORKLOADS) if args.workload == "all" else [args.workload]):
            fn = WORKLOADS[name]
            if name in ("spectral", "nd"):
                fn(D, torch, dev, args.steps, warm, cpu=not args.no_cpu)
            elif name == "groups":
                lens = [int(v) for v in args.lengths.split(",")] if args.lengths else None
                fn(D, torch, dev, args.steps, warm, lens, args.precisions.split(","))
            else:
                fn(D, torch, dev, args.steps, warm)
        if args.calibrate:
            print(json.dumps({"workload": "copy_1GiB", **copy_bandwidth(torch, dev, 3, 1)}), flush=True)
        return
    lengths = workload_lengths(0)
    total = sum(lengths)
    flops = flops_of(lengths)
    x64 = gaussian_complex(total, seed=0)

    r32 = run_dtype(D, torch, dev, "fp32", lengths, x64, args.steps, args.warmup)
    r64 = run_dtype(D, torch, dev, "fp64", lengths, x64, args.steps, args.warmup)
    roofline = pass_roofline(r32, D.L, args.steps, args.warmup)
    copy = copy_bandwidth(torch, dev, args.steps, args.warmup)
    cpu, ref = (None, None) if args.no_cpu else cpu_leg(lengths, x64)
    acc = accuracy(lengths, r32.pop("out"), r64.pop("out"), ref)
    del ref

    line = {
        "metric": "FFT GFLOP/s (5N*log2N) & HBM GB/s: FP32 batched vs FP64 plain, 1 GPU",
        "value": round(r32["gflops"], 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms"], 4),
        "ms_per_step_eager": round(r32["ms_eager"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (f32 butterflies, FP64-generated tables)",
        "data": "synthetic: re/im iid N(0,1) from numpy default_rng(0)",
        "config": {
            "workload": "mixed-length batched complex FFT: 4096 transforms, lengths uniform from "
                        "{1024,4096,65536,6720,1009,10007,20014}, one plan, inputs resident in HBM; "
                        "a step is one forward exec of the plan, replayed as its captured graph "
                        "(ms_per_step_eager: the same exec launched directly)",
            "transforms": COUNT, "points": total, "flops_per_step": flops,
            "parallelism": "1 GPU",
        },
        "hbm_gbs": round(r32["gbs"], 1),
        "fp64_plain": {"gflops": round(r64["gflops"], 2), "gbs": round(r64["gbs"], 1),
                       "ms_per_step": round(r64["ms"], 4), "ms_per_step_eager": round(r64["ms_eager"], 4)},
        "fp32_over_fp64": round(r32["gflops"] / r64["gflops"], 3),
        "launches_per_step": {"fp32": r32["launches"], "fp64": r64["launches"]},
        "groups_ms": {"fp32": {str(n): round(g["ms"], 4) for n, g in r32["groups"].items()},
                      "fp64": {str(n): round(g["ms"], 4) for n, g in r64["groups"].items()}},
        "copy_1GiB": copy,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": acc,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
