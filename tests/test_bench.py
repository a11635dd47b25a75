"""Host logic of bench.py (no GPU): the one-GPU refusal, the workload's
lengths, the accuracy figures, the pass-kernel roofline arithmetic and the
counter-summary calibration."""
import os
import subprocess
import sys
import types

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_workload_lengths():
    a = bench.workload_lengths(0)
    assert a == bench.workload_lengths(0) and len(a) == bench.COUNT
    assert set(a) == set(bench.LENGTHS)
    # uniform draw: every length holds about 1/7 of the transforms
    counts = [a.count(n) for n in bench.LENGTHS]
    assert min(counts) > 0.8 * bench.COUNT / 7 and max(counts) < 1.2 * bench.COUNT / 7
    assert bench.flops_of([1024]) == 5.0 * 1024 * 10


def test_accuracy_takes_the_worst_transform_of_each_length():
    lengths = [8, 16, 8, 8]
    rng = np.random.default_rng(3)
    ref = [rng.standard_normal(n) + 1j * rng.standard_normal(n) for n in lengths]
    flat = np.concatenate(ref)
    y64 = flat.copy()
    y32 = flat.astype(np.complex64)
    # plant an error of known size in the third transform (length 8)
    off = 8 + 16
    y64[off + 5] += 1e-13
    y32[off + 2] += np.complex64(1e-3)
    acc = bench.accuracy(lengths, y32, y64, ref)
    row8, row16 = acc["per_length"]["8"], acc["per_length"]["16"]
    assert row8["transforms"] == 3 and row16["transforms"] == 1
    want32 = np.linalg.norm(y32[off:off + 8].astype(np.complex128) - ref[2]) / np.linalg.norm(ref[2])
    assert row8["c64_rel_l2"] == pytest.approx(want32, rel=1e-2)
    assert row8["c64_max_abs"] == pytest.approx(1e-3, rel=1e-3)
    assert row8["c128_max_abs"] == pytest.approx(1e-13, rel=1e-2)
    assert row16["c64_rel_l2"] < 1e-6 and row16["c128_rel_l2"] == 0.0
    assert row8["c128_bar"] == pytest.approx(3e-14)
    # without the oracle's outputs, FP32 against the FP64 output only
    acc2 = bench.accuracy(lengths, y32, y64, None)
    assert "c128_rel_l2" not in acc2["per_length"]["8"]
    assert acc2["per_length"]["8"]["c64_max_abs"] > 9e-4


def _groups(spec):
    L = types.SimpleNamespace(KERNEL_PASS=0, KERNEL_CONV=1, KERNEL_REAL=2, KERNEL_AUX=3)
    groups = {}
    for n, (ms, pl, pb, conv) in spec.items():
        groups[n] = {"ms": ms, "work": {0: (pl, pb), 1: (conv, 0), 2: (0, 0), 3: (0, 0)}}
    return L, groups


def test_pass_roofline_uses_pass_only_groups():
    # two pass-only groups (3 launches, 0.3 ms, 3 GB) and one convolution group
    L, groups = _groups({1024: (0.1, 1, 1e9, 0), 65536: (0.2, 2, 2e9, 0), 1009: (5.0, 0, 0, 1)})
    r = {"groups": groups, "work": {0: (3, 3e9)}}
    roof = bench.pass_roofline(r, L, steps=20, warmup=3)
    assert roof["launches_per_step"] == 3
    assert roof["avg_launch_ms"] == pytest.approx(0.1)
    assert roof["achieved"] == pytest.approx(1e9 / 0.1e6, rel=1e-6)  # 10000 GB/s
    assert roof["frac"] == pytest.approx(roof["achieved"] / bench.HBM_PEAK_GBS, rel=1e-3)
    # eager and graph-replayed steps (3 + 20 calls each), then each group's
    # 20 + max(3, 20) replays
    assert roof["trace_window"] == {"skip": 2 * 23 * 3, "take": (20 + bench.GROUP_WARMUP) * 3}
    # a group mixing pass and convolution launches is left out, and then the
    # launch counts no longer agree with the step's
    L, groups = _groups({1024: (0.1, 1, 1e9, 0), 20014: (0.3, 2, 2e9, 1)})
    with pytest.raises(RuntimeError):
        bench.pass_roofline({"groups": groups, "work": {0: (3, 3e9)}}, L, 20, 3)


def _pmc_csv(path, counter, rows):
    with open(path, "w") as f:
        f.write("Kernel_Name,Grid_Size,Counter_Name,Counter_Value\n")
        for k, g, v in rows:
            f.write('"%s",%d,%s,%r\n' % (k, g, counter, v))


def test_summarize_pmc_calibrates_on_the_copy(tmp_path, capsys):
    fetch, write = tmp_path / "f.csv", tmp_path / "w.csv"
    gib_kib = bench.COPY_BYTES / 1024.0
    # the counter reads the copy at half its size (factor 2) and a smaller
    # copy that must not be used for calibration
    _pmc_csv(fetch, "FETCH_SIZE", [("__amd_rocclr_copyBuffer", 1 << 20, gib_kib / 2),
                                    ("__amd_rocclr_copyBuffer", 1 << 10, 7.0),
                                    ("void gofft::k_pass<float, 16, 0>(gofft::PassArgs)", 4096, 1000.0),
                                    ("void gofft::k_pass<float, 16, 2>(gofft::PassArgs)", 4096, 3000.0),
                                    ("void gofft::k_pass<double, 16, 0>(gofft::PassArgs)", 4096, 9e9)])
    _pmc_csv(write, "WRITE_SIZE", [("__amd_rocclr_copyBuffer", 1 << 20, gib_kib),
                                    ("void gofft::k_pass<float, 16, 0>(gofft::PassArgs)", 4096, 500.0)])
    bench.summarize_pmc(str(fetch), str(write))
    out = capsys.readouterr().out
    total = [line for line in out.splitlines() if line.startswith("traffic_bytes_per_launch:")]
    # fetch: mean 2000 KiB x 2; write: 500 KiB x 1
    assert float(total[0].split(":")[1]) == pytest.approx((2000 * 2 + 500) * 1024.0)
