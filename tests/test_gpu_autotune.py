"""GPU: plans built with the autotuner (GOFFT_FLAG_AUTOTUNE) keep parity, are
marked tuned, and record their choices in the GOFFT_TUNE_CACHE file, which a
later process reuses instead of timing again."""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import cgauss, rel_l2

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("n,dt", [(10007, "c64"), (20014, "c64"), (65536, "c64"), (9973, "c128"),
                                  (1 << 20, "c128")])
def test_autotuned_plans_keep_parity(oracle, n, dt):
    import torch
    from fft_amd import _lib as L
    from fft_amd import device as D
    batch = 3
    x = cgauss(n * batch, seed=n).reshape(batch, n)
    tdt = torch.complex64 if dt == "c64" else torch.complex128
    p = D.C2CPlan(n, batch, tdt, L.FLAG_AUTOTUNE)
    assert "tuned" in p.describe(), p.describe()
    xt = torch.from_numpy(x.astype(np.complex64 if dt == "c64" else np.complex128).ravel()).to("cuda")
    got = p.exec(xt).cpu().numpy().reshape(batch, n)
    bar = 1e-6 if dt == "c64" else 1e-14 * math.log2(n)
    assert rel_l2(got, oracle.fft_batch(x)) <= bar


def test_tune_cache_file(tmp_path):
    cache = tmp_path / "tune.txt"
    env = dict(os.environ, GOFFT_AUTOTUNE="1", GOFFT_TUNE_CACHE=str(cache))
    code = ("import sys; sys.path.insert(0, %r); import torch; from fft_amd import device as D; "
            "p = D.C2CPlan(10007, 64, torch.complex64); print(p.describe())" % ROOT)
    out1 = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out1.returncode == 0, out1.stderr[-2000:]
    lines1 = cache.read_text().split("\n")
    # tag, library version, device architecture, route, n, lines, precision, m, columns
    rows = [l.split() for l in lines1 if l.strip()]
    assert any(r[0] == "gofft-tune-v2" and r[3:7] == ["3", "10007", "64", "1"] for r in rows), lines1
    out2 = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out2.returncode == 0, out2.stderr[-2000:]
    assert cache.read_text().split("\n") == lines1  # reused, not timed again
    assert "tuned" in out2.stdout


def test_corrupt_tune_cache_is_ignored(tmp_path):
    """Entries the planner could not have made (a convolution length below 2q,
    one outside the cost model's window, unknown columns, another library
    version, garbage) are dropped: the plan is made afresh and keeps parity."""
    import torch
    arch = torch.cuda.get_device_properties(0).gcnArchName.replace(" ", "_")
    cache = tmp_path / "tune.txt"
    cache.write_text("\n".join([
        "gofft-tune-v2 0.2 %s 3 10007 4 1 1000 16" % arch,     # cl < 2q: would overflow the kernel table
        "gofft-tune-v2 0.2 %s 3 10007 5 1 40000 16" % arch,    # outside the window
        "gofft-tune-v2 0.2 %s 3 10007 6 1 20160 5" % arch,     # columns not a candidate
        "gofft-tune-v2 0.1 %s 3 10007 7 1 20160 16" % arch,    # another library version
        "garbage line",
    ]) + "\n")
    env = dict(os.environ, GOFFT_AUTOTUNE="1", GOFFT_TUNE_CACHE=str(cache))
    code = ("import sys; sys.path[:0] = [%r, %r]; import torch, numpy as np; from fft_amd import device as D; "
            "from conftest import cgauss, rel_l2; from oracle import oracle as O\n"
            "for b in (4, 5, 6, 7):\n"
            "    x = cgauss(10007 * b, seed=b).reshape(b, 10007)\n"
            "    p = D.C2CPlan(10007, b, torch.complex64)\n"
            "    y = p.exec(torch.from_numpy(x.astype(np.complex64).ravel()).to('cuda')).cpu().numpy()\n"
            "    e = rel_l2(y.reshape(b, 10007), O.fft_batch(x)); print(b, e, p.describe(), flush=True)\n"
            "    assert e < 1e-6\n" % (ROOT, os.path.join(ROOT, "tests")))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, (out.stdout[-1000:], out.stderr[-2000:])
