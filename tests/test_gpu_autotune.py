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
    assert any(l.startswith("3 10007 64 1 ") for l in lines1), lines1
    out2 = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out2.returncode == 0, out2.stderr[-2000:]
    assert cache.read_text().split("\n") == lines1  # reused, not timed again
    assert "tuned" in out2.stdout
