"""GPU: the complex64 pass-twiddle forms (GOFFT_FLAG_TWIDDLES_FP32: FP32
tables and products; GOFFT_FLAG_TWIDDLES_SPLIT: hi + lo float pairs in
double-float arithmetic) beside the default FP64-formed twiddles, against
the oracle on transforms whose passes carry twiddles: multi-pass 1-D
transforms (plain and real, whose fused first and last passes take the
forms too) and a strided two-pass N-D axis. Each stays within the FP32 bar
of rel-L2 1e-6; the split form's error is never above the default's by more
than rounding noise."""
import numpy as np
import pytest

from conftest import cgauss, rel_l2, rgauss

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from fft_amd import device as D
    from fft_amd import _lib as L
    return D, L, torch


def forms(L):
    return {"default": 0, "fp32": L.FLAG_TWIDDLES_FP32, "split": L.FLAG_TWIDDLES_SPLIT}


@pytest.mark.parametrize("n,batch", [(1 << 16, 8), (1 << 20, 2), (3 * (1 << 18), 2), (1 << 22, 1)])
def test_c2c_twiddle_forms(dev, oracle, n, batch):
    D, L, torch = dev
    x = cgauss(n * batch, seed=n).astype(np.complex64)
    want = oracle.fft_batch(x.astype(np.complex128).reshape(batch, n), threads=16)
    errs = {}
    for name, flag in forms(L).items():
        p = D.C2CPlan(n, batch, torch.complex64, flag)
        got = p.exec(torch.from_numpy(x).to("cuda")).cpu().numpy().reshape(batch, n)
        errs[name] = rel_l2(got, want)
        back = p.exec(torch.from_numpy(got.ravel()).to("cuda"), inverse=True).cpu().numpy()
        assert rel_l2(back, x) <= 1e-6, (name, n)
    assert max(errs.values()) <= 1e-6, errs
    assert errs["split"] <= errs["default"] * 1.05 + 1e-9, errs
    print(n, errs)


@pytest.mark.parametrize("n", [1 << 20, 1 << 24])
def test_real_twiddle_forms(dev, oracle, n):
    D, L, torch = dev
    x = rgauss(n, seed=3).astype(np.float32)
    want = oracle.rfft(x.astype(np.float64))
    for name, flag in forms(L).items():
        p = D.R2CPlan(n, 1, torch.float32, flag)
        X = p.forward(torch.from_numpy(x).to("cuda"))
        assert rel_l2(X.cpu().numpy(), want) <= 1e-6, name
        back = p.inverse(X).cpu().numpy()
        assert rel_l2(back, x) <= 1e-6, name


def test_rfft2_column_passes(dev, oracle):
    # 8192 rows: the column axis is two strided passes, the second twiddled
    D, L, torch = dev
    rows, cols = 8192, 256
    x = np.random.default_rng(7).uniform(0, 1, rows * cols).astype(np.float32)
    want = oracle.rfft2(x.astype(np.float64), (rows, cols))
    for name, flag in forms(L).items():
        p = D.NDPlan((rows, cols), L.ND_R2C, 1, torch.complex64, flag)
        X = p.rfft2(torch.from_numpy(x).to("cuda"))
        assert rel_l2(X.cpu().numpy(), want) <= 1e-6, name
        assert rel_l2(p.irfft2(X).cpu().numpy(), x) <= 1e-6, name
