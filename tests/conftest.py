import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def cgauss(n, seed=0, dtype=np.complex128):
    """re/im iid N(0,1): numpy PCG64 default_rng(seed), real parts then imaginary parts."""
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(n)
    im = rng.standard_normal(n)
    return (re + 1j * im).astype(dtype)


def rgauss(n, seed=0, dtype=np.float64):
    return np.random.default_rng(seed).standard_normal(n).astype(dtype)


def rel_l2(got, want):
    got = np.asarray(got, dtype=np.complex128).ravel()
    want = np.asarray(want, dtype=np.complex128).ravel()
    den = np.linalg.norm(want)
    return float(np.linalg.norm(got - want) / (den if den > 0 else 1.0))


def h2h_complex(n):
    """The reference harness's closed-form complex input (benchmarks/verify/main.go:17-23)."""
    i = np.arange(n)
    return ((i * 7 + 1) % 13) * 0.1 + 1j * (((i * 3 + 2) % 11) * 0.1)


def h2h_real(n):
    i = np.arange(n)
    return ((i * 7 + 1) % 13) * 0.1


def cmplx_signal(n):
    """cmplxSignal, plan_test.go:21-27."""
    i = np.arange(n, dtype=np.float64)
    return (np.sin(i * 0.7) + (np.arange(n) % 3)) + 1j * (np.cos(i * 0.3) - (np.arange(n) % 5))


def real_signal(n):
    """realSignal, rfft_test.go:10-16."""
    i = np.arange(n, dtype=np.float64)
    return np.sin(i * 0.7) + 0.3 * np.cos(i * 1.9) + (np.arange(n) % 4)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    # the line fan-out of the N-D transforms (parallel.go) on at most 16
    # threads, the host share of one GPU box
    O.set_workers(min(16, os.cpu_count() or 1))
    return O
