"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

The oracle is a C++ FP64 restatement of go-fft's CPU transform path (see the
header of gofft_oracle.cpp). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module — as the checker and as the CPU baseline,
never as the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

i64 = ctypes.c_longlong
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lp = ctypes.POINTER(ctypes.c_longlong)


def build():
    """Compile liboracle.so in place (g++, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sigs = {
            "or_set_workers": (None, [ctypes.c_int]),
            "or_get_workers": (ctypes.c_int, []),
            "or_route": (ctypes.c_int, [i64]),
            "or_factorize": (ctypes.c_int, [i64, _ip, ctypes.c_int]),
            "or_factors_are_small": (ctypes.c_int, [i64]),
            "or_next_smooth_conv": (i64, [i64]),
            "or_conv_cost": (ctypes.c_int, [i64, _dp]),
            "or_best_conv_len": (i64, [i64]),
            "or_primitive_root": (i64, [i64]),
            "or_is_prime": (ctypes.c_int, [i64]),
            "or_mod_pow": (i64, [i64, i64, i64]),
            "or_rader_info": (ctypes.c_int, [i64, _lp, _lp, _lp, _ip]),
            "or_bluestein_info": (i64, [i64, _lp]),
            "or_bitrev_for_radices": (None, [i64, _ip, ctypes.c_int, _lp]),
            "or_it_radices": (ctypes.c_int, [i64, _ip]),
            "or_bit_reverse": (None, [i64, _dp]),
            "or_radix2_rec": (None, [i64, _dp, ctypes.c_int]),
            "or_go_sincos": (None, [ctypes.c_double, _dp, _dp]),
            "or_radix2_stage_twiddles": (None, [i64, ctypes.c_int, _dp]),
            "or_twiddle_table": (None, [i64, _dp]),
            "or_fft": (ctypes.c_int, [i64, _dp, _dp, ctypes.c_int]),
            "or_fft_batch": (ctypes.c_int, [i64, i64, _dp, _dp, ctypes.c_int, ctypes.c_int]),
            "or_rfft": (ctypes.c_int, [i64, _dp, _dp]),
            "or_irfft": (ctypes.c_int, [i64, _dp, i64, _dp]),
            "or_fftn": (ctypes.c_int, [_lp, ctypes.c_int, i64, _dp, _dp, ctypes.c_int]),
            "or_rfft2": (ctypes.c_int, [i64, i64, i64, _dp, _dp]),
            "or_irfft2": (ctypes.c_int, [i64, i64, _dp, i64, _dp]),
            "or_psd": (ctypes.c_int, [i64, _dp, ctypes.c_double, _dp]),
            "or_spectrogram": (i64, [i64, _dp, i64, i64, _dp, i64, ctypes.c_double, _dp, ctypes.c_int]),
            "or_window": (ctypes.c_int, [ctypes.c_int, i64, _dp]),
            "or_fftfreq": (None, [i64, ctypes.c_double, _dp]),
            "or_rfftfreq": (None, [i64, ctypes.c_double, _dp]),
            "or_naive_dft": (ctypes.c_int, [i64, _dp, _dp, ctypes.c_int]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _c128(x):
    return np.ascontiguousarray(x, dtype=np.complex128)


def _f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def set_workers(w):
    lib().or_set_workers(int(w))


def route(n):
    return ["trivial", "iterative", "mixedradix", "rader", "bluestein"][lib().or_route(n)]


def factorize(n):
    buf = (ctypes.c_int * 64)()
    k = lib().or_factorize(n, buf, 64)
    return list(buf[:k])


def factors_are_small(n):
    return bool(lib().or_factors_are_small(n))


def next_smooth_conv(lo):
    return lib().or_next_smooth_conv(lo)


def conv_cost(m):
    c = ctypes.c_double()
    ok = lib().or_conv_cost(m, ctypes.byref(c))
    return c.value, bool(ok)


def best_conv_len(lo):
    return lib().or_best_conv_len(lo)


def primitive_root(p):
    return lib().or_primitive_root(p)


def is_prime(n):
    return bool(lib().or_is_prime(n))


def rader_info(n):
    perm = np.zeros(n - 1, dtype=np.int64)
    iperm = np.zeros(n - 1, dtype=np.int64)
    cl = ctypes.c_longlong()
    cyc = ctypes.c_int()
    rc = lib().or_rader_info(n, perm.ctypes.data_as(_lp), iperm.ctypes.data_as(_lp),
                             ctypes.byref(cl), ctypes.byref(cyc))
    if rc != 0:
        raise ValueError("not an odd prime: %d" % n)
    return perm, iperm, cl.value, bool(cyc.value)


def bluestein_info(n):
    ph = np.zeros(n, dtype=np.int64)
    m = lib().or_bluestein_info(n, ph.ctypes.data_as(_lp))
    return m, ph


def bitrev_for_radices(n, radices):
    r = (ctypes.c_int * len(radices))(*radices)
    out = np.zeros(n, dtype=np.int64)
    lib().or_bitrev_for_radices(n, r, len(radices), out.ctypes.data_as(_lp))
    return out


def it_radices(n):
    buf = (ctypes.c_int * 64)()
    k = lib().or_it_radices(n, buf)
    return list(buf[:k])


def bit_reverse(x):
    x = _c128(x).copy()
    lib().or_bit_reverse(len(x), _d(x.view(np.float64)))
    return x


def radix2_rec(x, inverse=False):
    x = _c128(x).copy()
    lib().or_radix2_rec(len(x), _d(x.view(np.float64)), int(inverse))
    return x


def go_sincos(x):
    s, c = ctypes.c_double(), ctypes.c_double()
    lib().or_go_sincos(float(x), ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def radix2_stage_twiddles(length, inverse=False):
    out = np.zeros(length // 2, dtype=np.complex128)
    lib().or_radix2_stage_twiddles(length, int(inverse), _d(out.view(np.float64)))
    return out


def twiddle_table(n):
    out = np.zeros(n, dtype=np.complex128)
    lib().or_twiddle_table(n, _d(out.view(np.float64)))
    return out


def fft(x, inverse=False):
    """Plan.FFT (or Plan.IFFT, normalised by 1/n) of one line."""
    x = _c128(x)
    out = np.zeros_like(x)
    if len(x):
        lib().or_fft(len(x), _d(x.view(np.float64)), _d(out.view(np.float64)), int(inverse))
    return out


def ifft(x):
    return fft(x, inverse=True)


def fft_batch(x, inverse=False, threads=1):
    """Row-wise Plan.FFT/IFFT over a (batch, n) array."""
    x = _c128(x)
    b, n = x.shape
    out = np.zeros_like(x)
    lib().or_fft_batch(n, b, _d(x.view(np.float64)), _d(out.view(np.float64)), int(inverse), int(threads))
    return out


def rfft(x):
    x = _f64(x)
    n = len(x)
    out = np.zeros(n // 2 + 1 if n else 0, dtype=np.complex128)
    if n:
        lib().or_rfft(n, _d(x), _d(out.view(np.float64)))
    return out


def irfft(spec, n):
    spec = _c128(spec)
    if n <= 0:
        return np.zeros(0)
    out = np.zeros(n, dtype=np.float64)
    sp = spec if len(spec) else np.zeros(1, dtype=np.complex128)
    lib().or_irfft(n, _d(sp.view(np.float64)), len(spec), _d(out))
    return out


def fftn(x, shape, inverse=False):
    x = _c128(x).reshape(-1)
    sh = np.asarray(shape, dtype=np.int64)
    out = np.zeros_like(x)
    src = x if len(x) else np.zeros(1, dtype=np.complex128)
    dst = out if len(out) else np.zeros(1, dtype=np.complex128)
    rc = lib().or_fftn(sh.ctypes.data_as(_lp), len(sh), len(x), _d(src.view(np.float64)),
                       _d(dst.view(np.float64)), int(inverse))
    if rc == -1:
        raise ValueError("fft: shape lengths must be positive")
    if rc == -2:
        raise ValueError("fft: shape product does not match len(data)")
    return out


def rfft2(x, shape):
    x = _f64(x).reshape(-1)
    rows, cols = shape
    if rows <= 0 or cols <= 0:
        raise ValueError("fft: shape lengths must be positive")
    if rows * cols != len(x):
        raise ValueError("fft: shape product does not match len(data)")
    out = np.zeros(rows * (cols // 2 + 1), dtype=np.complex128)
    lib().or_rfft2(rows, cols, len(x), _d(x), _d(out.view(np.float64)))
    return out


def irfft2(spec, shape):
    spec = _c128(spec).reshape(-1)
    rows, cols = shape
    if rows <= 0 or cols <= 0:
        raise ValueError("fft: shape lengths must be positive")
    out = np.zeros(rows * cols, dtype=np.float64)
    sp = spec if len(spec) else np.zeros(1, dtype=np.complex128)
    lib().or_irfft2(rows, cols, _d(sp.view(np.float64)), len(spec), _d(out))
    return out


def psd(x, d=1.0):
    x = _f64(x)
    n = len(x)
    out = np.zeros(n // 2 + 1 if n else 0)
    if n:
        lib().or_psd(n, _d(x), float(d), _d(out))
    return out


_SPEC_ERR = {
    -1: "fft: segment length must be positive",
    -2: "fft: window length must equal segment length",
    -3: "fft: overlap must satisfy 0 <= overlap < segment",
}


def spectrogram(x, segment, overlap, window, d=1.0, threads=1):
    """Returns a (frames, segment//2+1) array (frames may be 0)."""
    x = _f64(x)
    w = _f64(window)
    xs = x if len(x) else np.zeros(1)
    ws = w if len(w) else np.zeros(1)
    L = lib()
    frames = L.or_spectrogram(len(x), _d(xs), segment, overlap, _d(ws), len(w), float(d), None, 1)
    if frames < 0:
        raise ValueError(_SPEC_ERR[frames])
    out = np.zeros((frames, segment // 2 + 1))
    if frames:
        L.or_spectrogram(len(x), _d(xs), segment, overlap, _d(ws), len(w), float(d), _d(out), int(threads))
    return out


WINDOWS = {"hann": 0, "hamming": 1, "blackman": 2, "blackmanharris": 3, "bartlett": 4}


def window(kind, n):
    out = np.zeros(max(n, 0))
    if n > 0:
        lib().or_window(WINDOWS[kind], n, _d(out))
    return out


def fftfreq(n, d=1.0):
    out = np.zeros(max(n, 0))
    if n > 0:
        lib().or_fftfreq(n, float(d), _d(out))
    return out


def rfftfreq(n, d=1.0):
    out = np.zeros(n // 2 + 1 if n > 0 else 0)
    if n > 0:
        lib().or_rfftfreq(n, float(d), _d(out))
    return out


def naive_dft(x, inverse=False):
    """O(n^2) long-double DFT (independent check, unnormalised)."""
    x = _c128(x)
    out = np.zeros_like(x)
    lib().or_naive_dft(len(x), _d(x.view(np.float64)), _d(out.view(np.float64)), int(inverse))
    return out
