"""The reference's public transform API, backed by the gfx950 library.

Same names, argument meaning and error behaviour as go-fft's package `fft`
(plan.go, realplan.go, fft.go, rfft.go, ndfft.go, rfft2.go, freq.go,
window.go): slices become 1-D numpy arrays, Go panics become `Panic`
(a ValueError) carrying the reference's message text. Every transform runs on
the GPU: host arrays are copied to HBM, transformed by libgofft_hip.so, and
copied back. complex128/float64 is the default (the reference's only type);
pass dtype=np.complex64 for the FP32 form.
"""
import math
import threading

import numpy as np
import torch

from . import _lib as L
from . import device as D

__all__ = [
    "Panic", "Plan", "RealPlan", "NewPlan", "NewRealPlan", "FFT", "IFFT", "FFTReal", "RFFT",
    "IRFFT", "FFTN", "IFFTN", "FFT2", "IFFT2", "RFFT2", "IRFFT2", "PSD", "Spectrogram", "Hann",
    "Hamming", "Blackman", "BlackmanHarris", "Bartlett", "FFTFreq", "RFFTFreq",
    "FFTVJP", "IFFTVJP", "RFFTVJP", "IRFFTVJP", "FFTNVJP", "IFFTNVJP", "RFFT2VJP", "IRFFT2VJP",
]


class Panic(ValueError):
    """A Go panic of the reference API, with the reference's message."""


_DEV = "cuda"


def _ctype(dtype):
    return np.complex64 if np.dtype(dtype) in (np.complex64, np.float32) else np.complex128


def _rtype(dtype):
    return np.float32 if np.dtype(dtype) in (np.complex64, np.float32) else np.float64


def _to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(_DEV)


def _to_host(t):
    return t.cpu().numpy()


# Plan caches keyed like the reference's (plan.go:287-317, rfft.go:47-64),
# extended by dtype and batch.
_cache_mu = threading.Lock()
_c2c_cache = {}
_r2c_cache = {}


def _c2c_plan(n, dtype, batch=1):
    key = (n, np.dtype(dtype).str, batch)
    with _cache_mu:
        p = _c2c_cache.get(key)
    if p is None:
        p = D.C2CPlan(n, batch, torch.complex64 if _ctype(dtype) == np.complex64 else torch.complex128)
        with _cache_mu:
            p = _c2c_cache.setdefault(key, p)
    return p


def _r2c_plan(n, dtype, batch=1):
    key = (n, np.dtype(dtype).str, batch)
    with _cache_mu:
        p = _r2c_cache.get(key)
    if p is None:
        p = D.R2CPlan(n, batch, torch.float32 if _rtype(dtype) == np.float32 else torch.float64)
        with _cache_mu:
            p = _r2c_cache.setdefault(key, p)
    return p


class Plan:
    """A reusable transform of fixed length n (plan.go:19-109)."""

    def __init__(self, n, dtype=np.complex128):
        self.n = int(n)
        self.dtype = _ctype(dtype)
        self._dev = _c2c_plan(self.n, self.dtype) if self.n > 1 else None

    def Len(self):
        return self.n

    def _run(self, dst, src, inverse):
        src = np.asarray(src)
        if self.n <= 1:
            dst[: self.n] = src[: self.n]
            return dst
        x = _to_dev(src[: self.n].astype(self.dtype, copy=False))
        y = self._dev.exec(x, inverse=inverse)
        dst[: self.n] = _to_host(y)
        return dst

    def FFT(self, dst, src):
        """Forward DFT of src into dst (dst may alias src); returns dst."""
        return self._run(dst, src, False)

    def IFFT(self, dst, src):
        """Inverse DFT normalised by n; returns dst."""
        return self._run(dst, src, True)


def NewPlan(n, dtype=np.complex128):
    return Plan(n, dtype)


class RealPlan:
    """A reusable real-input transform of fixed length n (realplan.go:17-192)."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = _rtype(dtype)
        self._dev = _r2c_plan(self.n, self.dtype) if self.n >= 1 else None

    def Len(self):
        return self.n

    def RFFT(self, dst, src):
        """n/2+1 bins of the real signal src into dst; returns dst[:n/2+1]."""
        n = self.n
        if n == 0:
            return dst[:0]
        x = _to_dev(np.asarray(src[:n], dtype=self.dtype))
        dst[: n // 2 + 1] = _to_host(self._dev.forward(x))
        return dst[: n // 2 + 1]

    def IRFFT(self, dst, src):
        """Real signal of length n from min(len(src), n/2+1) bins; returns dst."""
        n = self.n
        if n <= 0:
            return dst[:0]
        src = np.asarray(src, dtype=_ctype(self.dtype))
        bins = min(len(src), n // 2 + 1)
        X = _to_dev(src[:bins] if bins > 0 else np.zeros(1, dtype=src.dtype))
        dst[:n] = _to_host(self._dev.inverse(X, bins=bins))
        return dst


def NewRealPlan(n, dtype=np.float64):
    return RealPlan(n, dtype)


def FFT(x, dtype=None):
    """Forward DFT (fft.go:14-23); the input is not modified."""
    x = np.asarray(x)
    dt = _ctype(dtype or (x.dtype if x.dtype.kind == "c" else np.complex128))
    out = np.zeros(len(x), dtype=dt)
    if len(x) <= 1:
        out[:] = x
        return out
    return Plan(len(x), dt).FFT(out, x)


def IFFT(x, dtype=None):
    """Inverse DFT normalised by n (fft.go:31-40)."""
    x = np.asarray(x)
    dt = _ctype(dtype or (x.dtype if x.dtype.kind == "c" else np.complex128))
    out = np.zeros(len(x), dtype=dt)
    if len(x) <= 1:
        out[:] = x
        return out
    return Plan(len(x), dt).IFFT(out, x)


def FFTReal(x):
    """Full complex spectrum of a real signal (fft.go:45-51)."""
    return FFT(np.asarray(x, dtype=np.float64).astype(np.complex128))


def RFFT(x, dtype=None):
    """n/2+1 non-redundant bins of a real signal (rfft.go:19-26)."""
    x = np.asarray(x)
    dt = _rtype(dtype or (x.dtype if x.dtype.kind == "f" else np.float64))
    n = len(x)
    if n == 0:
        return np.zeros(0, dtype=_ctype(dt))
    dst = np.zeros(n // 2 + 1, dtype=_ctype(dt))
    return RealPlan(n, dt).RFFT(dst, x)


def IRFFT(spectrum, n, dtype=None):
    """Real signal of length n from a half spectrum (rfft.go:39-45)."""
    spectrum = np.asarray(spectrum)
    dt = _rtype(dtype or (np.float32 if spectrum.dtype == np.complex64 else np.float64))
    if n <= 0:
        return np.zeros(0, dtype=dt)
    dst = np.zeros(n, dtype=dt)
    return RealPlan(n, dt).IRFFT(dst, spectrum)


def _validate_shape(shape, total):
    """validateShape, ndfft.go:50-62."""
    prod = 1
    for s in shape:
        if s <= 0:
            raise Panic("fft: shape lengths must be positive")
        prod *= s
    if prod != total:
        raise Panic("fft: shape product does not match len(data)")
    return prod


def _transform_n(data, shape, inverse, dtype):
    data = np.asarray(data)
    shape = [int(s) for s in shape]
    _validate_shape(shape, len(data))
    dt = _ctype(dtype or (data.dtype if data.dtype.kind == "c" else np.complex128))
    if len(data) <= 1:
        return data.astype(dt).copy()
    p = D.NDPlan(shape, L.ND_C2C, 1, torch.complex64 if dt == np.complex64 else torch.complex128)
    return _to_host(p.exec(_to_dev(data.astype(dt, copy=False)), inverse=inverse))


def FFTN(data, shape, dtype=None):
    """N-dimensional forward DFT of row-major data (ndfft.go:22-24)."""
    return _transform_n(data, shape, False, dtype)


def IFFTN(data, shape, dtype=None):
    """Inverse of FFTN, normalised by the product of the axis lengths (ndfft.go:31-33)."""
    return _transform_n(data, shape, True, dtype)


def FFT2(data, shape, dtype=None):
    return FFTN(data, list(shape), dtype)


def IFFT2(data, shape, dtype=None):
    return IFFTN(data, list(shape), dtype)


def RFFT2(data, shape, dtype=None):
    """Real 2-D forward transform, rows x (cols/2+1) bins (rfft2.go:22-74)."""
    rows, cols = int(shape[0]), int(shape[1])
    if rows <= 0 or cols <= 0:
        raise Panic("fft: shape lengths must be positive")
    data = np.asarray(data)
    if rows * cols != len(data):
        raise Panic("fft: shape product does not match len(data)")
    dt = _rtype(dtype or (data.dtype if data.dtype.kind == "f" else np.float64))
    p = D.NDPlan([rows, cols], L.ND_R2C, 1, torch.float32 if dt == np.float32 else torch.float64)
    return _to_host(p.rfft2(_to_dev(data.astype(dt, copy=False))))


def IRFFT2(data, shape, dtype=None):
    """Inverse of RFFT2 (rfft2.go:87-138); reads up to rows*(cols/2+1) bins."""
    rows, cols = int(shape[0]), int(shape[1])
    if rows <= 0 or cols <= 0:
        raise Panic("fft: shape lengths must be positive")
    data = np.asarray(data)
    dt = _rtype(dtype or (np.float32 if data.dtype == np.complex64 else np.float64))
    ct = _ctype(dt)
    p = D.NDPlan([rows, cols], L.ND_R2C, 1, torch.float32 if dt == np.float32 else torch.float64)
    bins = min(len(data), rows * (cols // 2 + 1))
    X = _to_dev(data[:bins].astype(ct) if bins > 0 else np.zeros(1, dtype=ct))
    return _to_host(p.irfft2(X, bins=bins))


# ---------------------------------------------------------------------------
# Vector-Jacobian products (no reference counterpart). For a real loss L, the
# gradient with respect to a complex array z is dL/dRe(z) + i dL/dIm(z).


def _vjp_c2c(gy, inverse, dtype):
    gy = np.asarray(gy)
    dt = _ctype(dtype or (gy.dtype if gy.dtype.kind == "c" else np.complex128))
    n = len(gy)
    if n == 0:
        return np.zeros(0, dtype=dt)
    p = _c2c_plan(n, dt)
    return _to_host(p.vjp(_to_dev(gy.astype(dt, copy=False)), inverse=inverse))


def FFTVJP(gy, dtype=None):
    """Gradient through y = FFT(x): n * IFFT(gy)."""
    return _vjp_c2c(gy, False, dtype)


def IFFTVJP(gy, dtype=None):
    """Gradient through y = IFFT(x): FFT(gy) / n."""
    return _vjp_c2c(gy, True, dtype)


def RFFTVJP(gy, n, dtype=None):
    """Gradient (real, length n) through y = RFFT(x) from the n/2+1 bins gy."""
    gy = np.asarray(gy)
    dt = _rtype(dtype or (np.float32 if gy.dtype == np.complex64 else np.float64))
    if n <= 0:
        return np.zeros(0, dtype=dt)
    if len(gy) != n // 2 + 1:
        raise Panic("fft: gradient length must be n/2+1")
    p = _r2c_plan(n, dt)
    return _to_host(p.vjp_forward(_to_dev(gy.astype(_ctype(dt), copy=False))))


def IRFFTVJP(gx, dtype=None):
    """Gradient (n/2+1 bins) through x = IRFFT(X, n) from the real gx of length n."""
    gx = np.asarray(gx)
    dt = _rtype(dtype or (gx.dtype if gx.dtype.kind == "f" else np.float64))
    n = len(gx)
    if n == 0:
        return np.zeros(0, dtype=_ctype(dt))
    p = _r2c_plan(n, dt)
    return _to_host(p.vjp_inverse(_to_dev(gx.astype(dt, copy=False))))


def _vjp_n(gy, shape, inverse, dtype):
    gy = np.asarray(gy)
    shape = [int(s) for s in shape]
    _validate_shape(shape, len(gy))
    dt = _ctype(dtype or (gy.dtype if gy.dtype.kind == "c" else np.complex128))
    if len(gy) <= 1:
        return gy.astype(dt).copy()
    p = D.NDPlan(shape, L.ND_C2C, 1, torch.complex64 if dt == np.complex64 else torch.complex128)
    return _to_host(p.vjp_c2c(_to_dev(gy.astype(dt, copy=False)), inverse=inverse))


def FFTNVJP(gy, shape, dtype=None):
    """Gradient through y = FFTN(x, shape)."""
    return _vjp_n(gy, shape, False, dtype)


def IFFTNVJP(gy, shape, dtype=None):
    """Gradient through y = IFFTN(x, shape)."""
    return _vjp_n(gy, shape, True, dtype)


def _r2c2_shape(shape):
    rows, cols = int(shape[0]), int(shape[1])
    if rows <= 0 or cols <= 0:
        raise Panic("fft: shape lengths must be positive")
    return rows, cols


def RFFT2VJP(gy, shape, dtype=None):
    """Gradient (real rows*cols) through y = RFFT2(x, shape) from rows*(cols/2+1) bins."""
    rows, cols = _r2c2_shape(shape)
    gy = np.asarray(gy)
    if len(gy) != rows * (cols // 2 + 1):
        raise Panic("fft: gradient length must be rows*(cols/2+1)")
    dt = _rtype(dtype or (np.float32 if gy.dtype == np.complex64 else np.float64))
    p = D.NDPlan([rows, cols], L.ND_R2C, 1, torch.float32 if dt == np.float32 else torch.float64)
    out = torch.empty(rows * cols, dtype=D.REAL_OF[p.code], device=_DEV)
    return _to_host(p.vjp(_to_dev(gy.astype(_ctype(dt), copy=False)), out, direction=0))


def IRFFT2VJP(gx, shape, dtype=None):
    """Gradient (rows*(cols/2+1) bins) through x = IRFFT2(X, shape) from the real gx."""
    rows, cols = _r2c2_shape(shape)
    gx = np.asarray(gx)
    if len(gx) != rows * cols:
        raise Panic("fft: shape product does not match len(data)")
    dt = _rtype(dtype or (gx.dtype if gx.dtype.kind == "f" else np.float64))
    p = D.NDPlan([rows, cols], L.ND_R2C, 1, torch.float32 if dt == np.float32 else torch.float64)
    out = torch.empty(rows * (cols // 2 + 1), dtype=D.COMPLEX_OF[p.code], device=_DEV)
    return _to_host(p.vjp(_to_dev(gx.astype(dt, copy=False)), out, direction=1))


def PSD(x, d=1.0, dtype=None):
    """One-sided periodogram density |X|^2 d/n, doubled except DC/Nyquist (freq.go:69-90)."""
    x = np.asarray(x)
    dt = _rtype(dtype or (x.dtype if x.dtype.kind == "f" else np.float64))
    n = len(x)
    if n == 0:
        return np.zeros(0, dtype=dt)
    p = D.PSDPlan(n, 1, d, torch.float32 if dt == np.float32 else torch.float64)
    return _to_host(p.exec(_to_dev(x.astype(dt, copy=False))))


def Spectrogram(x, segment, overlap, window, d=1.0, dtype=None):
    """Framed, windowed PSDs (freq.go:107-127). Returns a list of frames, or
    None when no frame fits (the reference's nil)."""
    if segment <= 0:
        raise Panic("fft: segment length must be positive")
    if len(window) != segment:
        raise Panic("fft: window length must equal segment length")
    if overlap < 0 or overlap >= segment:
        raise Panic("fft: overlap must satisfy 0 <= overlap < segment")
    x = np.asarray(x)
    dt = _rtype(dtype or (x.dtype if x.dtype.kind == "f" else np.float64))
    p = D.SpectrogramPlan(len(x), segment, overlap, window, d, torch.float32 if dt == np.float32 else torch.float64)
    if p.frames == 0:
        return None
    out = _to_host(p.exec(_to_dev(x.astype(dt, copy=False)))).reshape(p.frames, p.bins)
    return [row for row in out]


# --- host helpers (window.go, freq.go): plain host arithmetic --------------


def _cosine_window(n, a):
    """cosineWindow, window.go:59-78."""
    if n <= 0:
        return np.zeros(0)
    if n == 1:
        return np.ones(1)
    denom = float(n - 1)
    out = np.empty(n)
    for k in range(n):
        s = a[0]
        sign = -1.0
        for j in range(1, len(a)):
            s += sign * a[j] * math.cos(2 * math.pi * float(j) * float(k) / denom)
            sign = -sign
        out[k] = s
    return out


def Hann(n):
    return _cosine_window(n, [0.5, 0.5])


def Hamming(n):
    return _cosine_window(n, [0.54, 0.46])


def Blackman(n):
    return _cosine_window(n, [0.42, 0.5, 0.08])


def BlackmanHarris(n):
    return _cosine_window(n, [0.35875, 0.48829, 0.14128, 0.01168])


def Bartlett(n):
    """window.go:86-100."""
    if n <= 0:
        return np.zeros(0)
    if n == 1:
        return np.ones(1)
    denom = float(n - 1)
    half = denom / 2
    return np.array([(2 / denom) * (half - abs(float(k) - half)) for k in range(n)])


def FFTFreq(n, d):
    """freq.go:17-35."""
    if n <= 0:
        return np.zeros(0)
    val = 1.0 / (d * float(n))
    out = np.empty(n)
    half = (n - 1) // 2
    for i in range(half + 1):
        out[i] = float(i) * val
    j = -(n // 2)
    for i in range(half + 1, n):
        out[i] = float(j) * val
        j += 1
    return out


def RFFTFreq(n, d):
    """freq.go:45-56."""
    if n <= 0:
        return np.zeros(0)
    m = n // 2 + 1
    val = 1.0 / (d * float(n))
    return np.array([float(i) * val for i in range(m)])
