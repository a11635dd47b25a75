"""Device-resident plans over torch tensors (HBM in, HBM out).

Thin wrappers of the C ABI: torch provides device memory and the current HIP
stream; every transform runs in libgofft_hip.so's kernels. complex128/float64
tensors select the plain FP64 path, complex64/float32 the FP32 path.
"""
import ctypes
import math

import torch

from . import _lib as L

_C_TYPES = {torch.complex128: L.C128, torch.complex64: L.C64}
_R_TYPES = {torch.float64: L.C128, torch.float32: L.C64}
COMPLEX_OF = {L.C128: torch.complex128, L.C64: torch.complex64}
REAL_OF = {L.C128: torch.float64, L.C64: torch.float32}


def dtype_code(dtype):
    if dtype in (torch.complex128, torch.float64, "c128", "complex128", "f64", L.C128):
        return L.C128
    if dtype in (torch.complex64, torch.float32, "c64", "complex64", "f32", L.C64):
        return L.C64
    raise ValueError("unsupported dtype %r" % (dtype,))


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class _Plan:
    def __init__(self):
        self._h = ctypes.c_void_p()

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and L is not None and L._lib is not None:
            L._lib.gofft_plan_destroy(h)
            self._h = ctypes.c_void_p()

    def workspace_bytes(self):
        return L.lib().gofft_plan_workspace_bytes(self._h)

    def num_launches(self, direction=0):
        return L.lib().gofft_plan_num_launches(self._h, direction)

    def describe(self):
        """The plan's engines and launches as text."""
        n = L.lib().gofft_plan_describe(self._h, None, 0)
        buf = ctypes.create_string_buffer(max(1, n))
        L.lib().gofft_plan_describe(self._h, buf, n)
        return buf.value.decode()

    def kernel_work(self, kernel, direction=0):
        """(launches, data bytes) of kernel family `kernel` (L.KERNEL_*) in one exec."""
        n, b = ctypes.c_int64(), ctypes.c_int64()
        L.check(L.lib().gofft_plan_kernel_work(self._h, direction, kernel, ctypes.byref(n), ctypes.byref(b)))
        return n.value, b.value

    def index_table(self, n, table):
        """Integer index table `table` (L.TABLE_*) of the plan's length-n transform, as int64 numpy."""
        import numpy as np
        k = L.lib().gofft_plan_index_table(self._h, n, table, None, 0)
        if k < 0:
            raise L.GofftError(int(k), L.last_error())
        out = np.zeros(max(1, k), dtype=np.int64)
        L.lib().gofft_plan_index_table(self._h, n, table, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), k)
        return out[:k]

    def in_elems(self, direction=0):
        return L.lib().gofft_plan_in_elems(self._h, direction)

    def out_elems(self, direction=0):
        return L.lib().gofft_plan_out_elems(self._h, direction)


def run_op(plan, op, x, out, bins=0, inverse=False, stream=None):
    """One call of operation `op` (L.OP_*: the exec and VJP entry points) on
    device tensors, with `bins` and the direction where the op takes them."""
    lib, h, s = L.lib(), plan.handle, _stream(stream)
    d = int(bool(inverse))
    xi, oi = _ptr(x), _ptr(out)
    calls = {
        L.OP_C2C: lambda: lib.gofft_exec_c2c(h, xi, oi, d, s),
        L.OP_R2C: lambda: lib.gofft_exec_r2c(h, xi, oi, s),
        L.OP_C2R: lambda: lib.gofft_exec_c2r(h, xi, int(bins), oi, s),
        L.OP_ND_R2C: lambda: lib.gofft_exec_nd_r2c(h, xi, int(bins), oi, d, s),
        L.OP_PSD: lambda: lib.gofft_exec_psd(h, xi, oi, s),
        L.OP_SPECTROGRAM: lambda: lib.gofft_exec_spectrogram(h, xi, oi, s),
        L.OP_VJP_C2C: lambda: lib.gofft_vjp_c2c(h, xi, oi, d, s),
        L.OP_VJP_R2C: lambda: lib.gofft_vjp_r2c(h, xi, oi, s),
        L.OP_VJP_C2R: lambda: lib.gofft_vjp_c2r(h, xi, oi, s),
        L.OP_VJP_ND_R2C: lambda: lib.gofft_vjp_nd_r2c(h, xi, oi, d, s),
    }
    L.check(calls[op]())
    return out


class Graph:
    """One captured call of a plan on fixed tensors (gofft_graph_capture):
    operation `op` (L.OP_*, default the complex transform), with `bins` and
    the direction (`inverse`) where the op takes them. launch() replays it on
    a stream. The graph owns its workspace; it keeps the plan and tensors
    alive."""

    def __init__(self, plan, x, out, inverse=False, stream=None, op=L.OP_C2C, bins=0):
        self._g = ctypes.c_void_p()
        self._keep = (plan, x, out)
        L.check(L.lib().gofft_graph_capture(ctypes.byref(self._g), plan.handle, int(op), _ptr(x), int(bins),
                                            _ptr(out), int(bool(inverse)), _stream(stream)))

    def launch(self, stream=None):
        L.check(L.lib().gofft_graph_launch(self._g, _stream(stream)))

    def __del__(self):
        g = getattr(self, "_g", None)
        if g and L is not None and L._lib is not None:
            L._lib.gofft_graph_destroy(g)
            self._g = ctypes.c_void_p()


class C2CPlan(_Plan):
    """`batch` contiguous complex lines of length n (NewPlan, plan.go:42-73)."""

    def __init__(self, n, batch=1, dtype=torch.complex128, flags=L.FLAG_NONE):
        super().__init__()
        self.n, self.batch, self.code = int(n), int(batch), dtype_code(dtype)
        L.check(L.lib().gofft_plan_c2c(ctypes.byref(self._h), self.n, self.batch, self.code, flags))

    def exec(self, x, out=None, inverse=False, stream=None):
        if out is None:
            out = torch.empty_like(x)
        L.check(L.lib().gofft_exec_c2c(self._h, _ptr(x), _ptr(out), int(bool(inverse)), _stream(stream)))
        return out

    def vjp(self, gy, out=None, inverse=False, stream=None):
        if out is None:
            out = torch.empty_like(gy)
        L.check(L.lib().gofft_vjp_c2c(self._h, _ptr(gy), _ptr(out), int(bool(inverse)), _stream(stream)))
        return out


class ManyPlan(_Plan):
    """Transforms of uneven lengths packed back to back in one buffer."""

    def __init__(self, lengths, dtype=torch.complex128, flags=L.FLAG_NONE):
        super().__init__()
        self.lengths = [int(v) for v in lengths]
        self.code = dtype_code(dtype)
        arr = (ctypes.c_int64 * max(1, len(self.lengths)))(*self.lengths)
        L.check(L.lib().gofft_plan_c2c_many(ctypes.byref(self._h), arr, len(self.lengths), self.code, flags))
        self.total = sum(self.lengths)

    def exec(self, x, out=None, inverse=False, stream=None):
        if out is None:
            out = torch.empty_like(x)
        L.check(L.lib().gofft_exec_c2c(self._h, _ptr(x), _ptr(out), int(bool(inverse)), _stream(stream)))
        return out

    def vjp(self, gy, out=None, inverse=False, stream=None):
        if out is None:
            out = torch.empty_like(gy)
        L.check(L.lib().gofft_vjp_c2c(self._h, _ptr(gy), _ptr(out), int(bool(inverse)), _stream(stream)))
        return out


class R2CPlan(_Plan):
    """`batch` real lines of length n (NewRealPlan, realplan.go:37-58)."""

    def __init__(self, n, batch=1, dtype=torch.float64, flags=L.FLAG_NONE):
        super().__init__()
        self.n, self.batch, self.code = int(n), int(batch), dtype_code(dtype)
        L.check(L.lib().gofft_plan_r2c(ctypes.byref(self._h), self.n, self.batch, self.code, flags))

    def forward(self, x, out=None, stream=None):
        if out is None:
            out = torch.empty(self.batch * (self.n // 2 + 1), dtype=COMPLEX_OF[self.code], device=x.device)
        L.check(L.lib().gofft_exec_r2c(self._h, _ptr(x), _ptr(out), _stream(stream)))
        return out

    def inverse(self, X, bins=None, out=None, stream=None):
        if bins is None:
            bins = self.n // 2 + 1
        if out is None:
            out = torch.empty(self.batch * self.n, dtype=REAL_OF[self.code], device=X.device)
        L.check(L.lib().gofft_exec_c2r(self._h, _ptr(X), int(bins), _ptr(out), _stream(stream)))
        return out

    def vjp_forward(self, gy, out=None, stream=None):
        if out is None:
            out = torch.empty(self.batch * self.n, dtype=REAL_OF[self.code], device=gy.device)
        L.check(L.lib().gofft_vjp_r2c(self._h, _ptr(gy), _ptr(out), _stream(stream)))
        return out

    def vjp_inverse(self, gx, out=None, stream=None):
        if out is None:
            out = torch.empty(self.batch * (self.n // 2 + 1), dtype=COMPLEX_OF[self.code], device=gx.device)
        L.check(L.lib().gofft_vjp_c2r(self._h, _ptr(gx), _ptr(out), _stream(stream)))
        return out


class NDPlan(_Plan):
    """FFTN/FFT2 (kind ND_C2C) or RFFT2/IRFFT2 (kind ND_R2C) over `batch` arrays."""

    def __init__(self, shape, kind=L.ND_C2C, batch=1, dtype=torch.complex128, flags=L.FLAG_NONE):
        super().__init__()
        self.shape = [int(s) for s in shape]
        self.kind, self.batch, self.code = kind, int(batch), dtype_code(dtype)
        arr = (ctypes.c_int64 * max(1, len(self.shape)))(*self.shape)
        L.check(L.lib().gofft_plan_nd(ctypes.byref(self._h), arr, len(self.shape), kind, self.batch,
                                      self.code, flags))

    def exec(self, x, out=None, inverse=False, stream=None):
        if out is None:
            out = torch.empty_like(x)
        L.check(L.lib().gofft_exec_c2c(self._h, _ptr(x), _ptr(out), int(bool(inverse)), _stream(stream)))
        return out

    def rfft2(self, x, out=None, stream=None):
        rows, cols = self.shape
        if out is None:
            out = torch.empty(self.batch * rows * (cols // 2 + 1), dtype=COMPLEX_OF[self.code], device=x.device)
        L.check(L.lib().gofft_exec_nd_r2c(self._h, _ptr(x), 0, _ptr(out), 0, _stream(stream)))
        return out

    def irfft2(self, X, bins=None, out=None, stream=None):
        rows, cols = self.shape
        if bins is None:
            bins = rows * (cols // 2 + 1)
        if out is None:
            out = torch.empty(self.batch * rows * cols, dtype=REAL_OF[self.code], device=X.device)
        L.check(L.lib().gofft_exec_nd_r2c(self._h, _ptr(X), int(bins), _ptr(out), 1, _stream(stream)))
        return out

    def vjp(self, g, out, direction=0, stream=None):
        """ND_R2C plans: VJP of RFFT2 (direction 0: g complex half spectrum,
        out real) or of IRFFT2 (direction 1: g real, out complex)."""
        L.check(L.lib().gofft_vjp_nd_r2c(self._h, _ptr(g), _ptr(out), direction, _stream(stream)))
        return out

    def filter_loss_grad(self, x, H, y, chunk=0, stream=None):
        """ND_R2C plans: (loss per array, gradient) of ||IRFFT2(RFFT2(x) H) - y||^2
        (gofft_filter2_loss_grad), `chunk` arrays at a time (0: as many as fit)."""
        loss = torch.empty(self.batch, dtype=REAL_OF[self.code], device=x.device)
        gx = torch.empty_like(x)
        L.check(L.lib().gofft_filter2_loss_grad(self._h, _ptr(x), _ptr(H), _ptr(y), _ptr(loss), _ptr(gx),
                                                int(chunk), _stream(stream)))
        return loss, gx

    def vjp_c2c(self, gy, out=None, inverse=False, stream=None):
        """ND_C2C plans: VJP of FFTN (inverse=False) or IFFTN (inverse=True)."""
        if out is None:
            out = torch.empty_like(gy)
        L.check(L.lib().gofft_vjp_c2c(self._h, _ptr(gy), _ptr(out), int(bool(inverse)), _stream(stream)))
        return out


class PSDPlan(_Plan):
    """PSD (freq.go:69-90) of `batch` real lines."""

    def __init__(self, n, batch=1, d=1.0, dtype=torch.float64, flags=L.FLAG_NONE):
        super().__init__()
        self.n, self.batch, self.code = int(n), int(batch), dtype_code(dtype)
        L.check(L.lib().gofft_plan_psd(ctypes.byref(self._h), self.n, self.batch, float(d), self.code, flags))

    def exec(self, x, out=None, stream=None):
        if out is None:
            out = torch.empty(self.batch * (self.n // 2 + 1), dtype=REAL_OF[self.code], device=x.device)
        L.check(L.lib().gofft_exec_psd(self._h, _ptr(x), _ptr(out), _stream(stream)))
        return out


class SpectrogramPlan(_Plan):
    """Spectrogram (freq.go:107-127) of one real signal."""

    def __init__(self, length, segment, overlap, window, d=1.0, dtype=torch.float64, flags=L.FLAG_NONE):
        super().__init__()
        self.code = dtype_code(dtype)
        w = [float(v) for v in window]
        arr = (ctypes.c_double * max(1, len(w)))(*w)
        L.check(L.lib().gofft_plan_spectrogram(ctypes.byref(self._h), int(length), int(segment), int(overlap),
                                               arr, len(w), float(d), self.code, flags))
        self.frames = L.lib().gofft_plan_frames(self._h)
        self.bins = int(segment) // 2 + 1

    def exec(self, x, out=None, stream=None):
        if out is None:
            out = torch.empty(max(1, self.frames * self.bins), dtype=REAL_OF[self.code], device=x.device)
        L.check(L.lib().gofft_exec_spectrogram(self._h, _ptr(x), _ptr(out), _stream(stream)))
        return out
