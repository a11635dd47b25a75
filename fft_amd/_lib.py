"""ctypes binding of libgofft_hip.so (the C ABI in include/gofft.h).

The library must be built in place (fft_amd/build.py, or __graft_entry__.build()).
There is no fallback: if the shared library is missing, every call raises.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgofft_hip.so")

OK = 0
ERR_INVALID_ARG = -1
ERR_SHAPE = -2
ERR_OOM = -3
ERR_HIP = -4
ERR_UNSUPPORTED = -5

C128 = 0
C64 = 1
FORWARD = 0
INVERSE = 1
FLAG_NONE = 0
FLAG_REFERENCE_TWIDDLES = 1
FLAG_ACCURATE_TWIDDLES = 2
FLAG_AUTOTUNE = 4
FLAG_TWIDDLES_FP32 = 8
FLAG_TWIDDLES_SPLIT = 16
FLAG_QUEUE = 32
ND_C2C = 0
ND_R2C = 1
KERNEL_PASS, KERNEL_CONV, KERNEL_REAL, KERNEL_AUX = 0, 1, 2, 3
(TABLE_RADER_GENERATOR, TABLE_RADER_PERM, TABLE_RADER_IPERM, TABLE_RADER_CONV, TABLE_CHIRP_INDEX,
 TABLE_BLUESTEIN_CONV) = range(6)
(OP_C2C, OP_R2C, OP_C2R, OP_ND_R2C, OP_PSD, OP_SPECTROGRAM, OP_VJP_C2C, OP_VJP_R2C, OP_VJP_C2R,
 OP_VJP_ND_R2C) = range(10)

_i64 = ctypes.c_int64
_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); mirrors include/gofft.h
SIGNATURES = {
    "gofft_init": (ctypes.c_int, []),
    "gofft_last_error": (ctypes.c_char_p, []),
    "gofft_version": (ctypes.c_char_p, []),
    "gofft_route": (ctypes.c_int, [_i64]),
    "gofft_plan_c2c": (ctypes.c_int, [_pp, _i64, _i64, ctypes.c_int, ctypes.c_uint]),
    "gofft_plan_c2c_many": (ctypes.c_int, [_pp, ctypes.POINTER(_i64), _i64, ctypes.c_int, ctypes.c_uint]),
    "gofft_plan_r2c": (ctypes.c_int, [_pp, _i64, _i64, ctypes.c_int, ctypes.c_uint]),
    "gofft_plan_nd": (ctypes.c_int, [_pp, ctypes.POINTER(_i64), ctypes.c_int, ctypes.c_int, _i64,
                                     ctypes.c_int, ctypes.c_uint]),
    "gofft_plan_psd": (ctypes.c_int, [_pp, _i64, _i64, ctypes.c_double, ctypes.c_int, ctypes.c_uint]),
    "gofft_plan_spectrogram": (ctypes.c_int, [_pp, _i64, _i64, _i64, ctypes.POINTER(ctypes.c_double),
                                              _i64, ctypes.c_double, ctypes.c_int, ctypes.c_uint]),
    "gofft_plan_destroy": (ctypes.c_int, [_vp]),
    "gofft_plan_len": (_i64, [_vp]),
    "gofft_plan_frames": (_i64, [_vp]),
    "gofft_plan_in_elems": (_i64, [_vp, ctypes.c_int]),
    "gofft_plan_out_elems": (_i64, [_vp, ctypes.c_int]),
    "gofft_plan_workspace_bytes": (_i64, [_vp]),
    "gofft_plan_num_launches": (ctypes.c_int, [_vp, ctypes.c_int]),
    "gofft_plan_describe": (_i64, [_vp, ctypes.c_char_p, _i64]),
    "gofft_plan_index_table": (_i64, [_vp, _i64, ctypes.c_int, ctypes.POINTER(_i64), _i64]),
    "gofft_plan_kernel_work": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_i64),
                                              ctypes.POINTER(_i64)]),
    "gofft_exec_c2c": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "gofft_exec_r2c": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "gofft_exec_c2r": (ctypes.c_int, [_vp, _vp, _i64, _vp, _vp]),
    "gofft_exec_nd_r2c": (ctypes.c_int, [_vp, _vp, _i64, _vp, ctypes.c_int, _vp]),
    "gofft_exec_psd": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "gofft_exec_spectrogram": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "gofft_exec_c2c_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "gofft_vjp_c2c": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "gofft_vjp_r2c": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "gofft_vjp_c2r": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "gofft_vjp_nd_r2c": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "gofft_filter2_loss_grad": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "gofft_exec_c2c_host": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "gofft_exec_r2c_host": (ctypes.c_int, [_vp, _vp, _vp]),
    "gofft_exec_c2r_host": (ctypes.c_int, [_vp, _vp, _i64, _vp]),
    "gofft_vjp_c2c_host": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "gofft_vjp_r2c_host": (ctypes.c_int, [_vp, _vp, _vp]),
    "gofft_vjp_c2r_host": (ctypes.c_int, [_vp, _vp, _vp]),
    "gofft_exec_nd_r2c_host": (ctypes.c_int, [_vp, _vp, _i64, _vp, ctypes.c_int]),
    "gofft_exec_psd_host": (ctypes.c_int, [_vp, _vp, _vp]),
    "gofft_exec_spectrogram_host": (ctypes.c_int, [_vp, _vp, _vp]),
    "gofft_vjp_nd_r2c_host": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "gofft_graph_capture": (ctypes.c_int, [_pp, _vp, ctypes.c_int, _vp, _i64, _vp, ctypes.c_int, _vp]),
    "gofft_graph_c2c": (ctypes.c_int, [_pp, _vp, _vp, _vp, ctypes.c_int, _vp]),
    "gofft_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "gofft_graph_destroy": (ctypes.c_int, [_vp]),
    "gofft_sync": (ctypes.c_int, []),
    "gofft_probe_dft16": (ctypes.c_int, [ctypes.c_int, _vp, _vp, _i64, ctypes.c_int, _vp]),
}

_lib = None
_lock = threading.Lock()


class GofftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gofft error %d: %s" % (code, msg))
        self.code = code
        self.msg = msg


def lib():
    """Load libgofft_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    "libgofft_hip.so is not built (%s); run fft_amd/build.py" % LIB_PATH)
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def last_error():
    return lib().gofft_last_error().decode(errors="replace")


def check(rc):
    if rc != OK:
        raise GofftError(rc, last_error())
    return rc
