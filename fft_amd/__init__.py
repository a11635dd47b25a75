"""MI355X (gfx950) backend for the go-fft transform path.

`fft_amd.api` mirrors the reference's public API (NewPlan, Plan.FFT/IFFT,
NewRealPlan, RealPlan.RFFT/IRFFT, FFT/IFFT/FFTReal, RFFT/IRFFT, FFTN/IFFTN,
FFT2/IFFT2, RFFT2/IRFFT2, PSD, Spectrogram, windows and frequency axes);
`fft_amd.device` exposes device-resident plans over torch tensors. Both call
the C ABI of libgofft_hip.so (include/gofft.h).
"""
from . import _lib
from ._lib import GofftError


def library_path():
    return _lib.LIB_PATH


def __getattr__(name):
    # Import the torch-backed modules lazily so the C ABI can be inspected
    # without initialising a GPU runtime.
    if name in ("api", "device"):
        import importlib
        return importlib.import_module("." + name, __name__)
    raise AttributeError(name)
