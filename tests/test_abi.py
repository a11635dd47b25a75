"""The C ABI library loads and exports every entry point include/gofft.h
declares (no compute calls: CPU only)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gofft.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gofft_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for must in ["gofft_plan_c2c", "gofft_plan_c2c_many", "gofft_plan_r2c", "gofft_plan_nd",
                 "gofft_plan_psd", "gofft_plan_spectrogram", "gofft_exec_c2c", "gofft_exec_r2c",
                 "gofft_exec_c2r", "gofft_exec_nd_r2c", "gofft_exec_psd", "gofft_exec_spectrogram",
                 "gofft_vjp_c2c", "gofft_vjp_r2c", "gofft_vjp_c2r", "gofft_vjp_nd_r2c",
                 "gofft_plan_destroy", "gofft_last_error"]:
        assert must in syms


def test_library_exports_every_declared_symbol():
    from fft_amd import _lib
    L = _lib.lib()  # loads libgofft_hip.so; no GPU call
    for name in declared_symbols():
        assert hasattr(L, name), name
    # the ctypes signature table covers exactly the header
    assert set(_lib.SIGNATURES) == set(declared_symbols())


def test_pure_host_entry_points():
    from fft_amd import _lib
    L = _lib.lib()
    assert b"gfx950" in L.gofft_version()
    # routing mirrors NewPlan (plan.go:42-73) and needs no device
    assert [L.gofft_route(n) for n in [0, 1, 4096, 6720, 1009, 10007, 20014, 641, 769]] == \
        [0, 0, 1, 2, 3, 3, 4, 4, 3]
