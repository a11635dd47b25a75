"""CPU: the Go package go/gofft against the C ABI it binds. There is no Go
toolchain in this image or on the GPU box, so the binding is checked
statically: every C.gofft_* call names a function include/gofft.h declares
and passes as many arguments as the declaration takes, every C.GOFFT_*
constant is defined there, every exported name of the reference's transform
API has a Go counterpart, and the panic texts are the reference's."""
import os
import re

from conftest import ROOT

GO = os.path.join(ROOT, "go", "gofft")


def header_decls():
    text = open(os.path.join(ROOT, "include", "gofft.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"\b(gofft_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        args = m.group(2).strip()
        decls[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    consts = set(re.findall(r"#define\s+(GOFFT_[A-Z0-9_]+)", text))
    consts |= set(re.findall(r"\b(GOFFT_[A-Z0-9_]+)\s*=", text))
    return decls, consts


def go_sources(tests=False):
    out = {}
    for name in sorted(os.listdir(GO)):
        if name.endswith(".go") and name.endswith("_test.go") == tests:
            out[name] = open(os.path.join(GO, name)).read()
    return out


def call_args(text, start):
    """Top-level argument count of the call whose '(' is at text[start]."""
    depth, n, i, empty = 0, 1, start, True
    while True:
        c = text[i]
        if c in "([{":
            depth += 1
        elif c in ")]}":
            depth -= 1
            if depth == 0:
                return 0 if empty else n
        elif c == "," and depth == 1:
            n += 1
        elif not c.isspace() and depth >= 1 and i != start:
            empty = False
        i += 1


def test_every_c_call_matches_the_header():
    decls, consts = header_decls()
    seen = set()
    for name, text in go_sources().items():
        for m in re.finditer(r"\bC\.(gofft_[a-z0-9_]+)\s*\(", text):
            fn = m.group(1)
            assert fn in decls, (name, fn)
            got = call_args(text, m.end() - 1)
            assert got == decls[fn], (name, fn, got, decls[fn])
            seen.add(fn)
        for c in re.findall(r"\bC\.(GOFFT_[A-Z0-9_]+)", text):
            assert c in consts, (name, c)
    # the transforms, host-staged (the Go slices are host memory)
    for fn in ["gofft_plan_c2c", "gofft_plan_r2c", "gofft_plan_nd", "gofft_plan_psd", "gofft_plan_spectrogram",
               "gofft_plan_destroy", "gofft_last_error", "gofft_exec_c2c_host", "gofft_exec_r2c_host",
               "gofft_exec_c2r_host", "gofft_exec_nd_r2c_host", "gofft_exec_psd_host",
               "gofft_exec_spectrogram_host", "gofft_vjp_c2c_host", "gofft_vjp_r2c_host", "gofft_vjp_c2r_host",
               "gofft_vjp_nd_r2c_host"]:
        assert fn in seen, fn


def test_reference_api_is_covered():
    text = "\n".join(go_sources().values())
    exported = set(re.findall(r"^func (?:\([^)]*\) )?([A-Z][A-Za-z0-9]*)\(", text, flags=re.M))
    for name in ["NewPlan", "Len", "FFT", "IFFT", "FFTReal", "NewRealPlan", "RFFT", "IRFFT", "FFTN", "IFFTN",
                 "FFT2", "IFFT2", "RFFT2", "IRFFT2", "PSD", "Spectrogram", "FFTFreq", "RFFTFreq", "Hann",
                 "Hamming", "Blackman", "BlackmanHarris", "Bartlett", "FFTVJP", "IFFTVJP", "RFFTVJP",
                 "IRFFTVJP", "RFFT2VJP", "IRFFT2VJP"]:
        assert name in exported, name


def test_panic_texts_are_the_references():
    text = "\n".join(go_sources().values())
    for msg in ["fft: shape lengths must be positive", "fft: shape product does not match len(data)",
                "fft: segment length must be positive", "fft: window length must equal segment length",
                "fft: overlap must satisfy 0 <= overlap < segment"]:
        assert '"%s"' % msg in text, msg


def test_tests_port_the_reference_suites():
    tests = go_sources(tests=True)
    names = set()
    for text in tests.values():
        names |= set(re.findall(r"^func (Test[A-Za-z0-9]*)\(t \*testing\.T\)", text, flags=re.M))
        assert "package gofft" in text
    assert len(names) >= 25, sorted(names)


def test_go_files_are_balanced():
    # a cheap syntax sanity check: brackets balance outside strings and comments
    for name, text in {**go_sources(), **go_sources(tests=True)}.items():
        body = re.sub(r"//[^\n]*", "", text)
        body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
        body = re.sub(r'"(\\.|[^"\\])*"', '""', body)
        body = re.sub(r"`[^`]*`", "``", body)
        body = re.sub(r"'(\\.|[^'\\])'", "''", body)
        for o, c in ("()", "[]", "{}"):
            assert body.count(o) == body.count(c), (name, o)
