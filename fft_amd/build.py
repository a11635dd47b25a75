"""Builds fft_amd/libgofft_hip.so (the gfx950 transform library) in place.

hipcc compiles the HIP kernels and the host plan layer for gfx950 only. All
HIP code is compiled with packed FP32 instructions disabled (they are not used
on this path). Objects are cached under fft_amd/_build and rebuilt when their
source or a header changes; `build(force=True)` recompiles everything.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgofft_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["kernels.hip", "queue.hip", "plan.cpp"]
HEADERS = [os.path.join(CSRC, "kernels.h"), os.path.join(INCLUDE, "gofft.h")]
# sources that include another source file
INCLUDES = {"queue.hip": [os.path.join(CSRC, "kernels.hip")]}

FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE,
    "-I" + CSRC,
]


def _stale(obj, src):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src] + HEADERS + INCLUDES.get(os.path.basename(src), [])
    return any(os.path.getmtime(d) > t for d in deps)


def build(force=False, verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    procs = []
    for name in SOURCES:
        src = os.path.join(CSRC, name)
        obj = os.path.join(OBJDIR, name + ".o")
        objs.append(obj)
        if force or _stale(obj, src):
            lang = ["-x", "hip"] if name.endswith(".cpp") else []
            cmd = [HIPCC] + FLAGS + lang + ["-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((name, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    for name, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        noise = "is not a recognized feature for this target"
        lines = [l for l in text.splitlines() if noise not in l]
        if p.returncode != 0:
            sys.stderr.write("\n".join(lines) + "\n")
            raise RuntimeError("hipcc failed on %s" % name)
        if verbose and lines:
            print("\n".join(lines))
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
    print(LIB)
