"""Build libofdft_hip.so in-tree with hipcc for gfx950.

    python dftpy_amd/build.py        # compiles dftpy_amd/lib/libofdft_hip.so

Every translation unit is compiled to an object (in parallel), then linked
into one shared library with the C ABI of include/ofdft_hip.h.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libofdft_hip.so")

SOURCES = ["fft.hip", "fft_fixed.hip", "fft_fixed_f32.hip", "fft_fixed_f64.hip", "vec.hip", "eval.hip", "pseudo.hip", "opt.hip", "td.hip", "cnn.hip", "lmgp.hip", "density.hip", "capi.cpp", "dcsrch.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("OFDFT_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result", "-I" + os.path.join(ROOT, "include")]
# HIP device code: gfx950, packed-FP32 instructions disabled
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fno-gpu-rdc", "-munsafe-fp-atomics"]


def _compile(src, obj):
    path = os.path.join(SRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC] + COMMON + DEVICE + ["-c", path, "-o", obj]
    else:
        # host-only translation units: plain g++ against the HIP runtime headers
        cmd = ["g++"] + COMMON + ["-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r.stderr


def build(verbose=False, jobs=None):
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    jobs = jobs or min(len(SOURCES), os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = {ex.submit(_compile, s, o): s for s, o in zip(SOURCES, objs)}
        for f in cf.as_completed(futs):
            err = f.result()
            if verbose and err.strip():
                print(futs[f], err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    shutil.rmtree(objdir, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
