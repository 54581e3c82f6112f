"""Build libhfls.so (gfx950 HIP kernels + C++ host + C ABI) in-tree with hipcc.

Usage: python -m hfls_amd.build   (or hfls_amd.build.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libhfls.so")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")

# device code: gfx950 only; the compiler emits no packed-FP32 instructions (every
# FP32 operation is a plain VALU instruction)
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=fast",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-x", "c++", "-O2", "-std=c++17", "-fPIC", "-D__HIP_PLATFORM_AMD__",
              "-I" + os.path.join(ROCM, "include")]

SOURCES = [
    ("kernels.hip", HIP_FLAGS),
    ("kernels2.hip", HIP_FLAGS),
    ("symv.hip", HIP_FLAGS),
    ("tgemm.hip", HIP_FLAGS),
    ("tgemm16.hip", HIP_FLAGS),
    ("tgemm16b.hip", HIP_FLAGS),
    ("lrr.hip", HIP_FLAGS),
    ("block.hip", HIP_FLAGS),
    ("noalloc.hip", HIP_FLAGS),
    ("problem.cpp", HOST_FLAGS + ["-ffp-contract=off"]),
    ("engine.cpp", HOST_FLAGS),
    ("engine_block.cpp", HOST_FLAGS),
    ("serve.cpp", HOST_FLAGS),
    ("state.cpp", HOST_FLAGS),
    ("vtk.cpp", HOST_FLAGS),
    ("capi.cpp", HOST_FLAGS),
]


def _compile(src, flags):
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(lambda s: _compile(*s), SOURCES))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, OUT)
    # the reference's main() over the drop-in C++ facade (include/hfls/htfeti.hpp)
    root = os.path.dirname(HERE)
    exe = os.path.join(HERE, "htfeti")
    cmd = ["g++", "-std=c++17", "-O2", "-I" + os.path.join(root, "include"),
           os.path.join(CSRC, "htfeti_main.cpp"), "-L" + HERE, "-lhfls",
           "-Wl,-rpath,$ORIGIN", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade build failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
