"""Build libxnyz.so (gfx950) and the XnYZip CLI in-tree with hipcc.

    python -m xnyzip_amd.build          # or: python xnyzip_amd/build.py

Outputs: xnyzip_amd/lib/libxnyz.so, xnyzip_amd/bin/XnYZip.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib", "libxnyz.so")
CLI = os.path.join(PKG, "bin", "XnYZip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ZSTD_INC = os.environ.get("XNYZ_ZSTD_INC", "/opt/conda/include")
ZSTD_LIB = os.environ.get("XNYZ_ZSTD_LIB", "/opt/conda/lib")
ARCH = os.environ.get("XNYZ_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no fused multiply-add (the reference's x86-64 baseline
# arithmetic).  Packed FP32 instructions are disabled.  Denormals are kept.
COMMON = ["-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-function", f"-I{os.path.join(ROOT, 'include')}", f"-I{ZSTD_INC}"]
DEVICE = [f"--offload-arch={ARCH}", "-fno-gpu-flush-denormals-to-zero",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["kernels.hip", "mlp.hip", "zframe.hip", "noalloc.hip", "decode_dev.hip",
           "hilbert_sm.cpp", "api.cpp", "decode.cpp", "container.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def _compile(src: str) -> str:
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, "-c", path, "-o", out]
    else:
        cmd = [HIPCC, *COMMON, "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
               "-c", path, "-o", out]
    _run(cmd)
    return out


def build(verbose: bool = False) -> str:
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp,
          f"-L{ZSTD_LIB}", "-lzstd", f"-Wl,-rpath,{ZSTD_LIB}"])
    os.replace(tmp, LIB)
    cli_src = os.path.join(PKG, "cli", "xnyzip_main.cpp")
    if os.path.exists(cli_src):
        os.makedirs(os.path.dirname(CLI), exist_ok=True)
        _run([HIPCC, "-O2", "-std=c++20", *COMMON[3:], cli_src, "-o", CLI + ".tmp",
              f"-L{os.path.dirname(LIB)}", "-lxnyz", "-Wl,-rpath,$ORIGIN/../lib"])
        os.replace(CLI + ".tmp", CLI)
    shutil.rmtree(OBJ, ignore_errors=True)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
