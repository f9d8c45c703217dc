"""Build libinc_amd.so (HIP, gfx950) in-tree.

    python -m inc_amd.build          # compile csrc/*.hip + common.cpp

Objects are compiled in parallel with hipcc and linked into
``inc_amd/libinc_amd.so`` next to this file.  Packed-FP32 instructions are
disabled for every device pass (``-packed-fp32-ops``).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libinc_amd.so")
SOURCES = ["common.cpp", "burgers.hip", "ks.hip", "conv.hip", "unet.hip", "cfl.hip"]
ARCH = os.environ.get("INC_AMD_ARCH", "gfx950")


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found; ROCm is required to build libinc_amd")
    return exe


def _flags():
    return ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument"]


def _compile(src, outdir, verbose):
    obj = os.path.join(outdir, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".cpp") else []
    cmd = [hipcc()] + _flags() + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    if verbose and res.stderr.strip():
        msgs = [ln for ln in res.stderr.splitlines() if "packed-fp32-ops" not in ln]
        if msgs:
            print("\n".join(msgs), file=sys.stderr)
    return obj


def build(verbose=False):
    """Compile every HIP source and link the shared library; returns its path."""
    outdir = os.path.join(HERE, "_build")
    os.makedirs(outdir, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, outdir, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
