"""The C-ABI's host code under AddressSanitizer (no GPU needed).

Every source of libinc_amd is compiled for the host only (no device code)
with AddressSanitizer on the host pass, linked into a shared library, and
tests/asan_driver.cpp drives the argument-validation paths, the error-text
formatting, the workspace-size arithmetic and the per-thread error state
through it.  Nothing launches a kernel.  ASan aborts the driver on an
out-of-bounds access, a use after free or a double free; the driver itself
checks the status codes and messages.

A host-only object still registers its device-code bundle with the HIP
runtime at load, by symbol (__hip_fatbin_<hash>); a host-only build has no
bundle, so each such symbol is given an empty placeholder.  The runtime
reads a bundle only when one of its kernels is launched, which never
happens here.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "inc_amd", "csrc")
SOURCES = ["common.cpp", "burgers.hip", "ks.hip", "conv.hip", "unet.hip", "cfl.hip"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def _hipcc():
    return shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")


@pytest.mark.skipif(not os.path.exists(_hipcc()) or shutil.which("g++") is None, reason="needs hipcc and g++")
def test_c_abi_host_code_under_address_sanitizer(tmp_path):
    def compile_one(src):
        obj = str(tmp_path / (os.path.splitext(src)[0] + ".o"))
        cmd = [_hipcc(), "-x", "hip", "--cuda-host-only", "--offload-arch=gfx950", "-O1", "-g", "-std=c++17",
               "-fPIC", "-fno-omit-frame-pointer", "-Xarch_host", "-fsanitize=address", "-Wno-unused-result",
               "-c", os.path.join(CSRC, src), "-o", obj]
        res = subprocess.run(cmd, capture_output=True, text=True)
        assert res.returncode == 0, res.stderr[-4000:]
        return obj

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    nm = subprocess.run(["nm", "-u"] + objs, capture_output=True, text=True)
    assert nm.returncode == 0, nm.stderr
    syms = sorted({w for w in nm.stdout.split() if w.startswith("__hip_fatbin_")})
    stub = tmp_path / "no_device_code.c"
    stub.write_text("".join(f"unsigned char {sym}[64];\n" for sym in syms))
    objs.append(str(tmp_path / "no_device_code.o"))
    res = subprocess.run(["gcc", "-c", "-fPIC", str(stub), "-o", objs[-1]], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    lib = str(tmp_path / "libinc_amd_asan.so")
    res = subprocess.run(["g++", "-shared", "-fsanitize=address", "-o", lib] + objs
                         + [f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib"], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    exe = str(tmp_path / "asan_driver")
    res = subprocess.run(["g++", "-std=c++17", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-pthread",
                          os.path.join(ROOT, "tests", "asan_driver.cpp"), "-o", exe, lib,
                          f"-Wl,-rpath,{tmp_path}", f"-Wl,-rpath,{ROCM}/lib"], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1")
    res = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert res.returncode == 0, (res.stdout + res.stderr)[-6000:]
    assert "all checks passed" in res.stdout
