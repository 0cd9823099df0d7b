"""Build recipe for the native parts of torchproteinlibrary_amd (in-tree).

* ``libtpl_a2b.so``  -- HIP kernels + the C ABI of include/tpl_a2b.h, built by
  ``hipcc --offload-arch=gfx950`` (cross-compiles without a GPU).
* ``_ReducedModel<EXT_SUFFIX>`` -- the pybind11 torch extension that unpacks
  tensors and calls the C ABI; links libtpl_a2b.so with an $ORIGIN rpath.

Run ``python -m torchproteinlibrary_amd._build``.  Every call recompiles.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
import tempfile
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

LIB_NAME = "libtpl_a2b.so"
EXT_NAME = "_ReducedModel"


def lib_path() -> str:
    return os.path.join(PKG, LIB_NAME)


def ext_path() -> str:
    return os.path.join(PKG, EXT_NAME + sysconfig.get_config_var("EXT_SUFFIX"))


def _run(cmd, verbose):
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}):\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _atomic_out(final: str) -> str:
    fd, tmp = tempfile.mkstemp(prefix=os.path.basename(final) + ".", suffix=".tmp",
                               dir=os.path.dirname(final))
    os.close(fd)
    return tmp


def build_lib(verbose=True) -> str:
    out = lib_path()
    tmp = _atomic_out(out)
    cmd = [
        os.path.join(ROCM, "bin", "hipcc"), f"--offload-arch={ARCH}", "-O3", "-std=c++17",
        "-fPIC", "-shared", "-Wall", "-Werror", "-Wno-unused-command-line-argument",
        # No packed FP32 instructions anywhere in this library.
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", INCLUDE, os.path.join(CSRC, "a2b.hip"), "-o", tmp,
    ]
    try:
        _run(cmd, verbose)
        os.replace(tmp, out)
    finally:
        if os.path.exists(tmp):
            os.unlink(tmp)
    return out


def build_ext(verbose=True) -> str:
    import torch
    from torch.utils import cpp_extension

    out = ext_path()
    tmp = _atomic_out(out)
    torch_lib = os.path.join(os.path.dirname(torch.__file__), "lib")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    incs = []
    for p in cpp_extension.include_paths():
        incs += ["-isystem", p]
    incs += ["-isystem", sysconfig.get_paths()["include"], "-isystem",
             os.path.join(ROCM, "include"), "-I", INCLUDE]
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Werror",
        "-Wno-unused-variable", "-fvisibility=hidden",
        "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1", "-DTORCH_API_INCLUDE_EXTENSION_H",
        f"-DTORCH_EXTENSION_NAME={EXT_NAME}", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
        *incs, os.path.join(CSRC, "a2b_ext.cpp"), "-o", tmp,
        "-L", PKG, "-l:" + LIB_NAME, "-Wl,-rpath,$ORIGIN",
        "-L", torch_lib, "-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_python",
        "-Wl,-rpath," + torch_lib,
        "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
    ]
    try:
        _run(cmd, verbose)
        os.replace(tmp, out)
    finally:
        if os.path.exists(tmp):
            os.unlink(tmp)
    return out


def build(verbose=True):
    """Compile the HIP library, then the extension that links it."""
    build_lib(verbose)
    build_ext(verbose)


def built() -> bool:
    return os.path.exists(lib_path()) and os.path.exists(ext_path())


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
