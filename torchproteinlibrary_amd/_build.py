"""Build recipe for the native parts of torchproteinlibrary_amd (in-tree).

* ``libtpl_a2b.so``  -- Angles2Backbone: HIP kernels + the C ABI of
  include/tpl_a2b.h, built by ``hipcc --offload-arch=gfx950`` (cross-compiles
  without a GPU).
* ``libtpl_a2b_cpu.so`` -- Angles2Backbone on the host, FP64
  (include/tpl_a2b_cpu.h), built by g++;
* ``_ReducedModel<EXT_SUFFIX>`` -- the pybind11 torch extension that unpacks
  tensors and calls those C ABIs; links both libraries with an $ORIGIN rpath.
* ``libtpl_rmsd.so`` and ``_RMSD<EXT_SUFFIX>`` -- the same pair for
  Coords2RMSD (include/tpl_rmsd.h).
* ``libtpl_xform.so``, ``libtpl_a2c.so``, ``libtpl_pdb.so`` and
  ``_FullAtomModel<EXT_SUFFIX>`` -- the coordinate transforms
  (include/tpl_xform.h), the full-atom builder Angles2Coords
  (include/tpl_a2c.h) and the PDB reader (include/tpl_pdb.h, host code built
  by g++), one extension over the three libraries as in the reference's
  module of that name.

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
CPU_LIB_NAME = "libtpl_a2b_cpu.so"
EXT_NAME = "_ReducedModel"
RMSD_LIB_NAME = "libtpl_rmsd.so"
RMSD_EXT_NAME = "_RMSD"
XFORM_LIB_NAME = "libtpl_xform.so"
XFORM_EXT_NAME = "_FullAtomModel"
A2C_LIB_NAME = "libtpl_a2c.so"
PDB_LIB_NAME = "libtpl_pdb.so"

# ([(library, its HIP source)], extension, [its C++ sources])
PARTS = [
    ([(LIB_NAME, "a2b.hip"), (CPU_LIB_NAME, "a2b_cpu.cpp")], EXT_NAME, ["a2b_ext.cpp"]),
    ([(RMSD_LIB_NAME, "rmsd.hip")], RMSD_EXT_NAME, ["rmsd_ext.cpp"]),
    ([(XFORM_LIB_NAME, "xform.hip"), (A2C_LIB_NAME, "a2c.hip"), (PDB_LIB_NAME, "pdb.cpp")],
     XFORM_EXT_NAME, ["xform_ext.cpp", "a2c_ext.cpp", "pdb_ext.cpp"]),
]


# Per-source hipcc flags.  a2b.hip: its atomics are issued by one lane (the
# status bits, by the lanes that found an error), so LLVM's wave-level atomic
# combining gains nothing there and adds a readfirstlane, with a wait for the
# result, right after each atomic.  Without it the backward's ticket
# increment goes out before its loads and is awaited with them.
EXTRA_HIP_FLAGS = {"a2b.hip": ["-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]}


def lib_path(name: str = LIB_NAME) -> str:
    return os.path.join(PKG, name)


def ext_path(name: str = EXT_NAME) -> str:
    return os.path.join(PKG, name + sysconfig.get_config_var("EXT_SUFFIX"))


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


def build_lib(verbose=True, name: str = LIB_NAME, source: str = "a2b.hip") -> str:
    out = lib_path(name)
    tmp = _atomic_out(out)
    if source.endswith(".cpp"):  # host-only library
        cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread", "-Wall", "-Werror",
               "-I", INCLUDE, os.path.join(CSRC, source), "-o", tmp]
    else:
        cmd = [
            os.path.join(ROCM, "bin", "hipcc"), f"--offload-arch={ARCH}", "-O3", "-std=c++17",
            "-fPIC", "-shared", "-Wall", "-Werror", "-Wno-unused-command-line-argument",
            # No packed FP32 instructions anywhere in this library.
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        ] + EXTRA_HIP_FLAGS.get(source, []) + [
            "-I", INCLUDE, os.path.join(CSRC, source), "-o", tmp,
        ]
    try:
        _run(cmd, verbose)
        os.replace(tmp, out)
    finally:
        if os.path.exists(tmp):
            os.unlink(tmp)
    return out


def build_ext(verbose=True, name: str = EXT_NAME, source="a2b_ext.cpp", lib=LIB_NAME) -> str:
    """``source`` and ``lib`` are one name or a list of names."""
    sources = [source] if isinstance(source, str) else list(source)
    libs = [lib] if isinstance(lib, str) else list(lib)
    import torch
    from torch.utils import cpp_extension

    out = ext_path(name)
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
        f"-DTORCH_EXTENSION_NAME={name}", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
        *incs, *[os.path.join(CSRC, f) for f in sources], "-o", tmp,
        "-L", PKG, *["-l:" + f for f in libs], "-Wl,-rpath,$ORIGIN",
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
    """Compile each HIP library, then the extension that links it (the parts
    in parallel)."""
    def part(p):
        libs, ext, cpps = p
        with ThreadPoolExecutor(len(libs)) as pool:
            for f in [pool.submit(build_lib, verbose, lib, hip) for lib, hip in libs]:
                f.result()
        build_ext(verbose, ext, cpps, [lib for lib, _ in libs])
    with ThreadPoolExecutor(len(PARTS)) as pool:
        for f in [pool.submit(part, p) for p in PARTS]:
            f.result()


def built() -> bool:
    return all(all(os.path.exists(lib_path(lib)) for lib, _ in p[0]) and os.path.exists(ext_path(p[1]))
               for p in PARTS)


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
