"""ctypes binding of libtpl_rmsd.so (include/tpl_rmsd.h), as a maintainer of a
C/ctypes caller would write it."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tpl_rmsd.h")
LIB = os.path.join(ROOT, "torchproteinlibrary_amd", "libtpl_rmsd.so")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tpl_rmsd_\w+)\s*\(", text)))


def load():
    L = ctypes.CDLL(LIB)
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    L.tpl_rmsd_strerror.restype = ctypes.c_char_p
    L.tpl_rmsd_strerror.argtypes = [i32]
    L.tpl_rmsd_state_bytes.restype = ctypes.c_size_t
    L.tpl_rmsd_state_bytes.argtypes = [i32]
    L.tpl_rmsd_forward.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp]
    L.tpl_rmsd_backward.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp]
    return L
