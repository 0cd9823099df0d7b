"""ctypes binding of libtpl_a2c.so (include/tpl_a2c.h)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tpl_a2c.h")
LIB = os.path.join(ROOT, "torchproteinlibrary_amd", "libtpl_a2c.so")


def declared_functions():
    return sorted(set(re.findall(r"\b(tpl_a2c_\w+)\s*\(", open(HEADER).read())))


def load():
    L = ctypes.CDLL(LIB)
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    L.tpl_a2c_strerror.restype = ctypes.c_char_p
    L.tpl_a2c_strerror.argtypes = [i32]
    L.tpl_a2c_num_atoms.argtypes = [ctypes.c_char_p]
    L.tpl_a2c_forward.argtypes = [i32, vp, i32, vp, i32, vp, vp, vp, i32, vp, i32, vp]
    L.tpl_a2c_backward.argtypes = [i32, vp, i32, vp, vp, i32, vp, i32, vp, i32, vp]
    L.tpl_a2c_backward_form.argtypes = [i32, vp, i32, vp, i32, vp, vp, i32, vp, i32, vp, i32, vp]
    L.tpl_a2c_forward_ordered.argtypes = [vp, i32, vp, i32, vp, i32, vp, vp, vp, i32, vp, i32, vp]
    return L
