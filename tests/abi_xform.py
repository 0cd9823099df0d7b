"""ctypes binding of libtpl_xform.so (include/tpl_xform.h)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tpl_xform.h")
LIB = os.path.join(ROOT, "torchproteinlibrary_amd", "libtpl_xform.so")


def declared_functions():
    return sorted(set(re.findall(r"\b(tpl_xform_\w+)\s*\(", open(HEADER).read())))


def load():
    L = ctypes.CDLL(LIB)
    vp, i32 = ctypes.c_void_p, ctypes.c_int
    L.tpl_xform_strerror.restype = ctypes.c_char_p
    L.tpl_xform_strerror.argtypes = [i32]
    L.tpl_xform_translate.argtypes = [i32, vp, vp, vp, vp, vp, i32, i32, vp]
    L.tpl_xform_translate_backward.argtypes = [i32, vp, vp, vp, vp, i32, i32, vp]
    L.tpl_xform_rotate.argtypes = [i32, vp, vp, vp, i32, vp, vp, i32, i32, vp]
    L.tpl_xform_center.argtypes = [i32, vp, vp, vp, vp, i32, i32, vp]
    L.tpl_xform_center_backward.argtypes = [i32, vp, vp, vp, vp, i32, i32, vp]
    return L
