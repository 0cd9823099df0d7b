"""ctypes binding of libtpl_a2b.so, declared from include/tpl_a2b.h.  This is the
binding a ctypes user of the C ABI writes."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tpl_a2b.h")
LIB = os.path.join(ROOT, "torchproteinlibrary_amd", "libtpl_a2b.so")


class Config(ctypes.Structure):
    _fields_ = [("fwd_res_per_lane", ctypes.c_int32), ("bwd_res_per_lane", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("reserved", ctypes.c_int32)]


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tpl_a2b_\w+)\s*\(", text)))


def load():
    L = ctypes.CDLL(LIB)
    vp, i32, cfgp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Config)
    L.tpl_a2b_strerror.restype = ctypes.c_char_p
    L.tpl_a2b_strerror.argtypes = [i32]
    L.tpl_a2b_default_config.argtypes = [i32, i32, cfgp]
    L.tpl_a2b_chunk_residues.restype = ctypes.c_int64
    L.tpl_a2b_chunk_residues.argtypes = [cfgp, i32]
    L.tpl_a2b_sync_clear.argtypes = [vp, ctypes.c_size_t, vp]
    L.tpl_a2b_forward_noalloc.argtypes = [i32, vp, vp, vp, vp, i32, i32, cfgp, vp]
    L.tpl_a2b_noalloc_max_bytes.restype = ctypes.c_size_t
    L.tpl_a2b_noalloc_max_bytes.argtypes = []
    for f in (L.tpl_a2b_state_bytes, L.tpl_a2b_sync_bytes):
        f.restype = ctypes.c_size_t
        f.argtypes = [i32, i32, i32, cfgp]
    L.tpl_a2b_plain_frames_bytes.restype = ctypes.c_size_t
    L.tpl_a2b_plain_frames_bytes.argtypes = [i32, i32, i32]
    L.tpl_a2b_forward.argtypes = [i32, vp, vp, vp, vp, vp, vp, i32, i32, cfgp, vp]
    L.tpl_a2b_backward.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, i32, i32, cfgp, vp]
    L.tpl_a2b_forward_plain.argtypes = [i32, vp, vp, vp, vp, vp, i32, i32, vp]
    L.tpl_a2b_backward_plain.argtypes = [i32, vp, vp, vp, vp, vp, vp, i32, i32, vp]
    L.tpl_a2b_tune.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, i32, i32, cfgp, i32, i32,
                               cfgp, ctypes.POINTER(ctypes.c_float), vp]
    return L

