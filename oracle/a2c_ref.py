"""ctypes binding of oracle/_ref/libtplref_a2c.so, the reference's own
full-atom conformation code compiled from its sources (oracle/Makefile,
target ref_a2c; entry points in oracle/ref_a2c_shim.cpp).

TEST INFRASTRUCTURE ONLY: used by tests/ to pin oracle/a2c_oracle.py."""
import ctypes
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libtplref_a2c.so")
_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


def available() -> bool:
    return os.path.exists(PATH)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(PATH)
        L.tplref_a2c_forward.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, _dp]
        L.tplref_a2c_names.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p]
        L.tplref_a2c_backward.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, _dp, _dp]
        _lib = L
    return _lib


def forward_one(seq, angles):
    """coords [3 * atoms] of one sequence for angles [8, stride]."""
    a = np.ascontiguousarray(angles, dtype=np.float64)
    out = np.zeros(3 * 16 * (len(seq) + 1))
    n = lib().tplref_a2c_forward(seq.encode(), a.ctypes.data_as(_dp), a.shape[1], out.ctypes.data_as(_dp))
    return out[:3 * n]


def names_one(seq):
    n = 16 * (len(seq) + 1)
    an, rn = ctypes.create_string_buffer(4 * n), ctypes.create_string_buffer(n)
    k = lib().tplref_a2c_names(seq.encode(), len(seq) + 1, an, rn)
    return ([an.raw[4 * i:4 * i + 4].rstrip(b"\0").decode() for i in range(k)],
            [chr(c) for c in rn.raw[:k]])


def backward_one(seq, angles, grad_coords):
    a = np.ascontiguousarray(angles, dtype=np.float64)
    g = np.zeros(3 * 16 * (len(seq) + 1))
    g[:len(grad_coords)] = grad_coords
    ga = np.zeros(a.shape)
    lib().tplref_a2c_backward(seq.encode(), a.ctypes.data_as(_dp), a.shape[1], g.ctypes.data_as(_dp),
                              ga.ctypes.data_as(_dp))
    return ga
