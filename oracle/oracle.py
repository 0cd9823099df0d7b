"""FP64 CPU oracle for ReducedModel.Angles2Backbone -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, as the checker (and as the reported CPU
baseline).  The product package never imports it.

Two CPU implementations sit behind it:

* ``liba2b_oracle.so`` -- this repository's plain-C FP64 restatement of the
  reference CPU algorithm (oracle/a2b_oracle.c; each function there cites the
  reference file:line it follows).
* ``_ref/libtplref.so`` -- the reference's own CPU kernels
  (Layers/ReducedModel/cBackboneProteinCPUKernels.cpp with Math/cMatrix44.cpp,
  cMatrix33.cpp, cVector3.cpp), compiled from the reference sources by
  oracle/Makefile, wrapped by oracle/ref_shim.cpp.  Optional: present where it
  was built.

Array conventions follow the reference API: angles ``[B,3,L]`` float64 rows
(phi, psi, omega), lengths int32 ``[B]``, coords ``[B,9L]`` float64.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liba2b_oracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libtplref.so")

_lib = None
_ref = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


def _ensure_built():
    if not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", _HERE, "oracle"])


def lib():
    """Load (building once if absent) the C restatement."""
    global _lib
    if _lib is None:
        _ensure_built()
        L = ctypes.CDLL(_LIB_PATH)
        for name in ("a2bo_backward_pairwise", "a2bo_backward_adjoint"):
            getattr(L, name).argtypes = [_dp, _ip, _dp, _dp, ctypes.c_int, ctypes.c_int]
        L.a2bo_forward.argtypes = [_dp, _ip, _dp, ctypes.c_int, ctypes.c_int]
        L.a2bo_frames.argtypes = [_dp, _ip, _dp, ctypes.c_int, ctypes.c_int]
        L.a2bo_backward_dense.argtypes = [_dp, _ip, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(_REF_PATH)


def ref():
    """Load the reference's own CPU kernels (built under oracle/_ref)."""
    global _ref
    if _ref is None:
        R = ctypes.CDLL(_REF_PATH)
        R.tplref_forward.argtypes = [_dp, _dp, _dp, _ip, ctypes.c_int, ctypes.c_int]
        R.tplref_backward.argtypes = [_dp, _dp, _dp, _ip, _dp, _dp, ctypes.c_int, ctypes.c_int]
        _ref = R
    return _ref


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def forward(angles, lengths):
    """coords [B,9L] float64 (cBackboneProteinCPUKernels.cpp:14-60)."""
    a, n = _f64(angles), _i32(lengths)
    B, _, L = a.shape
    out = np.empty((B, 9 * L), np.float64)
    lib().a2bo_forward(_p(a), _p(n, _ip), _p(out), B, L)
    return out


def frames(angles, lengths):
    """Per-atom frames [B,3L,3,4] (rows of A_i without the homogeneous row)."""
    a, n = _f64(angles), _i32(lengths)
    B, _, L = a.shape
    out = np.empty((B, 3 * L, 12), np.float64)
    lib().a2bo_frames(_p(a), _p(n, _ip), _p(out), B, L)
    return out.reshape(B, 3 * L, 3, 4)


def backward_pairwise(angles, lengths, grad_coords):
    """Reference per-pair O(L^2) gradient (cBackboneProteinCPUKernels.cpp:62-145)."""
    a, n, g = _f64(angles), _i32(lengths), _f64(grad_coords)
    B, _, L = a.shape
    out = np.empty((B, 3, L), np.float64)
    lib().a2bo_backward_pairwise(_p(a), _p(n, _ip), _p(g), _p(out), B, L)
    return out


def backward_adjoint(angles, lengths, grad_coords):
    """O(L) closed-form gradient, long-double accumulation."""
    a, n, g = _f64(angles), _i32(lengths), _f64(grad_coords)
    B, _, L = a.shape
    out = np.empty((B, 3, L), np.float64)
    lib().a2bo_backward_adjoint(_p(a), _p(n, _ip), _p(g), _p(out), B, L)
    return out


def backward_dense(angles, lengths, grad_coords):
    """Reference memory behaviour: zero-filled [B,3,9L^2] tensor, then reduce."""
    a, n, g = _f64(angles), _i32(lengths), _f64(grad_coords)
    B, _, L = a.shape
    out = np.empty((B, 3, L), np.float64)
    dr = np.empty(B * 3 * 9 * L * L, np.float64)
    lib().a2bo_backward_dense(_p(a), _p(n, _ip), _p(g), _p(out), _p(dr), B, L)
    return out


def ref_forward_backward(angles, lengths, grad_coords):
    """Run the reference's own CPU kernels: returns (coords, grad_angles).

    Buffers are allocated and zero-filled the way
    TorchProteinLibrary/ReducedModel/Angles2Backbone/Angles2Backbone.py:64-65,83-84
    does.  A gets 16 spare zero doubles at its end: the reference reads frame
    A[3n] for the last omega of a chain (cBackboneProteinCPUKernels.cpp:83), which
    is past the buffer for the last chain when n == L.
    """
    a, n, g = _f64(angles), _i32(lengths), _f64(grad_coords)
    B, _, L = a.shape
    coords = np.zeros((B, 9 * L), np.float64)
    A = np.zeros(B * 48 * L + 16, np.float64)
    ga = np.zeros((B, 3, L), np.float64)
    dr = np.zeros(B * 3 * 9 * L * L, np.float64)
    R = ref()
    R.tplref_forward(_p(a), _p(coords), _p(A), _p(n, _ip), B, L)
    R.tplref_backward(_p(ga), _p(g), _p(a), _p(n, _ip), _p(A), _p(dr), B, L)
    return coords, ga


# ---------------------------------------------------------------- geometry


def dihedral(p0, p1, p2, p3):
    """Textbook IUPAC torsion angle (atan2 form), vectorised over leading dims."""
    b0 = p0 - p1
    b1 = p2 - p1
    b2 = p3 - p2
    b1n = b1 / np.linalg.norm(b1, axis=-1, keepdims=True)
    v = b0 - np.sum(b0 * b1n, -1, keepdims=True) * b1n
    w = b2 - np.sum(b2 * b1n, -1, keepdims=True) * b1n
    x = np.sum(v * w, -1)
    y = np.sum(np.cross(b1n, v) * w, -1)
    return np.arctan2(y, x)


def measured_torsions(coords_chain, n):
    """phi_j, psi_j, omega_j measured on one chain's coords [9L] (first n residues).

    phi_j = (C_{j-1}, N_j, CA_j, C_j), psi_j = (N_j, CA_j, C_j, N_{j+1}),
    omega_j = (CA_j, C_j, N_{j+1}, CA_{j+1}).  Entries whose atoms do not exist
    are NaN.
    """
    X = np.asarray(coords_chain, np.float64)[: 9 * n].reshape(n, 3, 3)
    N, CA, C = X[:, 0], X[:, 1], X[:, 2]
    phi = np.full(n, np.nan)
    psi = np.full(n, np.nan)
    omg = np.full(n, np.nan)
    if n >= 2:
        phi[1:] = dihedral(C[:-1], N[1:], CA[1:], C[1:])
        psi[:-1] = dihedral(N[:-1], CA[:-1], C[:-1], N[1:])
        omg[:-1] = dihedral(CA[:-1], C[:-1], N[1:], CA[1:])
    return phi, psi, omg


def wrap_angle(a):
    return (np.asarray(a) + math.pi) % (2 * math.pi) - math.pi
