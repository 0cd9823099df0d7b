"""CPU pin of the device eigen-solve: torchproteinlibrary_amd/csrc/rmsd_math.h,
compiled for the host with g++ from the same source the HIP kernel uses,
against numpy.linalg.eigh and the oracle's F and U."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle import rmsd_oracle as R

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("rmsd_math") / "librmsd_math_host.so"
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Werror",
                           os.path.join(HERE, "rmsd_math_host.cpp"), "-o", str(out)])
    L = ctypes.CDLL(str(out))
    dp = ctypes.POINTER(ctypes.c_double)
    L.rmsd_math_eig.argtypes = [dp, ctypes.c_double, dp]
    L.rmsd_math_f.argtypes = [dp, dp]
    L.rmsd_math_u.argtypes = [dp, dp]
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _eig(lib, Rm, bound):
    Rm = np.ascontiguousarray(Rm, dtype=np.float64)
    out = np.zeros(6)
    lib.rmsd_math_eig(_ptr(Rm), float(bound), _ptr(out))
    return out[0], out[1:5], bool(out[5])


def _centred(x, y, b, k):
    xs, ys = x[b, : 3 * k].reshape(k, 3), y[b, : 3 * k].reshape(k, 3)
    return xs - xs.mean(0), ys - ys.mean(0)


def test_f_and_u_match_the_oracle(lib):
    rng = np.random.default_rng(0)
    for _ in range(20):
        Rm = rng.normal(size=(3, 3))
        F = np.zeros(16)
        lib.rmsd_math_f(_ptr(np.ascontiguousarray(Rm)), _ptr(F))
        assert np.array_equal(F.reshape(4, 4), R.f_matrix(Rm))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        U = np.zeros(9)
        lib.rmsd_math_u(_ptr(q), _ptr(U))
        assert np.allclose(U.reshape(3, 3), R.u_from_q(q), atol=1e-15)


@pytest.mark.parametrize("noise", [0.5, 1e-3, 0.0])
def test_largest_eigenpair_matches_eigh(lib, noise):
    x, y, n, _ = R.workload(400, 150, seed=31, min_atoms=3, noise=noise)
    for b in range(400):
        xc, yc = _centred(x, y, b, int(n[b]))
        Rm = xc.T @ yc
        bound = 0.5 * ((xc * xc).sum() + (yc * yc).sum())
        lam, q, deg = _eig(lib, Rm, bound)
        w, V = np.linalg.eigh(R.f_matrix(Rm))
        assert not deg
        assert abs(lam - w[-1]) <= 1e-13 * w[-1], b
        assert min(np.abs(q - V[:, -1]).max(), np.abs(q + V[:, -1]).max()) <= 1e-10, b


def test_double_root_gives_a_top_eigenvector(lib):
    """Two atoms: the rotation about their axis is free and the largest
    eigenvalue of F is double.  The shifted adjugate still returns a unit
    vector of the top eigenspace."""
    rng = np.random.default_rng(2)
    for _ in range(50):
        xc = rng.normal(size=3)
        xc = np.stack([xc, -xc])
        yc = rng.normal(size=3)
        yc = np.stack([yc, -yc])
        Rm = xc.T @ yc
        F = R.f_matrix(Rm)
        lam, q, deg = _eig(lib, Rm, 0.5 * ((xc * xc).sum() + (yc * yc).sum()))
        w = np.linalg.eigvalsh(F)
        assert deg
        assert abs(lam - w[-1]) <= 1e-12 * w[-1]
        assert abs(np.linalg.norm(q) - 1) <= 1e-12
        assert np.linalg.norm(F @ q - lam * q) <= 1e-6 * lam


def test_zero_matrix_gives_identity(lib):
    lam, q, deg = _eig(lib, np.zeros((3, 3)), 0.0)
    assert lam == 0.0 and np.array_equal(q, [1.0, 0.0, 0.0, 0.0]) and not deg
