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


def _axis_quat(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis)
    return np.concatenate([[np.cos(angle / 2)], np.sin(angle / 2) * axis])


# Rotations whose quaternions have zero components: the identity, single-axis
# rotations, and 180-degree turns (q0 = 0).  With random rotations every
# component is O(1), so these are the cases where picking the adjugate row
# of a zero component (rounding noise only) would show.
SPECIAL_Q = {
    "identity": np.array([1.0, 0.0, 0.0, 0.0]),
    "z0.7": _axis_quat([0, 0, 1], 0.7),
    "x0.7": _axis_quat([1, 0, 0], 0.7),
    "y-1.3": _axis_quat([0, 1, 0], -1.3),
    "z180": _axis_quat([0, 0, 1], np.pi),
    "x180": _axis_quat([1, 0, 0], np.pi),
    "xy180": _axis_quat([1, 1, 0], np.pi),
}


def _solve(lib, xc, yc):
    Rm = xc.T @ yc
    bound = 0.5 * ((xc * xc).sum() + (yc * yc).sum())
    lam, q, deg = _eig(lib, Rm, bound)
    U = np.zeros(9)
    lib.rmsd_math_u(_ptr(np.ascontiguousarray(q)), _ptr(U))
    r2 = (xc * xc).sum() + (yc * yc).sum()
    rmsd = np.sqrt(max(r2 - 2 * abs(lam), 0.0) / len(xc))
    return lam, q, U.reshape(3, 3), rmsd, deg


@pytest.mark.parametrize("name", sorted(SPECIAL_Q))
def test_exact_rotations_with_zero_quaternion_components(lib, name):
    """y = Q x exactly: RMSD 0 and U = Q (the quaternion's zero components
    must not be the adjugate row the eigenvector is read from)."""
    Q = R.u_from_q(SPECIAL_Q[name])
    rng = np.random.default_rng(11)
    for k in (3, 10, 60):
        xs = np.cumsum(rng.normal(size=(k, 3)) * 2.0, axis=0)
        xc = xs - xs.mean(0)
        yc = xc @ Q.T
        lam, q, U, rmsd, deg = _solve(lib, xc, yc)
        assert np.all(np.isfinite(q)), (name, k)
        assert abs(np.linalg.norm(q) - 1) <= 1e-12
        assert np.abs(U - Q).max() <= 1e-9, (name, k, U, Q)
        assert rmsd <= 1e-6 * np.sqrt((xc * xc).sum() / k), (name, k, rmsd)


@pytest.mark.parametrize("name", sorted(SPECIAL_Q))
def test_noisy_special_rotations_match_kabsch(lib, name):
    """y = Q x + noise: the solve gives the Kabsch (SVD) rotation and RMSD."""
    Q = R.u_from_q(SPECIAL_Q[name])
    rng = np.random.default_rng(12)
    for k in (4, 25, 200):
        xs = np.cumsum(rng.normal(size=(k, 3)) * 3.8, axis=0)
        ys = xs @ Q.T + rng.normal(scale=0.3, size=(k, 3))
        xc, yc = xs - xs.mean(0), ys - ys.mean(0)
        lam, q, U, rmsd, deg = _solve(lib, xc, yc)
        r_k, Rot = R.kabsch(xs.reshape(-1), ys.reshape(-1), k)
        assert np.abs(U - Rot).max() <= 1e-8, (name, k)
        assert abs(rmsd - r_k) <= 1e-9 * max(r_k, 1.0), (name, k, rmsd, r_k)


def test_inverted_copy_matches_kabsch(lib):
    """y = -x (no proper rotation fits): the best rotation is the half turn
    about the axis of least spread (q0 = 0); RMSD and U as Kabsch gives."""
    rng = np.random.default_rng(13)
    for k in (5, 40):
        xs = rng.normal(size=(k, 3)) * np.array([5.0, 2.0, 0.7])
        xc = xs - xs.mean(0)
        yc = -xc
        lam, q, U, rmsd, deg = _solve(lib, xc, yc)
        r_k, Rot = R.kabsch(xc.reshape(-1), yc.reshape(-1), k)
        assert abs(q[0]) <= 1e-12
        assert np.abs(U - Rot).max() <= 1e-8
        assert abs(rmsd - r_k) <= 1e-9 * r_k
