"""Pins the FP64 oracle (oracle/a2b_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake fails it: a wrong element of
B(theta, beta, d), a shifted angle->atom map, a wrong bond constant, a sign
error in the derivative, an off-by-one in the gradient's summation range.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "a2b_golden.npz")
BETA = {"phi": math.pi - 1.9391, "psi": math.pi - 2.061, "omega": math.pi - 2.1186}
BOND = {"N-CA": 1.460, "CA-C": 1.525, "C-N": 1.330}


def _ragged(seed, B=6, L=11, scale=1.5):
    rng = np.random.default_rng(seed)
    a = rng.normal(0.0, scale, (B, 3, L))
    n = rng.integers(0, L + 1, B).astype(np.int32)
    n[0] = L
    g = rng.normal(0.0, 1.0, (B, 9 * L))
    for b in range(B):
        g[b, 9 * n[b]:] = 0
        a[b, :, n[b]:] = 0
    return a, n, g


# ------------------------------------------------------------ golden vectors

@pytest.mark.parametrize("case", ["helix", "zero", "random"])
def test_forward_matches_reference_golden(case):
    z = np.load(GOLDEN)
    x = O.forward(z[f"{case}_angles"], z[f"{case}_lengths"])
    np.testing.assert_allclose(x, z[f"{case}_coords"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("case", ["helix", "zero", "random"])
def test_backward_matches_reference_golden(case):
    z = np.load(GOLDEN)
    a, n, g = z[f"{case}_angles"], z[f"{case}_lengths"], z[f"{case}_grad"]
    want = z[f"{case}_grad_angles"]
    for fn in (O.backward_pairwise, O.backward_adjoint, O.backward_dense):
        got = fn(a, n, g)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-10 * max(1.0, np.abs(want).max()))


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
def test_against_reference_build_random():
    a, n, g = _ragged(11, B=7, L=13)
    n[0] = 12  # keep every chain shorter than the stride (see ref_forward_backward)
    xr, gr = O.ref_forward_backward(a, n, g)
    np.testing.assert_allclose(O.forward(a, n), xr, rtol=0, atol=1e-12)
    np.testing.assert_allclose(O.backward_adjoint(a, n, g), gr, rtol=0, atol=1e-10)


# --------------------------------------------- independent brute force (tiny)

def _rx(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1.0]])


def _ry(t):
    c, s = math.cos(t), math.sin(t)
    return np.array([[c, 0, s, 0], [0, 1, 0, 0], [-s, 0, c, 0], [0, 0, 0, 1.0]])


def _tx(d):
    m = np.eye(4)
    m[0, 3] = d
    return m


def _brute_chain(a, n):
    """Literal Ry(beta) Tx(d) Rx(theta) products, atom by atom."""
    params = [(0, BETA["phi"], BOND["N-CA"]), (1, BETA["psi"], BOND["CA-C"]),
              (2, BETA["omega"], BOND["C-N"])]
    A = np.eye(4)
    xs = [A[:3, 3].copy()]
    for i in range(1, 3 * n):
        k, beta, d = params[(i - 1) % 3]
        j = (i - 1) // 3
        A = A @ _ry(beta) @ _tx(d) @ _rx(a[k, j])
        xs.append(A[:3, 3].copy())
    return np.concatenate(xs)


def test_forward_matches_brute_force():
    a, n, _ = _ragged(3, B=5, L=8)
    x = O.forward(a, n)
    for b in range(len(n)):
        want = _brute_chain(a[b], n[b])
        np.testing.assert_allclose(x[b, : 9 * n[b]], want, rtol=0, atol=1e-12)
        assert np.all(x[b, 9 * n[b]:] == 0.0)


def test_frames_match_brute_force():
    """Per-atom frames A_i (rows of the 3x4 affine part) against the literal
    4x4 products; column 0 of A_i is the unit bond vector ending at atom i
    and the translation is atom i; zero past the chain."""
    a, n, _ = _ragged(4, B=5, L=9)
    fr = O.frames(a, n)
    x = O.forward(a, n)
    params = [(0, BETA["phi"], BOND["N-CA"]), (1, BETA["psi"], BOND["CA-C"]),
              (2, BETA["omega"], BOND["C-N"])]
    for b in range(len(n)):
        A = np.eye(4)
        for i in range(3 * n[b]):
            if i > 0:
                k, beta, d = params[(i - 1) % 3]
                A = A @ _ry(beta) @ _tx(d) @ _rx(a[b, k, (i - 1) // 3])
            np.testing.assert_allclose(fr[b, i], A[:3], rtol=0, atol=1e-12)
            np.testing.assert_array_equal(fr[b, i, :, 3], x[b, 3 * i: 3 * i + 3])
            if i > 0:
                u = x[b, 3 * i: 3 * i + 3] - x[b, 3 * i - 3: 3 * i]
                np.testing.assert_allclose(fr[b, i, :, 0], u / np.linalg.norm(u), rtol=0, atol=1e-12)
        assert np.all(fr[b, 3 * n[b]:] == 0.0)


# ------------------------------------------------------------ invariants

def _atoms(x, n):
    return x[: 9 * n].reshape(3 * n, 3)


def test_bond_lengths_and_angles():
    a, n, _ = _ragged(5, B=4, L=15)
    x = O.forward(a, n)
    for b in range(4):
        P = _atoms(x[b], n[b])
        d = np.linalg.norm(np.diff(P, axis=0), axis=1)
        want = np.tile([BOND["N-CA"], BOND["CA-C"], BOND["C-N"]], n[b])[: len(d)]
        np.testing.assert_allclose(d, want, rtol=0, atol=1e-12)
        # bond angle at atom i (between i-1 and i+1): N 1.9391, CA 2.061, C 2.1186
        v1 = P[:-2] - P[1:-1]
        v2 = P[2:] - P[1:-1]
        ang = np.arccos(np.sum(v1 * v2, 1) / np.linalg.norm(v1, axis=1) / np.linalg.norm(v2, axis=1))
        kinds = np.tile([1.9391, 2.061, 2.1186], n[b] + 1)[1: 1 + len(ang)]
        np.testing.assert_allclose(ang, kinds, rtol=0, atol=1e-10)


def test_torsions_equal_inputs():
    """The IUPAC torsions of the output reproduce phi, psi, omega with zero offset."""
    rng = np.random.default_rng(9)
    L = 12
    a = rng.uniform(-3.0, 3.0, (1, 3, L))
    n = np.array([L], np.int32)
    x = O.forward(a, n)[0]
    phi, psi, omg = O.measured_torsions(x, L)
    np.testing.assert_allclose(O.wrap_angle(phi[1:] - a[0, 0, 1:]), 0, atol=1e-10)
    np.testing.assert_allclose(O.wrap_angle(psi[:-1] - a[0, 1, :-1]), 0, atol=1e-10)
    np.testing.assert_allclose(O.wrap_angle(omg[:-1] - a[0, 2, :-1]), 0, atol=1e-10)


def test_fixed_frame():
    a, n, _ = _ragged(4, B=3, L=6)
    x = O.forward(a, n)
    b1 = BETA["phi"]
    assert np.all(x[0, :3] == 0.0)
    np.testing.assert_allclose(x[0, 3:6], [1.46 * math.cos(b1), 0.0, -1.46 * math.sin(b1)], atol=1e-14)


def test_helix_is_a_screw():
    L = 30
    a = np.zeros((1, 3, L))
    a[0, 0], a[0, 1], a[0, 2] = -1.047, -0.698, math.pi
    x = _atoms(O.forward(a, np.array([L], np.int32))[0], L)
    for c in range(3):
        d = np.linalg.norm(x[3 * 4 + c:: 3] - x[c: -3 * 4: 3][: len(x[3 * 4 + c:: 3])], axis=1)
        np.testing.assert_allclose(d, d[0], atol=1e-10)


def test_zero_angles_planar():
    a = np.zeros((1, 3, 10))
    x = _atoms(O.forward(a, np.array([10], np.int32))[0], 10)
    np.testing.assert_allclose(x[:, 1], 0.0, atol=1e-13)


def test_prefix_property_and_padding():
    a, n, _ = _ragged(8, B=4, L=9)
    x9 = O.forward(a, n)
    a2 = np.concatenate([a, np.zeros((4, 3, 5))], axis=2)
    x14 = O.forward(a2, n)
    for b in range(4):
        np.testing.assert_array_equal(x14[b, : 9 * 9], x9[b])
        assert np.all(x14[b, 9 * 9:] == 0)


def test_lengths_clamped():
    a, n, _ = _ragged(2, B=3, L=5)
    n2 = np.array([-3, 99, 2], np.int32)
    x = O.forward(a, n2)
    assert np.all(x[0] == 0)
    np.testing.assert_array_equal(x[1], O.forward(a[1:2], np.array([5], np.int32))[0])


# ------------------------------------------------------------ backward

def _fd(a, n, g, h=1e-6):
    fd = np.zeros_like(a)
    for b in range(a.shape[0]):
        for k in range(3):
            for j in range(a.shape[2]):
                ap, am = a.copy(), a.copy()
                ap[b, k, j] += h
                am[b, k, j] -= h
                fd[b, k, j] = (O.forward(ap, n)[b] - O.forward(am, n)[b]) @ g[b] / (2 * h)
    return fd


def test_backward_matches_finite_differences():
    a, n, g = _ragged(21, B=4, L=7)
    fd = _fd(a, n, g)
    for fn in (O.backward_pairwise, O.backward_adjoint, O.backward_dense):
        got = fn(a, n, g)
        np.testing.assert_allclose(got, fd, rtol=0, atol=1e-6 * max(1, np.abs(fd).max()))


def test_backward_exact_zeros_and_phi0():
    a, n, g = _ragged(22, B=5, L=9)
    n[:] = [9, 4, 1, 0, 2]
    for fn in (O.backward_pairwise, O.backward_adjoint, O.backward_dense):
        ga = fn(a, n, g)
        for b in range(5):
            if n[b] > 0:
                assert ga[b, 1, n[b] - 1] == 0.0 and ga[b, 2, n[b] - 1] == 0.0
            assert np.all(ga[b, :, n[b]:] == 0.0)
        assert ga[0, 0, 0] != 0.0  # the frame is anchored at N0: phi_0 moves atoms


def test_backward_linear_in_grad():
    a, n, g = _ragged(23, B=3, L=8)
    g2 = np.random.default_rng(1).normal(size=g.shape) * (g != 0)
    lhs = O.backward_adjoint(a, n, 2.0 * g - 3.0 * g2)
    rhs = 2.0 * O.backward_adjoint(a, n, g) - 3.0 * O.backward_adjoint(a, n, g2)
    np.testing.assert_allclose(lhs, rhs, atol=1e-9)
