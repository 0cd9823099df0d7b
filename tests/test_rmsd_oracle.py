"""Pins of the FP64 Coords2RMSD oracle (oracle/rmsd_oracle.py), on the CPU.

The reference holds no stored vectors for this layer; its tests are
tolerance-only (UnitTests/RMSD/Coords2RMSD/test.py:30-49: zero RMSD for a
rigidly moved copy; :67-97: gradcheck).  Here the restatement is pinned by an
independent solution of the same problem (Kabsch/SVD), rigid-motion
invariances, and finite differences.
"""
import numpy as np
import pytest

from oracle import rmsd_oracle as R


def test_horn_matches_kabsch():
    """Two different methods for the same least-squares optimum: the largest
    eigenpair of the 4x4 quaternion matrix (the reference's) and the SVD of
    the 3x3 correlation.  Same RMSD, same rotation."""
    x, y, n, _ = R.workload(24, 120, seed=1, min_atoms=3)
    for b in range(24):
        k = int(n[b])
        rmsd, UT, U, _, _ = R.forward_one(x[b], y[b], k)
        rk, rot = R.kabsch(x[b], y[b], k)
        # rmsd carries the reference's +1e-10 under the square root
        assert abs(rmsd ** 2 - (rk ** 2 + R.EPS_LIMIT / k)) <= 1e-9 * max(rk ** 2, 1.0), b
        # U takes the centred input onto the centred target, like Kabsch's
        assert np.abs(U - rot).max() <= 1e-8, b
        assert np.allclose(UT, U.T)


def test_rigid_copy_has_zero_rmsd():
    """The reference's own check (test.py:30-49): a rotated and translated
    copy is at RMSD 0 (here sqrt(1e-10 / n) from the reference's guard term)."""
    rng = np.random.default_rng(3)
    x, _, n, _ = R.workload(16, 64, seed=3, min_atoms=16)
    for b in range(16):
        k = int(n[b])
        xs = x[b, : 3 * k].reshape(k, 3)
        ys = xs @ R.random_rotation(rng).T + rng.normal(size=3) * 10
        y = np.zeros_like(x[b])
        y[: 3 * k] = ys.reshape(-1)
        rmsd = R.forward_one(x[b], y, k)[0]
        assert abs(rmsd - np.sqrt(R.EPS_LIMIT / k)) < 1e-6


def test_invariances():
    """RMSD does not change when either structure moves rigidly, and the
    optimal rotation follows the target's rotation."""
    rng = np.random.default_rng(5)
    x, y, n, _ = R.workload(8, 50, seed=5, min_atoms=4)
    for b in range(8):
        k = int(n[b])
        r0, _, U0, _, _ = R.forward_one(x[b], y[b], k)
        Q = R.random_rotation(rng)
        y2 = np.zeros_like(y[b])
        y2[: 3 * k] = (y[b, : 3 * k].reshape(k, 3) @ Q.T + rng.normal(size=3)).reshape(-1)
        x2 = np.zeros_like(x[b])
        x2[: 3 * k] = (x[b, : 3 * k].reshape(k, 3) + rng.normal(size=3)).reshape(-1)
        r1, _, U1, _, _ = R.forward_one(x2, y2, k)
        assert abs(r1 - r0) <= 1e-9 * r0
        assert np.abs(U1 - Q @ U0).max() <= 1e-8


def _fd(x, y, n, g, h=1e-6):
    out = np.zeros_like(x)
    for b in range(x.shape[0]):
        for j in range(3 * int(n[b])):
            xp, xm = x.copy(), x.copy()
            xp[b, j] += h
            xm[b, j] -= h
            rp = R.forward(xp[b:b + 1], y[b:b + 1], n[b:b + 1])[0][0]
            rm = R.forward(xm[b:b + 1], y[b:b + 1], n[b:b + 1])[0][0]
            out[b, j] = g[b] * (rp - rm) / (2 * h)
    return out


def test_backward_matches_central_differences():
    x, y, n, g = R.workload(4, 12, seed=7, min_atoms=3)
    ga = R.backward(x, y, n, g)
    fd = _fd(x, y, n, g)
    assert np.abs(ga - fd).max() <= 1e-7 * np.abs(fd).max()
    # slots past num_atoms get exactly 0
    for b in range(4):
        assert not ga[b, 3 * int(n[b]):].any()


def test_reference_example_structure():
    """The three-atom example of the reference's gradient script
    (UnitTests/RMSD/Coords2RMSD/test_backward.py:111-141), checked against
    central differences instead of a plot."""
    x = np.array([[2, 0, 0, 3, 0.1, 1, 3, 0.1, 0]], dtype=np.float64)
    t = np.array([[0, 0, 0, 1, 1, 0, 3, 0, 0]], dtype=np.float64)
    n = np.array([3], dtype=np.int32)
    g = np.array([1.0])
    ga = R.backward(x, t, n, g)
    fd = _fd(x, t, n, g)
    assert np.abs(ga - fd).max() <= 1e-7
    # the centred gradient sums to zero over the atoms (translation invariance)
    assert np.abs(ga.reshape(3, 3).sum(0)).max() <= 1e-12


@pytest.mark.parametrize("k", [0, 1, 2])
def test_degenerate_counts(k):
    x, y, _, _ = R.workload(1, 4, seed=11)
    rmsd, UT, _, _, _ = R.forward_one(x[0], y[0], k)
    if k == 0:
        assert rmsd == np.inf
    elif k == 1:
        assert rmsd == pytest.approx(np.sqrt(R.EPS_LIMIT))
        assert np.array_equal(UT, np.eye(3))
    else:
        assert np.isfinite(rmsd)
    assert not R.backward(x[:1], y[:1], np.array([k]), np.ones(1))[0, 3 * k:].any()
