"""Pins of the coordinate-transform oracle (oracle/xform_oracle.py), on the CPU."""
import numpy as np

from oracle import rmsd_oracle as R
from oracle import xform_oracle as X


def _data(B=6, A=40, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(B, 3 * A)) * 10
    n = rng.integers(0, A + 1, size=B)
    n[0], n[1] = 0, A
    Rm = np.stack([R.random_rotation(rng) for _ in range(B)])
    T = rng.normal(size=(B, 3)) * 5
    return x, n, Rm, T


def test_round_trips():
    x, n, Rm, T = _data()
    xm = X.translate_backward(x, n)  # x with zeros past n
    assert np.allclose(X.translate(X.translate(x, T, n), -T, n), xm, atol=1e-12)
    assert np.allclose(X.rotate(X.rotate(x, Rm, n), Rm, n, transpose=True), xm, atol=1e-12)


def test_centre_moves_with_the_structure():
    x, n, Rm, T = _data(seed=1)
    c = X.center(x, n)
    assert np.all(np.isnan(c[0]))  # 0 atoms: 0/0, as the reference's 1/n
    ok = n > 0
    assert np.allclose(X.center(X.translate(x, T, n), n)[ok], (c + T)[ok], atol=1e-12)
    assert np.allclose(X.center(X.rotate(x, Rm, n), n)[ok], np.einsum("brc,bc->br", Rm, c)[ok], atol=1e-12)
    # translating by minus the centre centres it (reference test.py TestCoordsTranslateForward)
    assert np.allclose(X.center(X.translate(x, -np.nan_to_num(c), n), n)[ok], 0.0, atol=1e-12)


def test_backwards_are_the_adjoints():
    """<F x, g> = <x, F^T g> for each linear map, exactly what central
    differences give for a linear function."""
    x, n, Rm, T = _data(seed=2)
    B, A = x.shape[0], x.shape[1] // 3
    g = np.random.default_rng(3).normal(size=x.shape)
    assert np.isclose((X.rotate(x, Rm, n) * g).sum(), (x * X.rotate(g, Rm, n, transpose=True)).sum())
    assert np.isclose((X.translate_backward(x, n) * g).sum(), (x * X.translate_backward(g, n)).sum())
    gc = np.random.default_rng(4).normal(size=(B, 3))
    ok = n > 0
    lhs = (np.nan_to_num(X.center(x, n)) * gc)[ok].sum()
    rhs = (x * X.center_backward(gc, n, A))[ok].sum()
    assert np.isclose(lhs, rhs)
