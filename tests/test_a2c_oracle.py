"""Pins the FP64 Angles2Coords oracle (oracle/a2c_oracle.py):
* against the reference's own CPU code, compiled from its sources into
  oracle/_ref (coordinates, atom names, gradients; all 20 residue types);
* against geometry the reference fixes: bond lengths and angles of the
  backbone, backbone torsions equal to the input angles, rigid side-chain
  groups;
* its backbone against the Angles2Backbone oracle (the reference's own
  cross-check, UnitTests/ReducedModel/Angles2Backbone/test.py:70);
* its gradient against central finite differences, and the reference's own
  backward test (UnitTests/FullAtomModel/Angles2Coords/test.py:83-142)."""
import numpy as np
import pytest

from oracle import a2c_oracle as O
from oracle import a2c_ref
from oracle import oracle as A2B

REF = pytest.mark.skipif(not a2c_ref.available(), reason="reference build oracle/_ref absent")


def _angles(rng, L, stride=None):
    a = np.zeros((8, stride or L))
    a[0, :L] = rng.normal(-1.1, 0.5, L)
    a[1, :L] = rng.normal(2.3, 0.5, L)
    a[2, :L] = np.pi + rng.normal(0, 0.1, L)
    a[3:, :L] = rng.uniform(-np.pi, np.pi, (5, L))
    return a


@REF
@pytest.mark.parametrize("seed", range(4))
def test_matches_reference_build(seed):
    rng = np.random.default_rng(seed)
    seq = O.AA20 if seed == 0 else "".join(rng.choice(list(O.AA20), size=int(rng.integers(1, 60))))
    ang = _angles(rng, len(seq), len(seq) + 2)
    c_ref = a2c_ref.forward_one(seq, ang)
    c, _, _ = O.forward_one(seq, ang)
    assert c.size == c_ref.size == 3 * O.num_atoms(seq)
    assert np.abs(c - c_ref).max() < 1e-12
    names, res1 = a2c_ref.names_one(seq)
    res3, atoms = O.names_one(seq)
    assert names == atoms
    assert [O.RES3[r] for r in res1] == res3
    g = rng.normal(size=c.size)
    ga_ref = a2c_ref.backward_one(seq, ang, g)
    ga = O.backward_one(seq, ang, g)
    assert np.abs(ga - ga_ref).max() <= 1e-11 * max(1.0, np.abs(ga_ref).max())


def _backbone(seq, coords):
    idx, k = [], 0
    for aa in seq:
        idx.append(k)
        k += len(O.residue_atoms(aa))
    X = coords.reshape(-1, 3)
    N = X[idx]
    CA = X[[i + 1 for i in idx]]
    C = X[[i + len(O.residue_atoms(aa)) - 2 for i, aa in zip(idx, seq)]]
    return N, CA, C


def test_backbone_geometry_and_torsions():
    rng = np.random.default_rng(11)
    seq = "".join(rng.choice(list(O.AA20), size=40))
    ang = _angles(rng, len(seq))
    c, _, _ = O.forward_one(seq, ang)
    N, CA, C = _backbone(seq, c)
    assert np.allclose(np.linalg.norm(CA - N, axis=1), 1.460)
    assert np.allclose(np.linalg.norm(C - CA, axis=1), 1.525)
    assert np.allclose(np.linalg.norm(N[1:] - C[:-1], axis=1), 1.330)
    ang_at = lambda a, b, c_: np.arccos(np.sum((a - b) * (c_ - b), 1) /  # noqa: E731
                                        np.linalg.norm(a - b, axis=1) / np.linalg.norm(c_ - b, axis=1))
    assert np.allclose(ang_at(N, CA, C), 2.061)
    assert np.allclose(ang_at(C[:-1], N[1:], CA[1:]), 1.9391)
    assert np.allclose(ang_at(CA[:-1], C[:-1], N[1:]), 2.1186)
    w = A2B.wrap_angle
    phi = A2B.dihedral(C[:-1], N[1:], CA[1:], C[1:])
    psi = A2B.dihedral(N[:-1], CA[:-1], C[:-1], N[1:])
    omg = A2B.dihedral(CA[:-1], C[:-1], N[1:], CA[1:])
    assert np.abs(w(phi - ang[0, 1:])).max() < 1e-9
    assert np.abs(w(psi - ang[1, :-1])).max() < 1e-9
    assert np.abs(w(omg - ang[2, :-1])).max() < 1e-9


def test_side_chain_torsion_moves_rigidly():
    """Changing xi1 of one residue moves exactly its side chain beyond CB, as
    a rotation about the CA-CB axis."""
    rng = np.random.default_rng(5)
    seq = "GKLW"
    ang = _angles(rng, 4)
    c0 = O.forward_one(seq, ang)[0].reshape(-1, 3)
    ang2 = ang.copy()
    ang2[3, 1] += 0.7  # xi1 of K
    c1 = O.forward_one(seq, ang2)[0].reshape(-1, 3)
    k0 = len(O.residue_atoms("G"))
    moved = np.where(np.abs(c1 - c0).max(1) > 1e-12)[0]
    assert list(moved) == list(range(k0 + 3, k0 + 7))  # CG, CD, CE, NZ of K
    ca, cb = c0[k0 + 1], c0[k0 + 2]
    d0 = np.linalg.norm(c0[moved] - cb, axis=1)
    assert np.allclose(np.linalg.norm(c1[moved] - cb, axis=1), d0)
    assert np.allclose(np.linalg.norm(c1[moved] - ca, axis=1), np.linalg.norm(c0[moved] - ca, axis=1))


def test_backbone_equals_angles2backbone():
    rng = np.random.default_rng(3)
    seqs = ["".join(rng.choice(list(O.AA20), size=n)) for n in (1, 2, 17, 50)]
    L = 50
    for seq in seqs:
        ang = _angles(rng, len(seq), L)
        c, _, _ = O.forward_one(seq, ang)
        N, CA, C = _backbone(seq, c)
        bb = A2B.forward(ang[None, :3, :], np.array([len(seq)], np.int32))[0][: 9 * len(seq)]
        bb = bb.reshape(-1, 3, 3)
        assert np.abs(bb[:, 0] - N).max() < 1e-12
        assert np.abs(bb[:, 1] - CA).max() < 1e-12
        assert np.abs(bb[:, 2] - C).max() < 1e-12


def test_gradient_finite_differences():
    rng = np.random.default_rng(8)
    seq = "MRKWHYFDE"
    ang = _angles(rng, len(seq), len(seq) + 1)
    c, _, _ = O.forward_one(seq, ang)
    g = rng.normal(size=c.size)
    ga = O.backward_one(seq, ang, g)
    h = 1e-6
    fd = np.zeros_like(ang)
    for r in range(8):
        for j in range(ang.shape[1]):
            ap, am = ang.copy(), ang.copy()
            ap[r, j] += h
            am[r, j] -= h
            fd[r, j] = (g @ O.forward_one(seq, ap)[0] - g @ O.forward_one(seq, am)[0]) / (2 * h)
    assert np.abs(ga - fd).max() < 1e-6 * max(1.0, np.abs(fd).max())
    # the angles nothing binds are exact zeros: omega of the last residue,
    # the padding column, unused side-chain rows (M binds xi1..xi4, H xi1..xi2)
    assert ga[2, len(seq) - 1] == 0 and not ga[:, len(seq)].any()
    assert ga[7, 0] == 0 and ga[5, 0] != 0
    assert not ga[5:, 4].any() and ga[4, 4] != 0


def test_reference_backward_test():
    """UnitTests/FullAtomModel/Angles2Coords/test.py:83-142: all 20 residues,
    phi -1.047, psi -0.698, omega pi, xi 110.4 degrees, L = sum(coords);
    forward differences dx = 1e-5 agree with the gradient to a mean of 0.01."""
    seq = "ACDEFGHIKLMNPQRSTVWY"
    ang = np.zeros((8, len(seq)))
    ang[0], ang[1], ang[2], ang[3:] = -1.047, -0.698, np.pi, 110.4 * np.pi / 180.0
    c, _, _ = O.forward_one(seq, ang)
    ga = O.backward_one(seq, ang, np.ones(c.size))
    y0 = c.sum()
    err = []
    for r in range(8):
        for j in range(len(seq)):
            a1 = ang.copy()
            a1[r, j] += 1e-5
            err.append(abs((O.forward_one(seq, a1)[0].sum() - y0) / 1e-5 - ga[r, j]))
    assert np.mean(err) < 0.01


def test_counts_and_names():
    assert O.num_atoms("") == 0
    assert O.num_atoms("G") == 4 and O.num_atoms("W") == 14 and O.num_atoms("AG") == 9
    res, atoms = O.names_one("GP")
    assert atoms == ["N", "CA", "C", "O", "N", "CA", "CB", "CG", "CD", "C", "O"]
    assert res == ["GLY"] * 4 + ["PRO"] * 7
    with pytest.raises(ValueError):
        O.build("AXG")


def test_fast_backward_equals_reference_form():
    """backward_fast (subtree sums) equals backward_one (the reference's
    per-angle sum of g . (F x) over the subtree) on every residue type."""
    rng = np.random.default_rng(21)
    for seq in (O.AA20, "".join(rng.choice(list(O.AA20), size=45)), "G", "PW"):
        ang = _angles(rng, len(seq), len(seq) + 2)
        c, _, _ = O.forward_one(seq, ang)
        g = rng.normal(size=c.size)
        a, b = O.backward_fast(seq, ang, g), O.backward_one(seq, ang, g)
        assert np.abs(a - b).max() <= 1e-11 * max(1.0, np.abs(b).max())


def test_matches_reference_fixtures():
    """Against outputs the reference's compiled code wrote
    (tests/golden/a2c_reference.npz, made by tests/golden/make_a2c_fixtures.py):
    the pin that holds where the reference sources are absent."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "a2c_reference.npz")
    z = np.load(path)
    k = 0
    while f"seq{k}" in z:
        seq = bytes(z[f"seq{k}"]).decode()
        ang = z[f"angles{k}"]
        c, _, _ = O.forward_one(seq, ang)
        assert np.abs(c - z[f"coords{k}"]).max() < 1e-12
        assert O.names_one(seq)[1] == list(z[f"atom_names{k}"])
        ga = O.backward_one(seq, ang, z[f"grad_coords{k}"])
        want = z[f"grad_angles{k}"]
        assert np.abs(ga - want).max() <= 1e-11 * max(1.0, np.abs(want).max())
        assert np.abs(O.backward_fast(seq, ang, z[f"grad_coords{k}"]) - want).max() <= 1e-11 * max(1.0, np.abs(want).max())
        k += 1
    assert k == 3
