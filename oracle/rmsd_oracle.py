"""FP64 CPU restatement of TorchProteinLibrary's Coords2RMSD, for tests only.

This module is test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use it.  The product path (torchproteinlibrary_amd)
never imports it.

It follows the reference step by step, in numpy FP64:
* centring both structures over their first n atoms (Coords2Center +
  CoordsTranslate, FullAtomModel/CoordsTransform/CoordsTransform.py:158-213,
  46-98; RMSD/Coords2RMSD/Coords2RMSD.py:71-77);
* the correlation matrix R[i][j] = sum_atoms xc_i yc_j
  (Layers/RMSD/Coords2RMSD/coords2rmsd_interface.cpp:135-150);
* the symmetric 4x4 matrix F built from R (:153-158), its largest eigenvalue
  and eigenvector q (:12-32; the reference uses torch's general eig, this
  uses numpy.linalg.eigh, which is exact for a symmetric matrix);
* U from q (:35-45) and UT[j][k] = U[k][j] (:90-93);
* R2 = sum |xc|^2 + |yc|^2 (:169-180) and
  rmsd = sqrt((R2 - 2 |lambda| + 1e-10) / n) (:182), with R2 - 2 |lambda|
  clamped at 0 (rounding can make it negative for identical structures,
  where the reference's formula gives NaN);
* the gradient (xc - UT yc) * g / (rmsd * n), zero past n
  (Coords2RMSD.py:47-56; CoordsRotate is dst = R src,
  Layers/FullAtomModel/TransformCUDAKernels.cu:41-50).

What pins it (tests/test_rmsd_oracle.py): an independent Kabsch/SVD solution
of the same least-squares problem (same RMSD and rotation), invariance under
rigid motions of either structure, zero RMSD for a rigidly moved copy (the
reference's own test, UnitTests/RMSD/Coords2RMSD/test.py:30-49), and central
finite differences of the forward against the backward.

Divergences shared with the HIP path: a structure whose F has no positive
eigenvalue (fewer than two distinct atoms) gets U = I, where the reference's
selection loop leaves U = 0 (coords2rmsd_interface.cpp:16-32); with n = 0 the
gradient is 0 where the reference's is NaN.
"""
from __future__ import annotations

import numpy as np

EPS_LIMIT = 1e-10  # coords2rmsd_interface.cpp:10


def _split(c, n):
    return np.asarray(c, dtype=np.float64)[: 3 * n].reshape(n, 3)


def f_matrix(R):
    """The reference's 4x4 matrix of the correlation R (coords2rmsd_interface.cpp:155-158)."""
    return np.array([
        [R[0, 0] + R[1, 1] + R[2, 2], R[1, 2] - R[2, 1], R[2, 0] - R[0, 2], R[0, 1] - R[1, 0]],
        [R[1, 2] - R[2, 1], R[0, 0] - R[1, 1] - R[2, 2], R[0, 1] + R[1, 0], R[0, 2] + R[2, 0]],
        [R[2, 0] - R[0, 2], R[0, 1] + R[1, 0], -R[0, 0] + R[1, 1] - R[2, 2], R[1, 2] + R[2, 1]],
        [R[0, 1] - R[1, 0], R[0, 2] + R[2, 0], R[1, 2] + R[2, 1], -R[0, 0] - R[1, 1] + R[2, 2]],
    ])


def u_from_q(q):
    """Rotation matrix of the unit quaternion q (coords2rmsd_interface.cpp:35-45)."""
    q0, q1, q2, q3 = q
    return np.array([
        [q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)],
        [2 * (q1 * q2 + q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * (q2 * q3 - q0 * q1)],
        [2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3],
    ])


def forward_one(x, y, n):
    """(rmsd, UT, U, cx, cy) of one structure pair over its first n atoms."""
    if n == 0:
        return np.inf, np.eye(3), np.eye(3), np.zeros(3), np.zeros(3)
    xs, ys = _split(x, n), _split(y, n)
    cx, cy = xs.mean(0), ys.mean(0)
    xc, yc = xs - cx, ys - cy
    R = xc.T @ yc
    w, V = np.linalg.eigh(f_matrix(R))
    k = int(np.argmax(w))
    lam = w[k]
    U = u_from_q(V[:, k]) if lam > 0 else np.eye(3)
    r2 = float((xc * xc).sum() + (yc * yc).sum())
    # clamped at 0 (FP64 rounding can make r2 - 2|lambda| negative for
    # identical structures; the reference's formula then gives NaN)
    rmsd = np.sqrt((max(r2 - 2.0 * abs(lam), 0.0) + EPS_LIMIT) / n)
    return rmsd, U.T.copy(), U, cx, cy


def forward(x, y, num_atoms):
    """Batched forward: x, y [B, 3A]; returns rmsd [B], UT [B, 3, 3]."""
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    B = x.shape[0]
    rmsd = np.zeros(B)
    UT = np.zeros((B, 3, 3))
    for b in range(B):
        rmsd[b], UT[b] = forward_one(x[b], y[b], int(num_atoms[b]))[:2]
    return rmsd, UT


def backward(x, y, num_atoms, grad_rmsd):
    """Gradient of sum_b grad_rmsd[b] * rmsd[b] with respect to x."""
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    out = np.zeros_like(x)
    for b in range(x.shape[0]):
        n = int(num_atoms[b])
        if n == 0:
            continue
        rmsd, UT, _, cx, cy = forward_one(x[b], y[b], n)
        xc, yc = _split(x[b], n) - cx, _split(y[b], n) - cy
        g = (xc - yc @ UT.T) * (float(grad_rmsd[b]) / (rmsd * n))
        out[b, : 3 * n] = g.reshape(-1)
    return out


def kabsch(x, y, n):
    """Independent check: least-squares rotation by SVD (Kabsch).  Returns
    (rmsd without the 1e-10 term, rotation taking centred x onto centred y)."""
    xs, ys = _split(x, n), _split(y, n)
    xc, yc = xs - xs.mean(0), ys - ys.mean(0)
    H = xc.T @ yc
    Uh, S, Vt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(Vt.T @ Uh.T))
    D = np.diag([1.0, 1.0, d])
    Rot = Vt.T @ D @ Uh.T
    diff = xc @ Rot.T - yc
    return float(np.sqrt((diff * diff).sum() / n)), Rot


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    return u_from_q(q)


def workload(B, A, seed, min_atoms=1, noise=0.5, dtype=np.float64):
    """Seeded structures: x [B, 3A] random walks (3.8 A steps, as a C-alpha
    trace), y = a random rigid motion of x plus Gaussian noise; num_atoms in
    [min_atoms, A]; slots past num_atoms are 0."""
    rng = np.random.default_rng(seed)
    n = rng.integers(min_atoms, A + 1, size=B).astype(np.int32)
    x = np.zeros((B, 3 * A))
    y = np.zeros((B, 3 * A))
    for b in range(B):
        steps = rng.normal(size=(A, 3))
        steps *= 3.8 / np.linalg.norm(steps, axis=1, keepdims=True)
        xs = np.cumsum(steps, axis=0) + rng.normal(scale=20.0, size=3)
        ys = xs @ random_rotation(rng).T + rng.normal(scale=20.0, size=3)
        ys += rng.normal(scale=noise, size=ys.shape)
        k = int(n[b])
        x[b, : 3 * k] = xs[:k].reshape(-1)
        y[b, : 3 * k] = ys[:k].reshape(-1)
    g = rng.normal(size=B)
    return x.astype(dtype), y.astype(dtype), n, g.astype(dtype)
