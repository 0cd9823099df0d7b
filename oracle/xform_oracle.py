"""FP64 CPU restatement of TorchProteinLibrary's coordinate transforms
(FullAtomModel/CoordsTransform), for tests only.

This module is test infrastructure: only tests/ (and __graft_entry__.smoke(),
bench.py's cpu_baseline leg) use it; the product path never imports it.

Per structure b, over its first n = num_atoms[b] atoms, zero past n:
* translate: x + T[b] (Layers/FullAtomModel/TransformCUDAKernels.cu:3-18);
  backward: the gradient passes through, T gets none (:21-37,
  CoordsTransform.py:91);
* rotate: R[b] x (mat33Vec3Mul, :40-53); backward: R[b]^T g (:56-75);
* centre: sum x / n (:77-86); backward: g[b] / n per atom (:88-98).
Pinned in tests/test_xform_oracle.py by identities (R^T R = I round trips,
translation round trips, the centre moving with the structure) and exact
finite differences of these linear maps.
"""
from __future__ import annotations

import numpy as np


def _mask(n, A):
    return (np.arange(A)[None, :] < np.asarray(n)[:, None])[..., None]


def translate(x, T, n):
    x = np.asarray(x, np.float64)
    B, A = x.shape[0], x.shape[1] // 3
    v = x.reshape(B, A, 3) + np.asarray(T, np.float64)[:, None, :]
    return np.where(_mask(n, A), v, 0.0).reshape(B, 3 * A)


def translate_backward(g, n):
    g = np.asarray(g, np.float64)
    B, A = g.shape[0], g.shape[1] // 3
    return np.where(_mask(n, A), g.reshape(B, A, 3), 0.0).reshape(B, 3 * A)


def rotate(x, R, n, transpose=False):
    x = np.asarray(x, np.float64)
    R = np.asarray(R, np.float64)
    B, A = x.shape[0], x.shape[1] // 3
    M = np.transpose(R, (0, 2, 1)) if transpose else R
    v = np.einsum("brc,bac->bar", M, x.reshape(B, A, 3))
    return np.where(_mask(n, A), v, 0.0).reshape(B, 3 * A)


def center(x, n):
    x = np.asarray(x, np.float64)
    B, A = x.shape[0], x.shape[1] // 3
    s = np.where(_mask(n, A), x.reshape(B, A, 3), 0.0).sum(1)
    with np.errstate(invalid="ignore", divide="ignore"):
        return s / np.asarray(n, np.float64)[:, None]


def center_backward(gc, n, A):
    gc = np.asarray(gc, np.float64)
    B = gc.shape[0]
    nn = np.asarray(n, np.float64)
    with np.errstate(invalid="ignore", divide="ignore"):
        v = np.broadcast_to((gc / np.where(nn > 0, nn, 1.0)[:, None])[:, None, :], (B, A, 3))
    return np.where(_mask(n, A), v, 0.0).reshape(B, 3 * A)
