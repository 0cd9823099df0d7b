"""Writes tests/golden/a2b_golden.npz: input/output vectors of the reference's
own CPU kernels (Layers/ReducedModel/cBackboneProteinCPUKernels.cpp, FP64),
compiled from the reference sources by oracle/Makefile into oracle/_ref.

Cases (seeded): a poly-residue alpha helix (the angles of
UnitTests/ReducedModel/Angles2Backbone/test_forward_simple.py: phi=-1.047,
psi=-0.698, omega=0), all-zero angles, and a random batch with ragged lengths
(including 0, 1 and 2 residues) drawn like UnitTests/ReducedModel/Angles2Backbone/test.py
(randn angles).  For every case the fixture stores angles, lengths, an
upstream gradient and the reference's coords and grad_angles.  The reference
reads frame A[3n] for omega[n-1] (cBackboneProteinCPUKernels.cpp:83), past the
chain.  So every chain here is shorter than the stride and that slot is the
zero padding: the fixture then holds the true value 0 there.

Run: python tests/golden/make_golden.py   (needs oracle/_ref/libtplref.so)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def case_helix():
    L = 21
    a = np.zeros((1, 3, L))
    n = np.array([20], np.int32)
    a[0, 0, :20] = -1.047
    a[0, 1, :20] = -0.698
    a[0, 2, :20] = 0.0
    return a, n


def case_zero():
    L = 9
    return np.zeros((2, 3, L)), np.array([8, 5], np.int32)


def case_random():
    rng = np.random.default_rng(20240611)
    L = 17
    n = np.array([16, 0, 1, 2, 9, 13, 3, 16], np.int32)
    a = rng.standard_normal((len(n), 3, L))
    return a, n


def main():
    assert O.ref_available(), "build oracle/_ref first (make -C oracle)"
    out = {}
    rng = np.random.default_rng(7)
    for name, fn in [("helix", case_helix), ("zero", case_zero), ("random", case_random)]:
        a, n = fn()
        B, _, L = a.shape
        g = rng.standard_normal((B, 9 * L))
        for b in range(B):
            g[b, 9 * n[b]:] = 0.0
        x, ga = O.ref_forward_backward(a, n, g)
        out[f"{name}_angles"] = a
        out[f"{name}_lengths"] = n
        out[f"{name}_grad"] = g
        out[f"{name}_coords"] = x
        out[f"{name}_grad_angles"] = ga
    np.savez_compressed(os.path.join(HERE, "a2b_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "a2b_golden.npz"))


if __name__ == "__main__":
    main()
