"""Writes tests/golden/a2c_reference.npz: outputs of the reference's own
full-atom code (compiled from its sources into oracle/_ref by
oracle/Makefile, target ref_a2c) for a few sequences, so that the oracle's
parity with it is checked where the reference sources are absent.

Run from the repository root: python tests/golden/make_a2c_fixtures.py"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import a2c_ref  # noqa: E402

SEQS = ["ACDEFGHIKLMNPQRSTVWY", "GPWYRKMHE", "S"]


def main():
    rng = np.random.default_rng(2024)
    out = {}
    for k, seq in enumerate(SEQS):
        L = len(seq)
        ang = np.zeros((8, L + 1))
        ang[0, :L] = rng.normal(-1.1, 0.5, L)
        ang[1, :L] = rng.normal(2.3, 0.5, L)
        ang[2, :L] = np.pi + rng.normal(0, 0.1, L)
        ang[3:, :L] = rng.uniform(-np.pi, np.pi, (5, L))
        c = a2c_ref.forward_one(seq, ang)
        g = rng.normal(size=c.size)
        out[f"seq{k}"] = np.frombuffer(seq.encode(), dtype=np.uint8)
        out[f"angles{k}"] = ang
        out[f"coords{k}"] = c
        out[f"grad_coords{k}"] = g
        out[f"grad_angles{k}"] = a2c_ref.backward_one(seq, ang, g)
        out[f"atom_names{k}"] = np.array(a2c_ref.names_one(seq)[0])
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "a2c_reference.npz"), **out)


if __name__ == "__main__":
    main()
