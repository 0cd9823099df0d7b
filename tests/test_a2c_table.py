"""The residue templates of the HIP full-atom builder
(torchproteinlibrary_amd/csrc/a2c_table.h), compiled for the host, against
the oracle's restatement of the reference's templates (oracle/a2c_oracle.py,
itself pinned to the reference's compiled code in test_a2c_oracle.py)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import a2c_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("a2c") / "table")
    subprocess.check_call(["g++", "-O1", "-std=c++17", "-Wall", "-Werror", "-I",
                           os.path.join(ROOT, "torchproteinlibrary_amd", "csrc"),
                           os.path.join(ROOT, "tests", "a2c_table_host.cpp"), "-o", exe])
    out = subprocess.check_output([exe], text=True)
    assert "E type" not in out
    res, atoms, sides = {}, {}, {}
    for line in out.splitlines():
        f = line.split()
        if f[0] == "R":
            res[f[1]] = (f[2], int(f[3]), int(f[4]), int(f[5]))
        elif f[0] == "A":
            atoms.setdefault(f[1], []).append((f[3], np.array([float(v) for v in f[4:7]])))
        elif f[0] == "S":
            sides.setdefault(f[1], []).append((int(f[3]), float(f[4]), float(f[5]), float(f[6]),
                                               int(f[7]), int(f[8])))
    return res, atoms, sides


@pytest.mark.parametrize("aa", list(O.AA20))
def test_residue_template(dump, aa):
    res, atoms, sides = dump
    name3, natoms, nca, nside = res[aa]
    assert name3 == O.RES3[aa]
    assert natoms == len(O.residue_atoms(aa))
    ca_group = O.GROUPS["pro"] if aa == "P" else [("CA", O.Z)]
    assert nca == len(ca_group)
    assert nside == len(O.SIDE[aa])
    want = [("N", O.Z)] + ca_group + [a for node in O.SIDE[aa] for a in node[4]] + O.GROUPS["carbonyl"]
    got = atoms[aa]
    assert [n for n, _ in got] == [n for n, _ in want]
    for (n, v), (_, w) in zip(got, want):
        assert np.abs(v - w).max() < 1e-12, (aa, n, v, w)
    # bound rows are 3, 4, ... in node order (the kernels take them in turn)
    bound = [row for row, *_ in sides.get(aa, []) if row >= 0]
    assert bound == list(range(3, 3 + len(bound)))
    # only the first side-chain node is without atoms (the backward's parent rule)
    assert all(cnt > 0 for k, (*_, cnt) in enumerate(sides.get(aa, [])) if k > 0)
    first = 1 + nca
    for k, ((row, alpha, beta, d, f0, cnt), node) in enumerate(zip(sides.get(aa, []), O.SIDE[aa])):
        parent, al, b, dd, group = node
        assert parent == ("CA" if k == 0 else k - 1)  # a linear chain from CA
        if isinstance(al, int):
            assert row == al
        else:
            assert row == -1 and alpha == pytest.approx(al, abs=1e-15)
        assert beta == pytest.approx(b, abs=1e-15) and d == pytest.approx(dd, abs=1e-15)
        assert (f0, cnt) == (first, len(group))
        first += len(group)
