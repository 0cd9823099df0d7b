"""FullAtomModel.PDB2Coords (host code): the native reader
(include/tpl_pdb.h) through its Python layer and its C ABI, against the
reference's own test fixture and the Python restatement in
oracle/pdb_oracle.py (itself pinned by that fixture)."""
import ctypes
import json
import os
import re

import numpy as np
import pytest
import torch

from oracle import pdb_oracle as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FIX = os.path.join(ROOT, "tests", "golden", "pdb2coords_reference_test.pdb")
EXPECT = json.load(open(os.path.join(ROOT, "tests", "golden", "pdb2coords_reference_test.json")))
LIB = os.path.join(ROOT, "torchproteinlibrary_amd", "libtpl_pdb.so")
HEADER = os.path.join(ROOT, "include", "tpl_pdb.h")


def _s(t):
    return bytes(t.numpy().astype(np.uint8)).split(b"\x00")[0].decode()


def _check_expectations(n, chains, res, atoms, xyz):
    assert n == EXPECT["num_atoms"]
    assert all(c == EXPECT["chain_names"] for c in chains[:n])
    for name, (a, b) in EXPECT["res_names"].items():
        assert all(r == name for r in res[a:b])
    assert set(atoms[:n]) == set(EXPECT["atom_name_set"])
    for j, v in EXPECT["coords"].items():
        assert np.round(np.abs(np.array(xyz[int(j)]) - v), EXPECT["places"]).max() == 0


def test_oracle_on_reference_fixture():
    xyz, chains, res, nums, atoms = P.read_file(FIX)
    _check_expectations(len(atoms), chains, res, atoms, xyz)


def test_layer_on_reference_fixture():
    """The reference's own test (UnitTests/FullAtomModel/PDB2Coords/test.py:62-93)."""
    from torchproteinlibrary_amd.FullAtomModel import PDB2CoordsUnordered
    coords, chains, res, nums, atoms, n = PDB2CoordsUnordered()([FIX])
    k = int(n[0])
    _check_expectations(k, [_s(chains[0, i]) for i in range(k)], [_s(res[0, i]) for i in range(k)],
                        [_s(atoms[0, i]) for i in range(k)], coords[0].reshape(-1, 3).numpy())


EDGE = """HEADER    SYNTHETIC
MODEL        1
ATOM      1  N   MET A  -2      11.104  13.207   2.100  1.00 20.00           N
ATOM      2  CA  MET A  -2      12.560  13.207   2.100  1.00 20.00           C
ATOM      3  H   MET A  -2      10.500  12.900   2.900  1.00 20.00           H
ATOM      4 1HB  MET A  -2      10.000  12.000   1.000  1.00 20.00           H
ATOM      5  SD  MET A  -2     -13.125  -1.500 105.250  1.00 20.00           S
HETATM    6  O   HOH A 101       1.000   2.000   3.000  1.00 20.00           O
TER       7      MET A  -2
ATOM      8  OXT GLY B 999       0.125  -0.250   0.375\r
ATOM      9  CD1 LEU B1000    1234.567-999.999   0.001
ATOM     10  C1' DA  C   5       1.1     2.2     3.3
ENDMDL
END
ATOM     11  NZ  LYS D   7      -1.000  -2.000  -3.000"""


def test_layer_matches_oracle_on_edge_records(tmp_path):
    from torchproteinlibrary_amd.FullAtomModel import PDB2CoordsUnordered
    f1, f2 = tmp_path / "edge.pdb", tmp_path / "ref.pdb"
    f1.write_text(EDGE)
    f2.write_text(open(FIX).read())
    coords, chains, res, nums, atoms, n = PDB2CoordsUnordered()([str(f1), str(f2)])
    for b, f in enumerate((f1, f2)):
        xyz, ch, rn, nm, an = P.read_file(str(f))
        k = int(n[b])
        assert k == len(an)
        assert np.array_equal(coords[b, :3 * k].numpy(), xyz.reshape(-1))  # bit-exact
        assert not coords[b, 3 * k:].any() and not atoms[b, k:].any()
        assert [_s(chains[b, i]) for i in range(k)] == ch
        assert [_s(res[b, i]) for i in range(k)] == rn
        assert nums[b, :k].tolist() == nm
        assert [_s(atoms[b, i]) for i in range(k)] == an
    # hydrogens, HETATM and non-ATOM lines skipped; CR stripped; name "C1'" kept
    assert [_s(atoms[0, i]) for i in range(int(n[0]))] == ["N", "CA", "SD", "OXT", "CD1", "C1'", "NZ"]


def test_write_read_round_trip(tmp_path):
    from torchproteinlibrary_amd.FullAtomModel import PDB2CoordsUnordered, pdb2sequence, writePDB
    p2c = PDB2CoordsUnordered()
    coords, chains, res, nums, atoms, n = p2c([FIX, FIX])
    coords[1] *= 2.0  # a second, different model
    out = str(tmp_path / "rt.pdb")
    writePDB(out, coords, chains, res, nums, atoms, n, add_model=False)
    c2, ch2, r2, n2, a2, k2 = p2c([out])
    assert int(k2[0]) == 2 * int(n[0])
    k = int(n[0])
    assert np.abs(c2[0].reshape(-1, 3)[:k].numpy() - coords[0].reshape(-1, 3)[:k].numpy()).max() < 5e-4
    assert np.abs(c2[0].reshape(-1, 3)[k:].numpy() - coords[1].reshape(-1, 3)[:k].numpy()).max() < 5e-4
    assert torch.equal(a2[0, :k], atoms[0, :k]) and torch.equal(n2[0, :k], nums[0, :k])
    assert pdb2sequence([FIX]) == ["QTAA"]


def test_errors(tmp_path):
    from torchproteinlibrary_amd.FullAtomModel import PDB2CoordsUnordered
    with pytest.raises(RuntimeError, match="cannot open"):
        PDB2CoordsUnordered()([str(tmp_path / "missing.pdb")])
    bad = tmp_path / "bad.pdb"
    bad.write_text("ATOM      1  CA  ALA A   1      1.000\n")
    with pytest.raises(RuntimeError, match="too short or unparsable"):
        PDB2CoordsUnordered()([str(bad)])


def test_c_abi():
    names = sorted(set(re.findall(r"\b(tpl_pdb_\w+)\s*\(", open(HEADER).read())))
    assert names == ["tpl_pdb_count", "tpl_pdb_read", "tpl_pdb_strerror"]
    lib = ctypes.CDLL(LIB)
    for name in names:
        assert hasattr(lib, name)
    lib.tpl_pdb_count.argtypes = [ctypes.c_char_p]
    vp = ctypes.c_void_p
    lib.tpl_pdb_read.argtypes = [ctypes.c_char_p, vp, vp, vp, vp, vp, ctypes.c_int]
    n = lib.tpl_pdb_count(FIX.encode())
    assert n == 26
    c = np.zeros(3 * n)
    ch, rn, an = (np.zeros((n, 4), np.uint8) for _ in range(3))
    nums = np.zeros(n, np.int32)
    assert lib.tpl_pdb_read(FIX.encode(), c.ctypes.data, ch.ctypes.data, rn.ctypes.data, nums.ctypes.data,
                            an.ctypes.data, n) == n
    assert np.array_equal(c, P.read_file(FIX)[0].reshape(-1))
    assert lib.tpl_pdb_read(FIX.encode(), c.ctypes.data, ch.ctypes.data, rn.ctypes.data, nums.ctypes.data,
                            an.ctypes.data, 5) == 5  # stops at max_atoms
    assert lib.tpl_pdb_count(None) == -3 and lib.tpl_pdb_count(b"/nonexistent/x.pdb") == -1
