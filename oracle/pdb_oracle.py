"""Python restatement of TorchProteinLibrary's PDB reader (PDB2CoordsUnordered).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's CPU-baseline leg
as the checker, never by the product path.

Restates Layers/FullAtomModel/cPDBLoader.cpp:23-53 (fixed columns of ATOM
records, heavy atoms only by Math/nUtil.cpp:166-171, coordinates parsed as
float by std::stof and widened) and the copy loop of
Layers/FullAtomModel/PDB2Coords/pdb2coords_interface.cpp:103-125.

The reference's loader is not built here: its header
(Layers/FullAtomModel/cPDBLoader.h:7) includes <TH/TH.h>, which the installed
torch no longer ships.  This module is pinned by the reference's own test
fixture (tests/golden/pdb2coords_reference_test.*) and by round trips.
"""
import numpy as np


def read(text):
    """Records of a PDB text: (coords [n, 3] float64, chain names, residue
    names, residue numbers, atom names)."""
    xyz, chains, res, nums, atoms = [], [], [], [], []
    for line in text.split("\n"):
        line = line.rstrip("\r")
        if line[0:4] != "ATOM":
            continue
        name = line[12:16].strip()
        if not name or name[0] not in "CNOS":
            continue
        xyz.append([float(np.float32(line[a:a + 8])) for a in (30, 38, 46)])
        res.append(line[17:20].strip())
        chains.append(line[21:22].strip())
        nums.append(int(line[22:26]))
        atoms.append(name)
    return np.array(xyz, dtype=np.float64).reshape(-1, 3), chains, res, nums, atoms


def read_file(filename):
    with open(filename) as f:
        return read(f.read())
