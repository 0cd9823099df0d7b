"""PDB files to coordinates (host side).

Drop-in for TorchProteinLibrary/FullAtomModel/PDB2Coords/PDB2Coords.py:
``PDB2CoordsUnordered()(filenames)`` returns ``(coords [B, 3A] float64,
chain_names [B, A, 4], res_names [B, A, 4], res_nums [B, A] int32,
atom_names [B, A, 4], num_atoms [B] int32)`` on the CPU, A the largest atom
count, zeros past each file's atoms (reference :76-89); ``writePDB`` writes
them back (:91-135); ``convertStringList``, ``string2tensor`` and
``tensor2string`` as there (:31-60).

The parsing is the native reader of include/tpl_pdb.h (records as the
reference's cPDBLoader reads them; files parsed in parallel).  Differences:
``pdb2sequence`` reads the sequence from the same ATOM records (the reference
uses Biopython, which this build does not depend on, and which also counts
HETATM residues); ``PDB2CoordsOrdered`` is not provided (the reference's
raises on use: it reads attributes it never sets, :63-74).
"""
from __future__ import annotations

import os

import numpy as np
import torch

from ..._native_xform import ext
from ..Angles2Coords.Angles2Coords import convertStringList

__all__ = ["PDB2CoordsUnordered", "writePDB", "pdb2sequence", "convertStringList", "string2tensor",
           "tensor2string"]

_ONE = {"GLY": "G", "ALA": "A", "SER": "S", "CYS": "C", "VAL": "V", "ILE": "I", "LEU": "L",
        "THR": "T", "ARG": "R", "LYS": "K", "ASP": "D", "ASN": "N", "GLU": "E", "GLN": "Q",
        "MET": "M", "HIS": "H", "PRO": "P", "PHE": "F", "TYR": "Y", "TRP": "W"}


def string2tensor(string):
    """A string -> 0-terminated uint8 tensor (reference :53-55)."""
    return torch.from_numpy(np.frombuffer(bytes(string + "\0", "ascii"), dtype=np.uint8).copy())


def tensor2string(tensor):
    """A 0-terminated uint8 tensor -> string (reference :58-60)."""
    return bytes(tensor.cpu().numpy().astype(np.uint8)).split(b"\x00")[0].decode("utf-8", "strict")


class PDB2CoordsUnordered:
    """Heavy-atom ATOM records of PDB files, in file order (reference :76-89)."""

    def __call__(self, filenames):
        names = convertStringList([os.fspath(f) for f in filenames])
        B = len(filenames)
        num_atoms = torch.zeros(B, dtype=torch.int32)
        coords = torch.zeros(B, 1, dtype=torch.double)
        chain_names = torch.zeros(B, 1, 1, dtype=torch.uint8)
        res_names = torch.zeros(B, 1, 1, dtype=torch.uint8)
        res_nums = torch.zeros(B, 1, dtype=torch.int32)
        atom_names = torch.zeros(B, 1, 1, dtype=torch.uint8)
        ext.PDB2CoordsUnordered(names, coords, chain_names, res_names, res_nums, atom_names, num_atoms)
        return coords, chain_names, res_names, res_nums, atom_names, num_atoms


def pdb2sequence(filenames):
    """One-letter sequences of PDB files from their ATOM records: a residue
    starts where the (chain, residue number) pair changes."""
    coords, chains, res, nums, _, n = PDB2CoordsUnordered()(filenames)
    out = []
    for b in range(len(filenames)):
        seq, last = [], None
        for j in range(int(n[b])):
            key = (tensor2string(chains[b, j]), int(nums[b, j]))
            if key != last:
                name = tensor2string(res[b, j])
                if name not in _ONE:
                    raise Exception("Can't extract sequence")
                seq.append(_ONE[name])
                last = key
        out.append("".join(seq))
    return out


def writePDB(filename, coords, chainnames, resnames, resnums, atomnames, num_atoms, bfactors=None,
             add_model=True, rewrite=True):
    """Append (or, with rewrite, write) the structures as PDB ATOM records,
    one MODEL per structure, numbering models and atoms after those already
    in the file when not rewriting (reference :91-135)."""
    last_model, last_atom = 0, 0
    if os.path.exists(filename):
        if rewrite:
            os.remove(filename)
        else:
            with open(filename) as fin:
                for line in fin:
                    if "MODEL" in line:
                        last_model = max(last_model, int(line.split()[1]))
                    if "ATOM" in line:
                        last_atom = max(last_atom, int(line[6:11]))
    coords = coords.detach().cpu()
    with open(filename, "a") as fout:
        for i in range(coords.size(0)):
            if add_model:
                fout.write("MODEL %d\n" % (i + last_model))
            for j in range(int(num_atoms[i])):
                chain = tensor2string(chainnames[i, j]) or "A"
                atom = tensor2string(atomnames[i, j])
                res = tensor2string(resnames[i, j])
                x, y, z = (float(coords[i, 3 * j + k]) for k in range(3))
                line = "ATOM  %5d  %-3s %3s %c%4d    %8.3f%8.3f%8.3f" % (
                    j + last_atom + 1, atom, res, chain[0], int(resnums[i, j]), x, y, z)
                if bfactors is not None:
                    line += "%6.2f%6.2f" % (1.0, float(bfactors[i, j]))
                fout.write(line + "\n")
            if add_model:
                fout.write("ENDMDL\n")
