"""Full-atom protein coordinates from angles on MI355X.

Drop-in for TorchProteinLibrary/FullAtomModel/Angles2Coords/Angles2Coords.py.
The names (``Angles2Coords``, ``Angles2CoordsFunction``,
``convertStringList``, ``convertString``), the call
``Angles2Coords()(input_angles, sequences)`` and what it returns are the
reference's (:34-87):

* ``input_angles`` ``[B, 8, L]`` (rows phi, psi, omega, xi1..xi5; residue i's
  N hangs on omega of residue i-1), float32 or float64, ``L`` at least the
  longest sequence; ``sequences`` a list of one-letter strings;
* returns ``(coords [B, 3*A], res_names [B, A, 4] uint8, atom_names
  [B, A, 4] uint8, num_atoms [B] int32)`` with ``A`` the largest atom count,
  coordinates and names 0 past each structure's atoms;
* the gradient flows to ``input_angles`` only; angles a structure does not
  use get 0;
* a NaN output raises ``Exception('Angles2CoordsFunction: forward Nan')`` /
  ``backward Nan`` (reference :53-54, 68-69).

What differs, by design: the work runs on the GPU (the angles must be a GPU
tensor; all four outputs are on its device), in FP64 inside whatever the
input dtype; a letter outside the 20 amino acids raises ``ValueError`` (the
reference silently skips the residue but still counts it); the backward is
O(atoms) per structure (the reference's is O(atoms x depth)) and works from
the saved output coordinates (bond axes and origins), so in float32 its
gradient carries the coordinates' rounding (about 1e-5 relative at 1000
residues).
"""
from __future__ import annotations

import math
import os

import numpy as np
import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable
from torch.nn.modules.module import Module

from ..._native_xform import ext

__all__ = ["Angles2Coords", "Angles2CoordsFunction", "convertStringList", "convertString"]

_STATUS_SEQ, _STATUS_LENGTH, _STATUS_ATOMS, _STATUS_ORDER = 1, 2, 4, 8
_LETTERS = set("ACDEFGHIKLMNPQRSTVWY")


def convertStringList(stringList):
    """List of strings -> 0-terminated uint8 tensor [len, longest + 1]
    (reference Angles2Coords.py:12-26)."""
    maxlen = max([len(s) + 1 for s in stringList] + [1])
    ar = np.zeros((len(stringList), maxlen), dtype=np.uint8)
    for i, s in enumerate(stringList):
        b = np.frombuffer(bytes(s, "ascii"), dtype=np.uint8)
        ar[i, :b.shape[0]] = b
    return torch.from_numpy(ar)


def convertString(string):
    """A string -> 0-terminated uint8 tensor (reference :28-30)."""
    return torch.from_numpy(np.frombuffer(bytes(string + "\0", "ascii"), dtype=np.uint8).copy())


def schedule(num_atoms):
    """The order in which the GPU takes the structures: the most atoms first
    (ties in batch order), as an int32 CPU tensor, or None when every
    structure has the same size (batch order is then as good).  One
    workgroup builds one structure and workgroups start in index order, so
    with mixed sizes this keeps the long structures from starting last:
    mixed lengths 30-1000 at B=1024 measured 59 -> 50 us (forward) and
    112 -> 77 us (backward)."""
    n = np.asarray(num_atoms, dtype=np.int64)
    if n.size < 2 or (n == n[0]).all():
        return None
    return torch.from_numpy(np.argsort(-n, kind="stable").astype(np.int32))


def _checks_enabled() -> bool:
    return os.environ.get("TPL_A2B_CHECKS", "1") != "0" and not torch.cuda.is_current_stream_capturing()


def _check(out, status, where):
    if status is not None:
        s = int(status.item())
        if s & _STATUS_SEQ:
            raise ValueError("Angles2CoordsFunction: a sequence holds a letter outside the 20 amino acids")
        if s & _STATUS_LENGTH:
            raise ValueError("Angles2CoordsFunction: a sequence is longer than input_angles.size(2)")
        if s & _STATUS_ATOMS:
            raise ValueError("Angles2CoordsFunction: a structure has more atoms than the output holds")
        if s & _STATUS_ORDER:
            raise RuntimeError("Angles2CoordsFunction: the scheduling order holds an index outside the batch")
        if math.isnan(out.sum()):
            raise Exception(f"Angles2CoordsFunction: {where} Nan")


class Angles2CoordsFunction(Function):
    """Protein angles -> coordinates (reference Angles2Coords.py:34-71).
    ``sequenceTensor`` is the uint8 [B, S] tensor of ``convertStringList`` on
    the angles' device.  ``order`` (optional, int32 [B]) is the order in
    which the GPU takes the structures (a permutation of 0..B-1; the outputs
    do not depend on it): ``Angles2Coords`` passes the structures from the
    most atoms to the fewest when their sizes differ."""

    @staticmethod
    def forward(ctx, input_angles, sequenceTensor, num_atoms, order=None):
        if input_angles.dim() != 3 or input_angles.size(1) != 8:
            raise ValueError("Angles2CoordsFunction: input_angles must be [batch, 8, length], got "
                             f"{tuple(input_angles.size())}")
        if not input_angles.is_cuda:
            raise RuntimeError("Angles2CoordsFunction: this build runs on the GPU only; "
                               "input_angles must be a GPU tensor")
        input_angles = input_angles.contiguous()
        ctx.angles_shape = input_angles.shape
        dev = input_angles.device
        seqs = sequenceTensor.to(dev).contiguous()
        B = input_angles.size(0)
        max_atoms = int(num_atoms.max()) if num_atoms.numel() else 0
        coords = torch.empty(B, 3 * max_atoms, dtype=input_angles.dtype, device=dev)
        res_names = torch.empty(B, max_atoms, 4, dtype=torch.uint8, device=dev)
        atom_names = torch.empty(B, max_atoms, 4, dtype=torch.uint8, device=dev)
        status = torch.zeros(1, dtype=torch.int32, device=dev) if _checks_enabled() else None
        if order is not None:
            order = order.to(dev).contiguous()
        ext.Angles2Coords_forward(seqs, input_angles, coords, res_names, atom_names, status, order)
        _check(coords, status, "forward")
        ctx.save_for_backward(seqs, coords)
        ctx.order = order
        ctx.mark_non_differentiable(res_names, atom_names, num_atoms)
        return coords, res_names, atom_names, num_atoms

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_atoms, *unused):
        seqs, coords = ctx.saved_tensors
        grad_atoms = grad_atoms.contiguous()
        grad_angles = torch.empty(ctx.angles_shape, dtype=coords.dtype, device=coords.device)
        status = torch.zeros(1, dtype=torch.int32, device=coords.device) if _checks_enabled() else None
        ext.Angles2Coords_backward_coords(grad_atoms, grad_angles, seqs, coords, status, ctx.order)
        _check(grad_angles, status, "backward")
        return grad_angles, None, None, None


class Angles2Coords(Module):
    """``Angles2Coords()(input_angles, sequences)`` (reference :73-87); keeps
    the last call's atom counts in ``self.num_atoms``."""

    def __init__(self):
        super().__init__()
        self.num_atoms = None

    def forward(self, input_angles, sequences):
        for s in sequences:
            if not set(s) <= _LETTERS:
                raise ValueError("Angles2Coords: a sequence holds a letter outside the 20 amino acids: "
                                 f"{sorted(set(s) - _LETTERS)}")
        self.num_atoms = [ext.getSeqNumAtoms(s) if s else 0 for s in sequences]
        num_atoms = torch.tensor(self.num_atoms, dtype=torch.int32, device=input_angles.device)
        stringListTensor = convertStringList(sequences)
        return Angles2CoordsFunction.apply(input_angles, stringListTensor, num_atoms, schedule(self.num_atoms))
