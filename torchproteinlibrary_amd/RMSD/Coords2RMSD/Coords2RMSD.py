"""Coords2RMSD: minimal RMSD over rigid rotations between two batches of
structures, on MI355X.

Drop-in for TorchProteinLibrary/RMSD/Coords2RMSD/Coords2RMSD.py.  The names
(``Coords2RMSD``, ``Coords2RMSDFunction``), the calls
``Coords2RMSD()(input, target, num_atoms)`` and
``Coords2RMSDFunction.apply(centered_input, centered_target, num_atoms)``, the
shapes and the error behaviour are the reference's:

* ``input``, ``target``: ``[B, 3*A]`` contiguous, xyz interleaved, float32 or
  float64; ``num_atoms``: int32 ``[B]`` on the same device;
* the Module returns ``rmsd [B]`` and keeps ``self.UT`` ``[B,3,3]`` (U^T, U the
  rotation taking the centred input onto the centred target); the Function
  returns ``(rmsd, UT)`` (reference :17-38, :71-78);
* the gradient flows to ``input`` only: ``(x_c - U^T y_c) * g / (rmsd * n)``,
  zero past ``num_atoms`` (reference :40-62);
* a wrong rank raises ``ValueError`` (reference :19-23); a NaN output raises
  ``Exception('Coords2RMSDFunction: forward Nan')`` / ``backward Nan``
  (reference :34-35, :59-60).

What differs, by design:

* centring, moments, the 4x4 eigenproblem, the RMSD and the gradient run in
  two device kernels with FP64 arithmetic, with no host round trip per
  structure (the reference solves each eigenproblem on the host and computes
  the gradient with a per-structure Python loop); the Module therefore needs
  no separate Coords2Center / CoordsTranslate launches, and the Function
  accepts raw as well as centred coordinates;
* the saved state is 16 doubles per structure, not centred copies of both
  inputs;
* a structure whose 4x4 matrix has no positive eigenvalue (fewer than two
  distinct atoms) gets U = I; the reference leaves U = 0 there.  With
  ``num_atoms == 0`` the RMSD is +inf as in the reference, and the gradient
  is 0 (the reference's is NaN);
* ``num_atoms`` outside ``[0, A]`` is clamped on the device and reported;
* CPU tensors are rejected: this build is the GPU path only.
"""
from __future__ import annotations

import os

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable
from torch.nn.modules.module import Module

from ..._native_rmsd import ext

_STATUS_NAN = 1
_STATUS_COUNT = 2


def _checks_enabled() -> bool:
    return os.environ.get("TPL_A2B_CHECKS", "1") != "0" and not torch.cuda.is_current_stream_capturing()


def _status_word(device):
    return torch.zeros(1, dtype=torch.int32, device=device) if _checks_enabled() else None


def _raise_on_status(status, where: str) -> None:
    if status is None:
        return
    bits = int(status.item())
    if bits & _STATUS_COUNT:
        raise ValueError("Coords2RMSDFunction: num_atoms outside [0, input.size(1)/3] "
                         "(clamped on the device)")
    if bits & _STATUS_NAN:
        raise Exception("Coords2RMSDFunction: %s Nan" % where)


class Coords2RMSDFunction(Function):
    """Protein coords, target coords -> (rmsd, UT).

    Reference: TorchProteinLibrary/RMSD/Coords2RMSD/Coords2RMSD.py:12-62.
    """

    @staticmethod
    def forward(ctx, centered_input, centered_target, num_atoms):
        if len(centered_input.size()) != 2:
            raise ValueError("Coords2RMSDFunction: ", "Incorrect input size:", centered_input.size())
        if not centered_input.is_cuda:
            raise RuntimeError("Coords2RMSD: this build runs on the GPU only; "
                               "input must be a GPU tensor")
        B = centered_input.size(0)
        dev, dt = centered_input.device, centered_input.dtype
        output = torch.empty(B, dtype=dt, device=dev)
        UT = torch.empty(B, 3, 3, dtype=dt, device=dev)
        state = (torch.empty(max(ext.state_bytes(B), 1), dtype=torch.uint8, device=dev)
                 if ctx.needs_input_grad[0] else None)
        status = _status_word(dev)
        ext.Coords2RMSDGPU_forward(centered_input, centered_target, output, num_atoms, UT, state,
                                   status)
        _raise_on_status(status, "forward")
        ctx.mark_non_differentiable(UT)
        if state is not None:
            ctx.save_for_backward(centered_input, centered_target, num_atoms, state)
        return output, UT

    @staticmethod
    @once_differentiable
    def backward(ctx, gradRMSD, gradUT):
        centered_input, centered_target, num_atoms, state = ctx.saved_tensors
        gradRMSD = gradRMSD.contiguous()
        if gradRMSD.dtype != centered_input.dtype:
            gradRMSD = gradRMSD.to(centered_input.dtype)
        grad_input = torch.empty_like(centered_input)
        status = _status_word(centered_input.device)
        ext.Coords2RMSDGPU_backward(grad_input, gradRMSD, centered_input, centered_target,
                                    num_atoms, state, status)
        _raise_on_status(status, "backward")
        return grad_input, None, None


class Coords2RMSD(Module):
    """``Coords2RMSD()(input [B,3A], target [B,3A], num_atoms int32 [B]) -> rmsd [B]``.

    Reference: TorchProteinLibrary/RMSD/Coords2RMSD/Coords2RMSD.py:64-78.
    """

    def __init__(self):
        super(Coords2RMSD, self).__init__()
        self.UT = None

    def forward(self, input, target, num_atoms):
        rmsd, self.UT = Coords2RMSDFunction.apply(input, target, num_atoms)
        return rmsd
