"""Coordinate transforms on MI355X: translate, rotate and centre batches of
structures.

Drop-in for TorchProteinLibrary/FullAtomModel/CoordsTransform/CoordsTransform.py
(GPU path).  The names (``CoordsTranslate``, ``CoordsRotate``,
``Coords2Center`` and their ``...Function`` classes, ``getRotation``,
``getRandomRotation``, ``getBBox``, ``getRandomTranslation``), the calls, the
shapes and the error behaviour are the reference's:

* coordinates ``[B, 3*A]`` contiguous, float32 or float64, ``num_atoms`` int32
  ``[B]`` on the same device; ``T``/centre ``[B, 3]``, ``R`` ``[B, 3, 3]``;
* outputs are zero past ``num_atoms`` (the reference zero-initialises them);
* gradients: the translation and the rotation get none
  (reference :91, :149); the centre's is ``grad / num_atoms`` per atom;
* a wrong rank raises ``ValueError``; a NaN output raises
  ``Exception('<Function>: forward Nan')`` / ``backward Nan``
  (reference :67-68, 122-123, 146-147, 179-180, 203-204).

What differs, by design: one thread per atom slot of one structure
(neighbouring threads, neighbouring atoms) instead of one thread per structure
and one block per atom index, and the centre summed in FP64 by a workgroup per
structure; ``num_atoms`` outside ``[0, A]`` is clamped and reported; CPU
tensors are rejected (the GPU path only).  ``getRotation``,
``getRandomRotation``, ``getBBox`` and ``getRandomTranslation`` are small
host utilities in the reference (Math/nUtil.cpp:27-99) and stay host-side
torch code here, with torch's random generator as in the reference.
"""
from __future__ import annotations

import math
import os

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable
from torch.nn.modules.module import Module

from ..._native_xform import ext

__all__ = ["CoordsTranslate", "CoordsTranslateFunction", "CoordsRotate", "CoordsRotateFunction",
           "Coords2Center", "Coords2CenterFunction", "getRotation", "getRandomRotation",
           "getBBox", "getRandomTranslation"]

_STATUS_COUNT = 2


def _checks_enabled() -> bool:
    return os.environ.get("TPL_A2B_CHECKS", "1") != "0" and not torch.cuda.is_current_stream_capturing()


def _status_word(device):
    return torch.zeros(1, dtype=torch.int32, device=device) if _checks_enabled() else None


def _check(out, status, name, where):
    if status is not None:
        if int(status.item()) & _STATUS_COUNT:
            raise ValueError(f"{name}: num_atoms outside [0, coords.size(1)/3] (clamped on the device)")
        if math.isnan(out.sum()):
            raise Exception(f"{name}: {where} Nan")


def _rank2(t, name):
    if len(t.size()) != 2:
        raise ValueError(f"{name}: ", "Incorrect input size:", t.size())
    if not t.is_cuda:
        raise RuntimeError(f"{name}: this build runs on the GPU only; input must be a GPU tensor")


class CoordsTranslateFunction(Function):
    """Coordinates translation (reference CoordsTransform.py:46-91)."""

    @staticmethod
    def forward(ctx, input_coords, T, num_atoms):
        _rank2(input_coords, "CoordsTranslateFunction")
        ctx.save_for_backward(T, num_atoms)
        out = torch.empty_like(input_coords)
        status = _status_word(input_coords.device)
        ext.CoordsTranslateGPU_forward(input_coords, out, T.contiguous(), num_atoms, status)
        _check(out, status, "CoordsTranslateFunction", "forward")
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output_coords):
        T, num_atoms = ctx.saved_tensors
        grad_output_coords = grad_output_coords.contiguous()
        grad_input = torch.empty_like(grad_output_coords)
        ext.CoordsTranslateGPU_backward(grad_output_coords, grad_input, T.contiguous(), num_atoms, None)
        return grad_input, None, None


class CoordsTranslate(Module):
    def forward(self, input_coords, T, num_atoms):
        return CoordsTranslateFunction.apply(input_coords, T, num_atoms)


class CoordsRotateFunction(Function):
    """Coordinates rotation, R x per atom (reference CoordsTransform.py:101-149)."""

    @staticmethod
    def forward(ctx, input_coords, R, num_atoms):
        _rank2(input_coords, "CoordsRotateFunction")
        R = R.contiguous()
        ctx.save_for_backward(R, num_atoms)
        out = torch.empty_like(input_coords)
        status = _status_word(input_coords.device)
        ext.CoordsRotateGPU_forward(input_coords, out, R, num_atoms, status)
        _check(out, status, "CoordsRotateFunction", "forward")
        return out

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output_coords):
        R, num_atoms = ctx.saved_tensors
        grad_output_coords = grad_output_coords.contiguous()
        grad_input = torch.empty_like(grad_output_coords)
        status = _status_word(grad_output_coords.device)
        ext.CoordsRotateGPU_backward(grad_output_coords, grad_input, R, num_atoms, status)
        _check(grad_input, status, "CoordsRotateFunction", "backward")
        return grad_input, None, None


class CoordsRotate(Module):
    def forward(self, input_coords, R, num_atoms):
        return CoordsRotateFunction.apply(input_coords, R, num_atoms)


class Coords2CenterFunction(Function):
    """Geometric centre of each structure (reference CoordsTransform.py:158-206)."""

    @staticmethod
    def forward(ctx, input_coords, num_atoms):
        _rank2(input_coords, "Coords2CenterFunction")
        ctx.save_for_backward(input_coords, num_atoms)
        center = torch.empty(input_coords.size(0), 3, dtype=input_coords.dtype, device=input_coords.device)
        status = _status_word(input_coords.device)
        ext.Coords2CenterGPU_forward(input_coords, center, num_atoms, status)
        _check(center, status, "Coords2CenterFunction", "forward")
        return center

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output_center):
        input_coords, num_atoms = ctx.saved_tensors
        grad_output_center = grad_output_center.contiguous()
        grad_input = torch.empty_like(input_coords)
        status = _status_word(input_coords.device)
        ext.Coords2CenterGPU_backward(grad_output_center, grad_input, num_atoms, status)
        _check(grad_input, status, "Coords2CenterFunction", "backward")
        return grad_input, None


class Coords2Center(Module):
    def forward(self, input_coords, num_atoms):
        return Coords2CenterFunction.apply(input_coords, num_atoms)


def getRotation(u):
    """Rotation matrices [B, 3, 3] (float64, on u's device) from u [B, 3] in
    [0, 1)^3: the unit quaternion (sqrt(1-u1) sin 2 pi u2, sqrt(1-u1) cos 2 pi u2,
    sqrt(u1) sin 2 pi u3, sqrt(u1) cos 2 pi u3), uniform over rotations for
    uniform u (reference Math/nUtil.cpp:60-79)."""
    u = u.to(torch.float64)
    u1, u2, u3 = u[:, 0], u[:, 1], u[:, 2]
    q0 = torch.sqrt(1 - u1) * torch.sin(2 * math.pi * u2)
    q1 = torch.sqrt(1 - u1) * torch.cos(2 * math.pi * u2)
    q2 = torch.sqrt(u1) * torch.sin(2 * math.pi * u3)
    q3 = torch.sqrt(u1) * torch.cos(2 * math.pi * u3)
    R = torch.stack([
        q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2),
        2 * (q1 * q2 + q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * (q2 * q3 - q0 * q1),
        2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3,
    ], dim=1)
    return R.reshape(-1, 3, 3)


def getRandomRotation(batch_size, generator=None):
    """Uniformly random rotations [B, 3, 3] (float64, CPU, like the reference's
    CoordsTransform.py:35-38)."""
    return getRotation(torch.rand(batch_size, 3, dtype=torch.float64, generator=generator))


def getBBox(input_coords, num_atoms):
    """Bounding box of each structure's first num_atoms atoms: (a, b), the
    per-axis minimum and maximum, [B, 3] each, in the dtype of the
    coordinates, on the CPU (reference CoordsTransform.py:12-22 and
    Math/nUtil.cpp:27-57).  A structure with no atoms gets a = +inf,
    b = -inf, the reference's starting values."""
    if len(input_coords.size()) != 2:
        raise ValueError('getBBox: ', 'Incorrect input size:', input_coords.size())
    x = input_coords.detach().cpu()
    B = x.size(0)
    xyz = x.reshape(B, -1, 3)
    n = num_atoms.detach().cpu().to(torch.int64).clamp(0, xyz.size(1))
    valid = (torch.arange(xyz.size(1))[None, :] < n[:, None]).unsqueeze(-1)
    inf = torch.tensor(float("inf"), dtype=x.dtype)
    a = torch.where(valid, xyz, inf).amin(dim=1)
    b = torch.where(valid, xyz, -inf).amax(dim=1)
    return a, b


def getRandomTranslation(a, b, volume_size, generator=None):
    """Random translations [B, 3] that keep each box [a, b] inside a cubic
    volume of side volume_size: per axis uniform in [-d, d] with
    d = max(0, volume_size / 2 - (b - a) / 2) / 2, d in float32 as the
    reference computes it (reference CoordsTransform.py:24-33,
    Math/nUtil.cpp:88-99)."""
    if len(a.size()) != 2:
        raise ValueError('getRandomTranslation: ', 'Incorrect input size:', a.size())
    a = a.detach().cpu()
    b = b.detach().cpu()
    half = torch.clamp(float(volume_size) / 2.0 - (b.to(torch.float64) - a.to(torch.float64)) / 2.0, min=0.0)
    d = (half * 0.5).to(torch.float32).to(torch.float64)
    u = torch.rand(a.size(0), 3, dtype=torch.float64, generator=generator)
    return (u * (2.0 * d) - d).to(a.dtype)
