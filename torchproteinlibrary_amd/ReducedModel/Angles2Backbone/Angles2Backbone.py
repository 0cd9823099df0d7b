"""Angles2Backbone: backbone dihedral angles -> N/CA/C coordinates, on MI355X.

Drop-in for TorchProteinLibrary/ReducedModel/Angles2Backbone/Angles2Backbone.py.
The names (``Angles2Backbone``, ``Angles2BackboneGPUFunction``), the call
``Angles2Backbone()(input, angles_length)``, the shapes, the output layout and
the error behaviour are the reference's:

* ``input``: ``[B, 3, L_max]`` contiguous, rows phi, psi, omega in radians,
  float32 (the reference GPU dtype) or float64; ``angles_length``: int32
  ``[B]`` on the same device (reference :12-30, angles2backbone_interface.cpp:12-15);
* returns ``[B, 9*L_max]`` coordinates N0, CA0, C0, N1, ... with exact zeros
  past each chain; gradient flows to ``input`` only (reference :52);
* device/contiguity/dtype violations raise ``RuntimeError`` (Math/nUtil.h:49-57),
  a wrong rank raises ``Exception`` (reference :22-23), and a NaN output raises
  ``Exception('Angles2BackboneFunction: forward Nan')`` / ``backward Nan``
  (reference :27-28, :49-50).

What differs, by design:

* the forward is a parallel FP64 prefix scan over per-residue rigid transforms
  and the backward an O(L) reverse-scan adjoint over the forward's output
  coordinates (no ``A``/``dr_dangle`` buffers: the reference's ``dr_dangle``
  is O(B*L^2), 110 GB at B=1024, L=1000); for float32 the forward also keeps
  2 bytes per coordinate of FP64 rounding residual for the backward, and only
  when the input requires grad;
* gradients of psi[len-1] and omega[len-1] are exactly 0 (the reference reads
  frames past the chain there);
* lengths outside [0, L_max] are clamped on the device and reported (the
  reference writes out of bounds);
* the NaN / length checks read one status word written by the kernels (one
  host sync, like the reference's ``output.sum()``); they are skipped while a
  stream is being captured, and ``TPL_A2B_CHECKS=0`` turns them off;
* CPU tensors are rejected: this build is the GPU path only.
"""
from __future__ import annotations

import os

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable
from torch.nn.modules.module import Module

from ... import _tuning
from ..._native import ext

_STATUS_NAN = 1
_STATUS_LENGTH = 2
_STATUS_TIMEOUT = 4


def _checks_enabled() -> bool:
    return os.environ.get("TPL_A2B_CHECKS", "1") != "0" and not torch.cuda.is_current_stream_capturing()


def _status_word(device):
    return torch.zeros(1, dtype=torch.int32, device=device) if _checks_enabled() else None


def _raise_on_status(status, where: str) -> None:
    if status is None:
        return
    bits = int(status.item())
    if bits & _STATUS_TIMEOUT:
        raise RuntimeError("Angles2BackboneFunction: %s look-back wait timed out "
                           "(result invalid)" % where)
    if bits & _STATUS_LENGTH:
        raise ValueError("Angles2BackboneFunction: angles_length outside [0, input.size(2)] "
                         "(clamped on the device)")
    if bits & _STATUS_NAN:
        raise Exception("Angles2BackboneFunction: %s Nan" % where)


def _check_rank(input_angles):
    if len(input_angles.size()) != 3:
        raise Exception("Angles2BackboneFunction: ", "Incorrect input size:", input_angles.size())


def _empty_bytes(n, device):
    return torch.empty(int(n), dtype=torch.uint8, device=device)


class Angles2BackboneGPUFunction(Function):
    """Backbone angles -> backbone coordinates (scan forward, O(L) adjoint backward).

    Reference: TorchProteinLibrary/ReducedModel/Angles2Backbone/Angles2Backbone.py:8-52.
    """

    @staticmethod
    def forward(ctx, input_angles, angles_length):
        _check_rank(input_angles)
        B, _, L = input_angles.shape
        cfg = _tuning.config_for(input_angles, angles_length)
        dev = input_angles.device
        coords = torch.empty(B, 9 * L, dtype=input_angles.dtype, device=dev)
        st_bytes, sy_bytes = ext.workspace_bytes(input_angles.dtype == torch.float64, B, L,
                                                 list(cfg))
        state = _empty_bytes(st_bytes, dev) if ctx.needs_input_grad[0] else None
        sync = _empty_bytes(sy_bytes, dev)
        status = _status_word(dev)
        ext.Angles2BackboneGPU_forward(input_angles, coords, angles_length, state, sync, status,
                                       list(cfg))
        _raise_on_status(status, "forward")
        if state is not None:
            ctx.save_for_backward(coords, angles_length, state)
        ctx.cfg = cfg
        return coords

    @staticmethod
    @once_differentiable
    def backward(ctx, gradOutput):
        coords, angles_length, state = ctx.saved_tensors
        gradOutput = gradOutput.contiguous()
        if gradOutput.dtype != coords.dtype:
            gradOutput = gradOutput.to(coords.dtype)
        B, L = coords.shape[0], coords.shape[1] // 9
        dev = coords.device
        gradInput = torch.empty(B, 3, L, dtype=coords.dtype, device=dev)
        status = _status_word(dev)
        ext.Angles2BackboneGPU_backward(gradInput, gradOutput, coords, angles_length, state,
                                        status, list(ctx.cfg))
        _raise_on_status(status, "backward")
        return gradInput, None


class Angles2BackbonePlainGPUFunction(Function):
    """The plain form: one thread per chain runs the serial recurrence.  The
    backward is the O(L^2) per-pair formula with one thread per angle.  It is
    the baseline the scan path is measured against (same API)."""

    @staticmethod
    def forward(ctx, input_angles, angles_length):
        _check_rank(input_angles)
        B, _, L = input_angles.shape
        dev = input_angles.device
        coords = torch.empty(B, 9 * L, dtype=input_angles.dtype, device=dev)
        frames = _empty_bytes(ext.plain_frames_bytes(input_angles.dtype == torch.float64, B, L), dev)
        status = _status_word(dev)
        ext.Angles2BackboneGPU_forward_plain(input_angles, coords, angles_length, frames, status)
        _raise_on_status(status, "forward")
        ctx.save_for_backward(input_angles, angles_length, frames)
        return coords

    @staticmethod
    @once_differentiable
    def backward(ctx, gradOutput):
        input_angles, angles_length, frames = ctx.saved_tensors
        gradOutput = gradOutput.contiguous().to(input_angles.dtype)
        B, _, L = input_angles.shape
        gradInput = torch.empty(B, 3, L, dtype=input_angles.dtype, device=input_angles.device)
        status = _status_word(input_angles.device)
        ext.Angles2BackboneGPU_backward_plain(gradInput, gradOutput, input_angles, angles_length,
                                              frames, status)
        _raise_on_status(status, "backward")
        return gradInput, None


class Angles2Backbone(Module):
    """``Angles2Backbone()(input [B,3,L], angles_length int32 [B]) -> [B, 9L]``.

    Reference: TorchProteinLibrary/ReducedModel/Angles2Backbone/Angles2Backbone.py:96-104.
    ``form="plain"`` selects the baseline kernels instead of the scan.
    """

    def __init__(self, form: str = "scan"):
        super(Angles2Backbone, self).__init__()
        if form not in ("scan", "plain"):
            raise ValueError("form must be 'scan' or 'plain'")
        self.form = form

    def forward(self, input, angles_length):
        if not input.is_cuda:
            raise RuntimeError("Angles2Backbone: this build runs on the GPU only; "
                               "input must be a GPU tensor")
        if self.form == "plain":
            return Angles2BackbonePlainGPUFunction.apply(input, angles_length)
        return Angles2BackboneGPUFunction.apply(input, angles_length)
