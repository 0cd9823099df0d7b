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
  and the backward an O(L) reverse-scan adjoint that recomputes the frames
  from the input angles and one FP64 checkpoint frame per 64-256 residues
  saved by the forward (no ``A``/``dr_dangle`` buffers: the reference's ``A``
  is 192 bytes per residue and its ``dr_dangle`` O(B*L^2), 110 GB at B=1024,
  L=1000); the checkpoints are kept only when the input requires grad;
* gradients of psi[len-1] and omega[len-1] are exactly 0 (the reference reads
  frames past the chain there);
* lengths outside [0, L_max] are clamped on the device and reported (the
  reference writes out of bounds);
* the NaN / length checks read one status word written by the kernels (one
  host sync, like the reference's ``output.sum()``); they are skipped while a
  stream is being captured, and ``TPL_A2B_CHECKS=0`` turns them off;
* the look-back words of the kernels live in one zero-initialised buffer per
  (device, stream), whose tickets, counters and flags every call returns to
  zero;
* CPU tensors (float64, as the reference's CPU path requires) go to a host
  library: the same FP64 recurrence per chain and the O(L) adjoint from the
  forward's coordinates, chains spread over torch.get_num_threads() threads,
  where the reference builds an O(L^2) derivative tensor on one thread.
"""
from __future__ import annotations

import os
import threading
import warnings

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable
from torch.nn.modules.module import Module

from ... import _tuning
from ..._native import ext

_STATUS_NAN = 1
_STATUS_LENGTH = 2
_STATUS_TIMEOUT = 4
_STATUS_SYNC = 8


def _checks_enabled() -> bool:
    return os.environ.get("TPL_A2B_CHECKS", "1") != "0" and not torch.cuda.is_current_stream_capturing()


def _status_word(device):
    return torch.zeros(1, dtype=torch.int32, device=device) if _checks_enabled() else None


def _run_checked(launch, status, where: str, sync) -> None:
    """Runs launch(status) and raises on the status bits.  A look-back
    timeout or unclean look-back words make the result invalid: the words are
    cleared and the call is run once more, with a RuntimeWarning, and only a
    second failure raises.  The kernels' waits are bounded (a wave that never
    sees its record gives up and reports TIMEOUT instead of hanging), so this
    is the recovery path for a call that ran out of wait time."""
    launch(status)
    if status is None:
        return
    bits = int(status.item())
    if bits & (_STATUS_TIMEOUT | _STATUS_SYNC):
        ext.sync_clear(sync)  # look-back words may be left set
        what = "look-back wait timed out" if bits & _STATUS_TIMEOUT else "look-back words were not clean"
        warnings.warn("Angles2BackboneFunction: %s %s; cleared and repeated" % (where, what),
                      RuntimeWarning)
        status.zero_()
        launch(status)
        bits = int(status.item())
        if bits & (_STATUS_TIMEOUT | _STATUS_SYNC):
            ext.sync_clear(sync)
            what = "look-back wait timed out" if bits & _STATUS_TIMEOUT else "look-back words were not clean"
            raise RuntimeError("Angles2BackboneFunction: %s %s (result invalid)" % (where, what))
    _raise_on_bits(bits, where)


def _raise_on_status(status, where: str) -> None:
    if status is None:
        return
    bits = int(status.item())
    if bits & (_STATUS_TIMEOUT | _STATUS_SYNC):
        raise RuntimeError("Angles2BackboneFunction: %s look-back failed (result invalid)" % where)
    _raise_on_bits(bits, where)


def _raise_on_bits(bits: int, where: str) -> None:
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


_sync_lock = threading.Lock()
_sync_bufs: dict = {}


def _sync_buffer(device, nbytes: int):
    """The look-back buffer of (device, current stream), zeroed when made;
    every call returns its tickets, counters and flags to zero.  Grown (a new
    zeroed buffer) when too small.  Calls on one stream are ordered, so they
    can share it."""
    stream = torch.cuda.current_stream(device)
    key = (device.index, stream.cuda_stream)
    with _sync_lock:
        buf = _sync_bufs.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.zeros(max(int(nbytes), 256), dtype=torch.uint8, device=device)
            _sync_bufs[key] = buf
    return buf


class Angles2BackboneGPUFunction(Function):
    """Backbone angles -> backbone coordinates (chunk scan forward, O(L)
    adjoint backward from checkpoints).

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
        sync = _sync_buffer(dev, sy_bytes)
        status = _status_word(dev)
        _run_checked(lambda st: ext.Angles2BackboneGPU_forward(input_angles, coords, angles_length,
                                                               state, sync, st, list(cfg)),
                     status, "forward", sync)
        if state is not None:
            ctx.save_for_backward(input_angles, angles_length, state)
        ctx.cfg = cfg
        return coords

    @staticmethod
    @once_differentiable
    def backward(ctx, gradOutput):
        input_angles, angles_length, state = ctx.saved_tensors
        gradOutput = gradOutput.contiguous()
        if gradOutput.dtype != input_angles.dtype:
            gradOutput = gradOutput.to(input_angles.dtype)
        B, _, L = input_angles.shape
        dev = input_angles.device
        gradInput = torch.empty(B, 3, L, dtype=input_angles.dtype, device=dev)
        _, sy_bytes = ext.workspace_bytes(input_angles.dtype == torch.float64, B, L, list(ctx.cfg))
        sync = _sync_buffer(dev, sy_bytes)
        status = _status_word(dev)
        _run_checked(lambda st: ext.Angles2BackboneGPU_backward(gradInput, gradOutput, input_angles,
                                                                angles_length, state, sync, st,
                                                                list(ctx.cfg)),
                     status, "backward", sync)
        return gradInput, None


class Angles2BackboneCPUFunction(Function):
    """Backbone angles -> backbone coordinates on the host, float64.

    Reference: TorchProteinLibrary/ReducedModel/Angles2Backbone/Angles2Backbone.py:54-93
    (same names, float64 only, same error strings).
    """

    @staticmethod
    def forward(ctx, input, angles_length):
        _check_rank(input)
        B, _, L = input.shape
        coords = torch.empty(B, 9 * L, dtype=torch.double)
        bits = ext.Angles2BackboneCPU_forward(input, coords, angles_length, torch.get_num_threads())
        if bits & _STATUS_LENGTH:
            raise ValueError("Angles2BackboneFunction: angles_length outside [0, input.size(2)] "
                             "(clamped)")
        if bits & _STATUS_NAN:
            raise Exception('Angles2BackboneFunction: output_coords_cpu forward Nan')
        ctx.save_for_backward(coords, angles_length)
        return coords

    @staticmethod
    @once_differentiable
    def backward(ctx, gradOutput_cpu):
        coords, angles_length = ctx.saved_tensors
        gradOutput_cpu = gradOutput_cpu.contiguous().to(torch.double)
        B, L = coords.shape[0], coords.shape[1] // 9
        gradInput_cpu = torch.empty(B, 3, L, dtype=torch.double)
        bits = ext.Angles2BackboneCPU_backward(gradInput_cpu, gradOutput_cpu, coords, angles_length,
                                               torch.get_num_threads())
        if bits & _STATUS_NAN:
            raise Exception('Angles2BackboneFunction: gradInput_cpu backward Nan')
        return gradInput_cpu, None


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

    Reference: TorchProteinLibrary/ReducedModel/Angles2Backbone/Angles2Backbone.py:96-104:
    GPU tensors go to the chunk kernels, CPU tensors to the host library.
    ``form="plain"`` selects the GPU baseline kernels instead of the scan.
    """

    def __init__(self, form: str = "scan"):
        super(Angles2Backbone, self).__init__()
        if form not in ("scan", "plain"):
            raise ValueError("form must be 'scan' or 'plain'")
        self.form = form

    def forward(self, input, angles_length):
        if not input.is_cuda:
            if self.form == "plain":
                raise RuntimeError("Angles2Backbone: the plain form runs on the GPU only")
            return Angles2BackboneCPUFunction.apply(input, angles_length)
        if self.form == "plain":
            return Angles2BackbonePlainGPUFunction.apply(input, angles_length)
        return Angles2BackboneGPUFunction.apply(input, angles_length)


class Angles2BackboneStep:
    """Angles2Backbone forward and backward for fixed shapes, captured once
    into one HIP graph and replayed.

    For loops that run the same shapes many times (fitting, serving): one
    ``replay()`` runs the forward (``angles`` -> ``coords``) and the backward
    (``grad_coords`` -> ``grad_angles``) on the current stream, with nothing
    allocated, no host synchronisation and one launch per step.  The graph
    reads ``angles``, ``angles_length`` and ``grad_coords`` from their memory
    at every replay, so a loop updates them in place between replays; it
    writes ``coords`` and ``grad_angles``.  The chunk sizes are the tuned ones
    for these inputs, or ``config`` (forward, backward residues per lane).

    Replays do not check for NaN or out-of-range lengths (the checks of the
    layer read a status word on the host); lengths must stay within
    ``[0, angles.size(2)]``.  There is no counterpart in the reference, whose
    layer synchronises the host on every call (Angles2Backbone.py:27,49).
    """

    def __init__(self, angles, angles_length, config=None):
        _check_rank(angles)
        if not angles.is_cuda or not angles_length.is_cuda:
            raise RuntimeError("Angles2BackboneStep: angles and angles_length must be GPU tensors")
        if not angles.is_contiguous() or not angles_length.is_contiguous():
            raise RuntimeError("Angles2BackboneStep: angles and angles_length must be contiguous")
        if angles.dtype not in (torch.float32, torch.float64) or angles_length.dtype != torch.int32:
            raise RuntimeError("Angles2BackboneStep: angles must be float32 or float64, angles_length int32")
        B, _, L = angles.shape
        cfg = tuple(config) if config is not None else _tuning.config_for(angles, angles_length)
        dev, dt = angles.device, angles.dtype
        self.angles = angles
        self.angles_length = angles_length
        self.config = cfg
        self.coords = torch.zeros(B, 9 * L, dtype=dt, device=dev)
        self.grad_coords = torch.zeros(B, 9 * L, dtype=dt, device=dev)
        self.grad_angles = torch.zeros(B, 3, L, dtype=dt, device=dev)
        st_bytes, sy_bytes = ext.workspace_bytes(dt == torch.float64, B, L, list(cfg))
        self._state = _empty_bytes(max(int(st_bytes), 256), dev)
        self._sync = torch.zeros(max(int(sy_bytes), 256), dtype=torch.uint8, device=dev)
        self._graph = ext.StepGraph(angles, angles_length, self.grad_coords, self.coords,
                                    self.grad_angles, self._state, self._sync, list(cfg))

    def replay(self):
        """Runs one forward+backward step on the current stream."""
        self._graph.replay()
