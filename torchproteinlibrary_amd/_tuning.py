"""Launch-shape autotuner cache for the Angles2Backbone scan kernels.

The reference has one fixed launch shape per kernel (cBackboneProteinCUDAKernels.cu:166,171-177).
Here a tile is 64*waves*res_per_lane residues.  For each (device, dtype,
batch, max_length) the first eager call times every candidate shape on the
caller's own angles and lengths (forward+backward, median of a few runs, in
the native tuner ``tpl_a2b_tune``) and caches the fastest.  Calls made while a
stream is being captured never tune: they use the cached shape or the
untuned default.

Environment: ``TPL_A2B_TUNE=0`` disables tuning (default shape always).
"""
from __future__ import annotations

import os
import threading

import torch

from ._native import ext

_lock = threading.Lock()
_cache: dict = {}
_timings: dict = {}

WAVES = (1, 2, 4)
RES_PER_LANE = (1, 2, 4, 8)


MAX_TILE = {torch.float32: 1024, torch.float64: 512}  # register-tile limit per dtype


def candidates(max_length: int, dtype=torch.float32, cus: int = 0, batch: int = 0):
    """Launch shapes worth timing for chains padded to max_length: every
    (waves, res_per_lane) whose tile is at most the next power of two >= L
    (larger tiles only add idle lanes) and within the dtype's tile limit, with
    one workgroup per tile; and, where one tile holds a whole chain, the same
    shape on a persistent forward grid of 2, 3 or 4 workgroups per CU, when
    that is fewer workgroups than chains (each then loads its next chain while
    it works on the current one)."""
    cap = 64
    while cap < max_length:
        cap *= 2
    cap = min(cap, MAX_TILE.get(dtype, 512))
    out = []
    for w in WAVES:
        for r in RES_PER_LANE:
            tile = 64 * w * r
            if tile <= max(cap, 64):
                out.append((w, r, 0))
                if cus > 0 and tile >= max_length:
                    out.extend((w, r, k * cus) for k in (2, 3, 4) if k * cus < batch)
    if not out:
        out.append((1, 1, 0))
    return out


def _key(angles: torch.Tensor):
    B, _, L = angles.shape
    return (angles.device.index, angles.dtype, int(B), int(L))


def set_config(angles_like: torch.Tensor, config) -> None:
    """Pin the launch shape used for inputs shaped like angles_like."""
    with _lock:
        _cache[_key(angles_like)] = tuple(int(c) for c in config)


def cached_config(angles: torch.Tensor):
    return _cache.get(_key(angles))


def timings(angles: torch.Tensor):
    """{config: ms} measured when the shape of angles was tuned (or None)."""
    return _timings.get(_key(angles))


def clear() -> None:
    with _lock:
        _cache.clear()
        _timings.clear()


def default_config(angles: torch.Tensor):
    B, _, L = angles.shape
    w, r, g = ext.default_config(int(B), int(L))
    while 64 * w * r > MAX_TILE.get(angles.dtype, 512):
        r //= 2
    return (w, max(r, 1), g)


def config_for(angles: torch.Tensor, lengths: torch.Tensor):
    key = _key(angles)
    cfg = _cache.get(key)
    if cfg is not None:
        return cfg
    if os.environ.get("TPL_A2B_TUNE", "1") == "0" or torch.cuda.is_current_stream_capturing():
        return default_config(angles)
    with _lock:
        cfg = _cache.get(key)
        if cfg is None:
            cfg = _tune(angles, lengths)
            _cache[key] = cfg
    return cfg


def tune_now(angles: torch.Tensor, lengths: torch.Tensor, reps: int = 5):
    """Tune for these inputs now (replacing any cached shape); returns it."""
    with _lock:
        cfg = _tune(angles, lengths, reps)
        _cache[_key(angles)] = cfg
    return cfg


def _tune(angles, lengths, reps: int = 5):
    B, _, L = angles.shape
    if B == 0 or L == 0:
        return default_config(angles)
    cus = torch.cuda.get_device_properties(angles.device).multi_processor_count
    cands = candidates(int(L), angles.dtype, cus, int(B))
    st = sy = 0
    for c in cands:
        s, y = ext.workspace_bytes(angles.dtype == torch.float64, int(B), int(L), list(c))
        st, sy = max(st, s), max(sy, y)
    dev, dt = angles.device, angles.dtype
    coords = torch.empty(B, 9 * L, dtype=dt, device=dev)
    gin = torch.empty(B, 3, L, dtype=dt, device=dev)
    gout = torch.zeros(B, 9 * L, dtype=dt, device=dev)
    state = torch.empty(max(st, 1), dtype=torch.uint8, device=dev)
    sync = torch.empty(max(sy, 1), dtype=torch.uint8, device=dev)
    best, ms = ext.tune(angles.detach().contiguous(), lengths.contiguous(), gout, coords, gin,
                        state, sync, [list(c) for c in cands], int(reps))
    _timings[_key(angles)] = {tuple(c): m for c, m in zip(cands, ms)}
    return tuple(best)
