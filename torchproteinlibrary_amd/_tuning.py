"""Chunk-size autotuner cache for the Angles2Backbone kernels.

The reference has one fixed launch shape per kernel (cBackboneProteinCUDAKernels.cu:166,171-177).
Here a forward chunk is 64*fwd_res_per_lane residues and a backward chunk
(and checkpoint stride) 64*bwd_res_per_lane; one wave takes one chunk.  The
best sizes depend on the device and on the mix of chain lengths, so the cache
key is (device name, dtype, batch, max_length, length bucket), the bucket
being the fraction of chains in each power-of-two length band, in eighths.

The bucket needs the lengths on the host.  It is computed once per lengths
tensor (its storage pointer and version), so a loop that reuses its lengths
tensor reads them once; calls made while a stream is being captured never
read them and never tune: they use the shape's last bucket and cached config,
or the untuned default.  The first eager call of a key times every candidate
on the caller's own angles and lengths and caches the fastest: the native
tuner ``tpl_a2b_tune`` runs each candidate's forward+backward steps back to
back, as a training loop does, with the candidates taking turns over a few
passes, and keeps each one's median.

Environment: ``TPL_A2B_TUNE=0`` disables tuning (default config always).
"""
from __future__ import annotations

import os
import threading

import torch

from ._native import ext

_lock = threading.Lock()
_cache: dict = {}
_timings: dict = {}
_bucket_memo: dict = {}
_last_bucket: dict = {}

FWD_RES_PER_LANE = (1, 2, 4, 8)
BWD_RES_PER_LANE = (1, 2, 4)
MAX_RES = {torch.float32: (8, 4), torch.float64: (4, 2)}  # (forward, backward) register limits


def candidates(max_length: int, dtype=torch.float32):
    """(fwd_res_per_lane, bwd_res_per_lane) pairs worth timing for chains
    padded to max_length: chunks of at most the next power of two >= L
    (larger chunks only add idle lanes) within the dtype's register limits."""
    cap = 64
    while cap < max_length:
        cap *= 2
    fmax, bmax = MAX_RES.get(dtype, (4, 2))
    out = [(f, b) for f in FWD_RES_PER_LANE for b in BWD_RES_PER_LANE
           if f <= fmax and b <= bmax and 64 * f <= cap and 64 * b <= cap]
    return out or [(1, 1)]


def length_bucket(lengths_host, max_length: int):
    """Fractions (in eighths, rounded) of chains in the length bands
    [0, 64), [64, 128), [128, 256), ... up to max_length."""
    import numpy as np
    n = np.asarray(lengths_host).reshape(-1)
    if n.size == 0:
        return ()
    edges = [0, 64]
    while edges[-1] < max_length:
        edges.append(edges[-1] * 2)
    edges.append(1 << 62)
    h, _ = np.histogram(np.clip(n, 0, None), bins=edges)
    return tuple(int(round(8.0 * c / n.size)) for c in h)


def _shape_key(angles: torch.Tensor):
    B, _, L = angles.shape
    dev = angles.device
    name = torch.cuda.get_device_name(dev) if dev.type == "cuda" else "cpu"
    return (name, dev.index, angles.dtype, int(B), int(L))


def _bucket(angles: torch.Tensor, lengths: torch.Tensor, may_read: bool):
    shape = _shape_key(angles)
    memo = (lengths.data_ptr(), lengths._version, lengths.numel())
    hit = _bucket_memo.get(memo)
    if hit is not None and hit[0] == shape:
        return hit[1]
    if not may_read:
        return _last_bucket.get(shape, ())
    b = length_bucket(lengths.detach().cpu().numpy(), int(angles.shape[2]))
    _bucket_memo[memo] = (shape, b)
    if len(_bucket_memo) > 4096:
        _bucket_memo.clear()
    _last_bucket[shape] = b
    return b


def set_config(angles_like: torch.Tensor, config, lengths: torch.Tensor | None = None) -> None:
    """Pin the config used for inputs shaped like angles_like (for every
    length mix, or for the mix of `lengths` only)."""
    with _lock:
        shape = _shape_key(angles_like)
        cfg = tuple(int(c) for c in config)
        if lengths is None:
            _cache[shape + ("any",)] = cfg
        else:
            _cache[shape + (_bucket(angles_like, lengths, True),)] = cfg


def cached_config(angles: torch.Tensor, lengths: torch.Tensor | None = None):
    shape = _shape_key(angles)
    if lengths is not None:
        c = _cache.get(shape + (_bucket(angles, lengths, False),))
        if c is not None:
            return c
    return _cache.get(shape + ("any",))


def timings(angles: torch.Tensor, lengths: torch.Tensor):
    """{config: ms} measured when this shape and length mix were tuned (or None)."""
    return _timings.get(_shape_key(angles) + (_bucket(angles, lengths, False),))


def clear() -> None:
    with _lock:
        _cache.clear()
        _timings.clear()
        _bucket_memo.clear()
        _last_bucket.clear()


def default_config(angles: torch.Tensor):
    B, _, L = angles.shape
    f, b = ext.default_config(int(B), int(L))
    fmax, bmax = MAX_RES.get(angles.dtype, (4, 2))
    return (min(f, fmax), min(b, bmax))


def config_for(angles: torch.Tensor, lengths: torch.Tensor):
    capturing = torch.cuda.is_current_stream_capturing()
    shape = _shape_key(angles)
    pinned = _cache.get(shape + ("any",))
    if pinned is not None:
        return pinned
    tune = os.environ.get("TPL_A2B_TUNE", "1") != "0" and not capturing
    key = shape + (_bucket(angles, lengths, tune),)
    cfg = _cache.get(key)
    if cfg is not None:
        return cfg
    if not tune:
        return default_config(angles)
    with _lock:
        cfg = _cache.get(key)
        if cfg is None:
            cfg = _tune(angles, lengths, key)
            _cache[key] = cfg
    return cfg


def tune_now(angles: torch.Tensor, lengths: torch.Tensor, reps: int = 5):
    """Tune for these inputs now (replacing any cached config); returns it."""
    key = _shape_key(angles) + (_bucket(angles, lengths, True),)
    with _lock:
        cfg = _tune(angles, lengths, key, reps)
        _cache[key] = cfg
    return cfg


def _tune(angles, lengths, key, reps: int = 5):
    B, _, L = angles.shape
    if B == 0 or L == 0:
        return default_config(angles)
    cands = candidates(int(L), angles.dtype)
    st = sy = 0
    for c in cands:
        s, y = ext.workspace_bytes(angles.dtype == torch.float64, int(B), int(L), list(c))
        st, sy = max(st, s), max(sy, y)
    dev, dt = angles.device, angles.dtype
    coords = torch.empty(B, 9 * L, dtype=dt, device=dev)
    gin = torch.empty(B, 3, L, dtype=dt, device=dev)
    gout = torch.zeros(B, 9 * L, dtype=dt, device=dev)
    state = torch.empty(max(st, 256), dtype=torch.uint8, device=dev)
    sync = torch.zeros(max(sy, 256), dtype=torch.uint8, device=dev)
    best, ms = ext.tune(angles.detach().contiguous(), lengths.contiguous(), gout, coords, gin,
                        state, sync, [list(c) for c in cands], int(reps))
    _timings[key] = {tuple(c): m for c, m in zip(cands, ms)}
    return tuple(best)
