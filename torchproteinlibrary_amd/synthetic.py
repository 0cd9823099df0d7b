"""Seeded synthetic Angles2Backbone workloads (angles, lengths, upstream grad).

Five named workloads; each draws everything in float64 with
``numpy.random.default_rng(seed)`` and casts the angles and gradients to
float32, so a CPU oracle fed ``angles.astype(float64)`` sees exactly the bits
the GPU sees.

=========  =====  ===================================  =======
name       B      lengths                              L_max
=========  =====  ===================================  =======
uniform200 32     all 200                              200
uniform700 256    all 700                              700
mixed1000  1024   floor(exp(U(ln 30, ln 1001))), <=1000 1000
long16k    64     uniform integer [4096, 16384]        16384
fit500     512    uniform integer [50, 500]            500
=========  =====  ===================================  =======

Valid slots: phi ~ N(-1.1, 0.3), psi ~ N(2.3, 0.4), omega = pi + N(0, 0.05)
(radians); padded slots are 0.  Upstream gradient on coordinates: N(0, 1) on
atoms of the chain, 0 on padding.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class Spec:
    name: str
    batch: int
    max_length: int
    seed: int
    kind: str
    lo: int = 0
    hi: int = 0


SPECS = {
    "uniform200": Spec("uniform200", 32, 200, 1, "const"),
    "uniform700": Spec("uniform700", 256, 700, 2, "const"),
    "mixed1000": Spec("mixed1000", 1024, 1000, 3, "loguniform", 30, 1000),
    "long16k": Spec("long16k", 64, 16384, 4, "uniform", 4096, 16384),
    "fit500": Spec("fit500", 512, 500, 5, "uniform", 50, 500),
}


def draw_lengths(spec: Spec, rng) -> np.ndarray:
    B = spec.batch
    if spec.kind == "const":
        n = np.full(B, spec.max_length)
    elif spec.kind == "loguniform":
        u = rng.uniform(math.log(spec.lo), math.log(spec.hi + 1), B)
        n = np.clip(np.floor(np.exp(u)), spec.lo, spec.hi)
    elif spec.kind == "uniform":
        n = rng.integers(spec.lo, spec.hi + 1, B)
    else:
        raise ValueError(spec.kind)
    return n.astype(np.int32)


def draw_angles(lengths: np.ndarray, max_length: int, rng) -> np.ndarray:
    B = len(lengths)
    phi = rng.normal(-1.1, 0.3, (B, max_length))
    psi = rng.normal(2.3, 0.4, (B, max_length))
    omg = math.pi + rng.normal(0.0, 0.05, (B, max_length))
    a = np.stack([phi, psi, omg], axis=1)
    mask = np.arange(max_length)[None, :] < lengths[:, None]
    a *= mask[:, None, :]
    return a.astype(np.float32)


def draw_grad(lengths: np.ndarray, max_length: int, rng) -> np.ndarray:
    B = len(lengths)
    g = rng.normal(0.0, 1.0, (B, 9 * max_length))
    mask = np.arange(9 * max_length)[None, :] < 9 * lengths[:, None]
    return (g * mask).astype(np.float32)


def workload(name: str, batch: int | None = None):
    """(angles f32 [B,3,L], lengths i32 [B], grad f32 [B,9L]) of a named workload.

    ``batch`` keeps the first ``batch`` chains (same draws)."""
    spec = SPECS[name]
    rng = np.random.default_rng(spec.seed)
    n = draw_lengths(spec, rng)
    a = draw_angles(n, spec.max_length, rng)
    g = draw_grad(n, spec.max_length, rng)
    if batch is not None:
        n, a, g = n[:batch], a[:batch], g[:batch]
    return a, n, g


def custom(lengths, max_length: int, seed: int):
    """Same angle/grad distributions for explicit lengths."""
    rng = np.random.default_rng(seed)
    n = np.asarray(lengths, np.int32)
    return draw_angles(n, max_length, rng), n, draw_grad(n, max_length, rng)
