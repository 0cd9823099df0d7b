"""Loads the native extension ``_ReducedModel`` (and through it libtpl_a2b.so).

There is no fallback: if the extension is not built, importing this module
raises ImportError that says how to build it.
"""
from __future__ import annotations

import importlib
import os

from . import _build

try:
    _ext = importlib.import_module("torchproteinlibrary_amd._ReducedModel")
except ImportError as e:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "torchproteinlibrary_amd: the native extension _ReducedModel is not built "
        f"(expected {_build.ext_path()} linking {_build.lib_path()}); run "
        "`python -m torchproteinlibrary_amd._build` (hipcc for gfx950 is required)"
    ) from e

ext = _ext
LIB_PATH = _build.lib_path()
EXT_PATH = os.path.abspath(_ext.__file__)
