"""Loads the native extension ``_FullAtomModel`` (and through it libtpl_xform.so).

There is no fallback: if the extension is not built, importing this module
raises ImportError that says how to build it.
"""
from __future__ import annotations

import importlib
import os

from . import _build

try:
    _ext = importlib.import_module("torchproteinlibrary_amd._FullAtomModel")
except ImportError as e:  # pragma: no cover - exercised only on broken installs
    raise ImportError(
        "torchproteinlibrary_amd: the native extension _FullAtomModel is not built (expected "
        f"{_build.ext_path(_build.XFORM_EXT_NAME)} linking {_build.lib_path(_build.XFORM_LIB_NAME)}); "
        "run `python -m torchproteinlibrary_amd._build` (hipcc for gfx950 is required)"
    ) from e

ext = _ext
LIB_PATH = _build.lib_path(_build.XFORM_LIB_NAME)
EXT_PATH = os.path.abspath(_ext.__file__)
