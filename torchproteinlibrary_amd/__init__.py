"""torchproteinlibrary_amd -- MI355X-native ReducedModel.Angles2Backbone,
RMSD.Coords2RMSD and the FullAtomModel.CoordsTransform ops.

Drop-in for ``TorchProteinLibrary.ReducedModel.Angles2Backbone``
(TorchProteinLibrary/ReducedModel/__init__.py:1) and
``TorchProteinLibrary.RMSD.Coords2RMSD`` (TorchProteinLibrary/RMSD/__init__.py):
replace ``from TorchProteinLibrary import ReducedModel, RMSD`` with
``from torchproteinlibrary_amd import ReducedModel, RMSD``; the coordinate
transforms are ``torchproteinlibrary_amd.FullAtomModel.CoordsTransform``.

Submodules load lazily, so the build recipe (``python -m
torchproteinlibrary_amd._build``) runs before the native extension exists.
"""
import importlib

__all__ = ["ReducedModel", "RMSD", "FullAtomModel"]


def __getattr__(name):
    if name in __all__:
        return importlib.import_module("." + name, __name__)
    raise AttributeError(name)
