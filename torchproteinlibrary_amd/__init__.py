"""torchproteinlibrary_amd -- MI355X-native ReducedModel.Angles2Backbone and
RMSD.Coords2RMSD.

Drop-in for ``TorchProteinLibrary.ReducedModel.Angles2Backbone``
(TorchProteinLibrary/ReducedModel/__init__.py:1) and
``TorchProteinLibrary.RMSD.Coords2RMSD`` (TorchProteinLibrary/RMSD/__init__.py):
replace ``from TorchProteinLibrary import ReducedModel, RMSD`` with
``from torchproteinlibrary_amd import ReducedModel, RMSD``.

Submodules load lazily, so the build recipe (``python -m
torchproteinlibrary_amd._build``) runs before the native extension exists.
"""
import importlib

__all__ = ["ReducedModel", "RMSD"]


def __getattr__(name):
    if name in __all__:
        return importlib.import_module("." + name, __name__)
    raise AttributeError(name)
