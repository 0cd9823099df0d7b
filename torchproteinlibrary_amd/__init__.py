"""torchproteinlibrary_amd -- MI355X-native ReducedModel.Angles2Backbone.

Drop-in for ``TorchProteinLibrary.ReducedModel.Angles2Backbone``
(TorchProteinLibrary/ReducedModel/__init__.py:1): replace
``from TorchProteinLibrary import ReducedModel`` with
``from torchproteinlibrary_amd import ReducedModel``.

Submodules load lazily, so the build recipe (``python -m
torchproteinlibrary_amd._build``) runs before the native extension exists.
"""
import importlib

__all__ = ["ReducedModel"]


def __getattr__(name):
    if name == "ReducedModel":
        return importlib.import_module(".ReducedModel", __name__)
    raise AttributeError(name)
