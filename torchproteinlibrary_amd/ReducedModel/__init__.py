"""ReducedModel layers (reference: TorchProteinLibrary/ReducedModel/__init__.py:1)."""
from .Angles2Backbone import Angles2Backbone  # noqa: F401
