"""Reference: TorchProteinLibrary/ReducedModel/Angles2Backbone/__init__.py:1."""
from .Angles2Backbone import (  # noqa: F401
    Angles2Backbone,
    Angles2BackboneCPUFunction,
    Angles2BackboneGPUFunction,
    Angles2BackbonePlainGPUFunction,
    Angles2BackboneStep,
)
