"""FullAtomModel layers built for MI355X: the full-atom builder
Angles2Coords and the coordinate transforms (reference:
TorchProteinLibrary/FullAtomModel/__init__.py:1-3).  PDB2Coords and
Coords2TypedCoords are not part of this build."""
from .Angles2Coords import *  # noqa: F401,F403
from .CoordsTransform import *  # noqa: F401,F403
