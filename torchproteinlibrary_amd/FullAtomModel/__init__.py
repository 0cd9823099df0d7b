"""FullAtomModel layers built for MI355X: the full-atom builder
Angles2Coords, the coordinate transforms and the PDB reader (reference:
TorchProteinLibrary/FullAtomModel/__init__.py:1-4).  Coords2TypedCoords is not
part of this build."""
from .Angles2Coords import *  # noqa: F401,F403
from .CoordsTransform import *  # noqa: F401,F403
from .PDB2Coords import *  # noqa: F401,F403
