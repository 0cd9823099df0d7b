"""FullAtomModel layers built for MI355X: the coordinate transforms
(reference: TorchProteinLibrary/FullAtomModel/__init__.py:3).  Angles2Coords,
PDB2Coords and Coords2TypedCoords are not part of this build."""
from .CoordsTransform import *  # noqa: F401,F403
