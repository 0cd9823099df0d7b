"""RMSD layers (reference: TorchProteinLibrary/RMSD/__init__.py)."""
from .Coords2RMSD import Coords2RMSD  # noqa: F401
