from .Coords2RMSD import Coords2RMSD, Coords2RMSDFunction  # noqa: F401
