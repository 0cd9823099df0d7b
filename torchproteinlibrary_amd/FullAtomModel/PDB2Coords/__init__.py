from .PDB2Coords import *  # noqa: F401,F403
