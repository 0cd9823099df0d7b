from .CoordsTransform import *  # noqa: F401,F403
