"""MI355X (gfx950) implementation of SUNSET's LABFM hot path.

Python host of the C ABI in include/sunset_amd.h: node sets and formats
(`nodes`), the device solver (`core.Solver`)."""
from .core import Ensemble, Solver, SunsetError, lib  # noqa: F401
from . import nodes  # noqa: F401
