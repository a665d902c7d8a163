"""Error classes of the wave path, with the same names and bases as nefi's (``nefi/errors.py:6-27``).

Callers catch these by class (``ConfigError`` for bad settings, ``ShapeError`` for wrong field
shapes, ``OperatorError`` for invalid inputs at run time), so the hierarchy is kept identical.
"""

from __future__ import annotations


class NefiError(Exception):
    """Root of every error raised by this package."""


class ShapeError(NefiError, ValueError):
    """Inconsistent tensor or grid shapes."""


class ConfigError(NefiError, ValueError):
    """Invalid operator or solver settings."""


class RegistryError(NefiError, KeyError):
    """Unknown registry kind or name."""


class SolverError(NefiError, RuntimeError):
    """The optimisation cannot go on."""


class OperatorError(NefiError, RuntimeError):
    """Invalid input to a forward operator (for example a sound speed above ``c_max``)."""


class NativeError(NefiError, RuntimeError):
    """The HIP library is missing, failed to load, or returned an error code."""
