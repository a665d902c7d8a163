"""Error classes of the wave path: nefi's own classes where nefi is installed (``nefi/errors.py:6-27``).

nefi's CLI and solver catch ``NefiError`` and its subclasses by class, so the errors raised here
must be those classes, not look-alikes. When the ``nefi`` package can be imported (the plugin
case: nefi loads this package), ``NefiError``, ``ShapeError``, ``ConfigError``,
``RegistryError``, ``SolverError`` and ``OperatorError`` are nefi's classes themselves.
Otherwise (this package used on its own) a hierarchy of the same names and bases is defined
here. ``NativeError`` (the HIP library failed) is this package's own, derived from
``NefiError`` either way. ``NEFI_AMD_STANDALONE_ERRORS=1`` keeps the local classes.
"""

from __future__ import annotations

import importlib
import importlib.util
import os

__all__ = ["NefiError", "ShapeError", "ConfigError", "RegistryError", "SolverError",
           "OperatorError", "NativeError", "BOUND_TO_NEFI"]


def _nefi_errors():
    """nefi's error module if nefi is importable (and not switched off), else None."""
    if os.environ.get("NEFI_AMD_STANDALONE_ERRORS", "0") == "1":
        return None
    try:
        if importlib.util.find_spec("nefi") is None:
            return None
        return importlib.import_module("nefi.errors")
    except (ImportError, ValueError):
        return None


_ref = _nefi_errors()
BOUND_TO_NEFI = _ref is not None

if _ref is not None:
    NefiError = _ref.NefiError
    ShapeError = _ref.ShapeError
    ConfigError = _ref.ConfigError
    RegistryError = _ref.RegistryError
    SolverError = _ref.SolverError
    OperatorError = _ref.OperatorError
else:

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
