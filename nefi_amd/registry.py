"""Name → class registry, so operators build from ``{"type": name, ...}`` configs.

Same kinds, lookup rules (case-insensitive names) and ``build`` merge rule as
``nefi/registry.py:38-105``; only the kinds the wave path touches matter here, but every kind
name is accepted so configs written for nefi load unchanged.
"""

from __future__ import annotations

from collections.abc import Callable
from typing import Any

from .errors import ConfigError, RegistryError

KINDS = (
    "field",
    "encoding",
    "head",
    "operator",
    "loss",
    "postprocess",
    "instance",
    "baseline",
    "metric",
    "callback",
    "scene",
    "datagen",
)

_TABLE: dict[str, dict[str, Any]] = {k: {} for k in KINDS}


def register(kind: str, name: str | None = None) -> Callable:
    """Class decorator: file ``obj`` under ``kind`` as ``name`` (default: its ``__name__``)."""
    if kind not in _TABLE:
        raise RegistryError(f"unknown registry kind {kind!r}; known: {KINDS}")

    def _file(obj):
        key = name if name is not None else getattr(obj, "__name__", None)
        if key is None:
            raise RegistryError("cannot infer a registry name; pass one explicitly")
        _TABLE[kind][key.lower()] = obj
        return obj

    return _file


def get(kind: str, name: str) -> Any:
    table = _TABLE.get(kind)
    if table is None:
        raise RegistryError(f"unknown registry kind {kind!r}; known: {KINDS}")
    obj = table.get(name.lower())
    if obj is None:
        known = ", ".join(sorted(table)) or "(none)"
        raise RegistryError(f"no {kind} registered as {name!r}; known: {known}")
    return obj


def list_registered(kind: str | None = None) -> dict[str, list[str]]:
    kinds = (kind,) if kind else KINDS
    return {k: sorted(_TABLE[k]) for k in kinds}


def build(kind: str, cfg: Any, **extra: Any) -> Any:
    """Instantiate from a name, a ``{"type": name, **kwargs}`` dict, or pass an object through.

    Keyword arguments in ``cfg`` take precedence over ``extra``.
    """
    if isinstance(cfg, str):
        return get(kind, cfg)(**extra)
    if isinstance(cfg, dict):
        kwargs = dict(cfg)
        name = kwargs.pop("type", None)
        if name is None:
            raise ConfigError(f"config for {kind} needs a 'type' key: {cfg}")
        return get(kind, name)(**{**extra, **kwargs})
    return cfg
