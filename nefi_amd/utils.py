"""Small tensor helpers the wave operators need (``nefi/utils/tensor.py:15-69``)."""

from __future__ import annotations

from collections.abc import Sequence

import torch
import torch.nn.functional as F

from .errors import ShapeError


def shape_tuple(shape) -> tuple[int, ...]:
    if isinstance(shape, int):
        return (int(shape),)
    return tuple(int(n) for n in shape)


_LINEAR = {1: "linear", 2: "bilinear", 3: "trilinear"}


def resample(x: torch.Tensor, shape: Sequence[int], mode: str = "auto") -> torch.Tensor:
    """``x`` with its trailing ``len(shape)`` dims resampled to ``shape``.

    The rule of ``nefi/utils/tensor.py:35-69``: ``auto`` takes cell averages (``area``) when no
    axis grows and cell-centred (bi/tri)linear interpolation otherwise; ``area`` and ``nearest``
    can be asked for directly. Used on the host, once, for the known sound speed of the
    photoacoustic operator at another resolution.
    """
    target = shape_tuple(shape)
    d = len(target)
    if d not in _LINEAR:
        raise ShapeError(f"resample supports 1-3 spatial dims, got target shape {target}")
    if x.ndim < d:
        raise ShapeError(f"tensor with shape {tuple(x.shape)} has fewer dims than target {target}")
    source = tuple(x.shape[-d:])
    if source == target:
        return x
    if mode == "auto":
        mode = "area" if all(s >= t for s, t in zip(source, target)) else "linear"
    if mode == "linear":
        how = dict(mode=_LINEAR[d], align_corners=False)
    elif mode in ("area", "nearest"):
        how = dict(mode=mode)
    else:
        raise ValueError(f"unknown resample mode {mode!r}")
    out = F.interpolate(x.reshape((-1, 1) + source), size=target, **how)
    return out.reshape(tuple(x.shape[:-d]) + target)
