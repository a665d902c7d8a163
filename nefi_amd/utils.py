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


def resample(x: torch.Tensor, shape: Sequence[int], mode: str = "auto") -> torch.Tensor:
    """Resample the trailing ``len(shape)`` dims to ``shape``.

    ``auto`` averages over areas when every axis shrinks and interpolates (bi/tri)linearly with
    cell-centred alignment otherwise, as ``nefi/utils/tensor.py:35-69`` does. The wave operators
    call it on the host, once, for the known sound speed of the photoacoustic operator.
    """
    shape = shape_tuple(shape)
    d = len(shape)
    if d not in (1, 2, 3):
        raise ShapeError(f"resample supports 1-3 spatial dims, got target shape {shape}")
    if x.ndim < d:
        raise ShapeError(f"tensor with shape {tuple(x.shape)} has fewer dims than target {shape}")
    old = tuple(x.shape[-d:])
    if old == shape:
        return x
    lead = tuple(x.shape[:-d])
    xb = x.reshape(-1, 1, *old)
    if mode == "auto":
        mode = "area" if all(o >= n for o, n in zip(old, shape)) else "linear"
    if mode == "linear":
        kind = ("linear", "bilinear", "trilinear")[d - 1]
        y = F.interpolate(xb, size=shape, mode=kind, align_corners=False)
    elif mode in ("area", "nearest"):
        y = F.interpolate(xb, size=shape, mode=mode)
    else:
        raise ValueError(f"unknown resample mode {mode!r}")
    return y.reshape(*lead, *shape)
