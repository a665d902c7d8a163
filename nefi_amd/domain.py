"""Rectangular physical domain on a uniform cell-centred grid (``nefi/domain.py:19-156``).

The wave path needs the spacing, the resolution change ``at`` and the coordinate grids; their
conventions (cell centres at ``lo + (i + ½)·h``, normalised coordinates in ``[-1, 1]``) are the
reference's, because receiver and source stencils are derived from them.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass, field

import torch

from .errors import ShapeError
from .utils import shape_tuple


@dataclass(frozen=True)
class Domain:
    """Grid ``shape`` over the box ``extent = ((lo, hi), ...)``; optional axis names."""

    shape: tuple[int, ...]
    extent: tuple[tuple[float, float], ...]
    axes: tuple[str, ...] | None = field(default=None)

    def __post_init__(self) -> None:
        shape = shape_tuple(self.shape)
        ext = tuple((float(lo), float(hi)) for lo, hi in self.extent)
        object.__setattr__(self, "shape", shape)
        object.__setattr__(self, "extent", ext)
        if len(ext) != len(shape):
            raise ShapeError(f"extent has {len(ext)} axes but shape {shape} has {len(shape)}")
        if any(not hi > lo for lo, hi in ext):
            raise ShapeError(f"extent must satisfy hi > lo, got {ext}")
        if self.axes is not None:
            names = tuple(self.axes)
            if len(names) != len(shape):
                raise ShapeError("axes must have one name per dimension")
            object.__setattr__(self, "axes", names)

    @staticmethod
    def unit(shape, axes: Sequence[str] | None = None) -> Domain:
        shape = shape_tuple(shape)
        return Domain(shape, tuple((0.0, 1.0) for _ in shape), tuple(axes) if axes else None)

    @staticmethod
    def from_spacing(shape, spacing, origin=0.0, axes: Sequence[str] | None = None) -> Domain:
        shape = shape_tuple(shape)
        d = len(shape)
        sp = tuple(spacing) if isinstance(spacing, (tuple, list)) else (float(spacing),) * d
        og = tuple(origin) if isinstance(origin, (tuple, list)) else (float(origin),) * d
        ext = tuple((float(o), float(o + n * s)) for n, s, o in zip(shape, sp, og))
        return Domain(shape, ext, tuple(axes) if axes else None)

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> tuple[float, ...]:
        return tuple(hi - lo for lo, hi in self.extent)

    @property
    def numel(self) -> int:
        out = 1
        for n in self.shape:
            out *= n
        return out

    def spacing(self, shape=None) -> tuple[float, ...]:
        """Cell size per axis at ``shape`` (default: the native shape)."""
        shape = self.shape if shape is None else shape_tuple(shape)
        if len(shape) != self.ndim:
            raise ShapeError(f"shape {shape} does not match domain ndim {self.ndim}")
        return tuple((hi - lo) / n for (lo, hi), n in zip(self.extent, shape))

    def at(self, shape) -> Domain:
        """The same box at another resolution."""
        shape = shape_tuple(shape)
        if len(shape) != self.ndim:
            raise ShapeError(f"shape {shape} does not match domain ndim {self.ndim}")
        return Domain(shape, self.extent, self.axes)

    def coarsen(self, factor: int = 2, min_size: int = 4) -> Domain:
        return self.at(tuple(max(min_size, n // factor) for n in self.shape))

    def refine(self, factor: int = 2) -> Domain:
        return self.at(tuple(n * factor for n in self.shape))

    def axis_coords(self, shape=None, normalized: bool = True, device=None, dtype=None):
        shape = self.shape if shape is None else shape_tuple(shape)
        dtype = dtype or torch.get_default_dtype()
        out = []
        for (lo, hi), n in zip(self.extent, shape):
            i = torch.arange(n, device=device, dtype=dtype)
            out.append(-1.0 + (2.0 * i + 1.0) / n if normalized else lo + (i + 0.5) * (hi - lo) / n)
        return out

    def coords(self, shape=None, device=None, dtype=None) -> torch.Tensor:
        """Normalised cell centres, ``(*shape, ndim)`` in ``[-1, 1]``."""
        mesh = torch.meshgrid(*self.axis_coords(shape, True, device, dtype), indexing="ij")
        return torch.stack(mesh, dim=-1)

    def physical_coords(self, shape=None, device=None, dtype=None) -> torch.Tensor:
        """Physical cell centres, ``(*shape, ndim)``."""
        mesh = torch.meshgrid(*self.axis_coords(shape, False, device, dtype), indexing="ij")
        return torch.stack(mesh, dim=-1)

    def to_normalized(self, x: torch.Tensor) -> torch.Tensor:
        lo = torch.tensor([e[0] for e in self.extent], device=x.device, dtype=x.dtype)
        hi = torch.tensor([e[1] for e in self.extent], device=x.device, dtype=x.dtype)
        return 2.0 * (x - lo) / (hi - lo) - 1.0

    def to_physical(self, u: torch.Tensor) -> torch.Tensor:
        lo = torch.tensor([e[0] for e in self.extent], device=u.device, dtype=u.dtype)
        hi = torch.tensor([e[1] for e in self.extent], device=u.device, dtype=u.dtype)
        return lo + (u + 1.0) * 0.5 * (hi - lo)

    def to_dict(self) -> dict:
        return {
            "shape": list(self.shape),
            "extent": [list(e) for e in self.extent],
            "axes": list(self.axes) if self.axes else None,
        }

    @staticmethod
    def from_dict(d: dict) -> Domain:
        axes = d.get("axes")
        return Domain(tuple(d["shape"]), tuple(tuple(e) for e in d["extent"]),
                      tuple(axes) if axes else None)
