"""The box a field lives on and its uniform grid of cells (the contract of ``nefi/domain.py``).

Only what the wave operators read is provided: the grid ``shape``, the box ``extent``, the cell
size per axis, the same box at another resolution, and the cell-centre coordinates. The
conventions are nefi's because receiver and source stencils are computed from them: cell ``i``
of an axis spanning ``[lo, hi)`` with ``n`` cells is centred at ``lo + (i + 1/2)(hi - lo)/n``,
and normalised coordinates map the box onto ``[-1, 1]``.
"""

from __future__ import annotations

from collections.abc import Sequence

import torch

from .errors import ShapeError
from .utils import shape_tuple


class Domain:
    """``shape`` cells over the box ``extent = ((lo, hi), ...)``, optionally named ``axes``.

    Immutable and hashable: two domains with the same grid and box are equal.
    """

    __slots__ = ("_shape", "_extent", "_axes")

    def __init__(self, shape, extent, axes: Sequence[str] | None = None) -> None:
        grid = shape_tuple(shape)
        box = tuple((float(a), float(b)) for a, b in extent)
        if len(box) != len(grid):
            raise ShapeError(f"extent has {len(box)} axes but shape {grid} has {len(grid)}")
        bad = [ax for ax, (a, b) in enumerate(box) if not b > a]
        if bad:
            raise ShapeError(f"extent must satisfy hi > lo on every axis, not on axes {bad}")
        names = None if axes is None else tuple(str(n) for n in axes)
        if names is not None and len(names) != len(grid):
            raise ShapeError(f"{len(names)} axis names for a {len(grid)}-D domain")
        object.__setattr__(self, "_shape", grid)
        object.__setattr__(self, "_extent", box)
        object.__setattr__(self, "_axes", names)

    def __setattr__(self, name, value):
        raise AttributeError("Domain is immutable")

    @classmethod
    def unit(cls, shape, axes: Sequence[str] | None = None) -> Domain:
        """``shape`` cells over the unit box ``[0, 1]^d``."""
        grid = shape_tuple(shape)
        return cls(grid, [(0.0, 1.0)] * len(grid), axes)

    # ---- the grid and the box ------------------------------------------------------------
    @property
    def shape(self) -> tuple[int, ...]:
        return self._shape

    @property
    def extent(self) -> tuple[tuple[float, float], ...]:
        return self._extent

    @property
    def axes(self) -> tuple[str, ...] | None:
        return self._axes

    @property
    def ndim(self) -> int:
        return len(self._shape)

    def _grid(self, shape) -> tuple[int, ...]:
        grid = self._shape if shape is None else shape_tuple(shape)
        if len(grid) != self.ndim:
            raise ShapeError(f"a {len(grid)}-D shape {grid} on a {self.ndim}-D domain")
        return grid

    def spacing(self, shape=None) -> tuple[float, ...]:
        """Cell size per axis on a grid of ``shape`` cells (default: this domain's grid)."""
        grid = self._grid(shape)
        return tuple((b - a) / n for (a, b), n in zip(self._extent, grid))

    def at(self, shape) -> Domain:
        """The same box on a grid of ``shape`` cells."""
        return Domain(self._grid(shape), self._extent, self._axes)

    # ---- coordinates of the cell centres -------------------------------------------------
    def _centres(self, shape, normalized: bool, device, dtype) -> list[torch.Tensor]:
        dtype = dtype or torch.get_default_dtype()
        out = []
        for (a, b), n in zip(self._extent, self._grid(shape)):
            half = torch.arange(n, device=device, dtype=dtype) + 0.5  # i + 1/2
            out.append(2.0 * half / n - 1.0 if normalized else a + half * (b - a) / n)
        return out

    def coords(self, shape=None, device=None, dtype=None) -> torch.Tensor:
        """Normalised cell centres in ``[-1, 1]``, ``(*shape, ndim)``."""
        axes = self._centres(shape, True, device, dtype)
        return torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1)

    def physical_coords(self, shape=None, device=None, dtype=None) -> torch.Tensor:
        """Physical cell centres, ``(*shape, ndim)``."""
        axes = self._centres(shape, False, device, dtype)
        return torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=-1)

    # ---- value semantics -------------------------------------------------------------------
    def __eq__(self, other) -> bool:
        return isinstance(other, Domain) and (self._shape, self._extent, self._axes) == (
            other._shape, other._extent, other._axes)

    def __hash__(self) -> int:
        return hash((self._shape, self._extent, self._axes))

    def __repr__(self) -> str:
        return f"Domain(shape={self._shape}, extent={self._extent}, axes={self._axes})"
