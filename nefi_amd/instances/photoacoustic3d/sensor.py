"""Finite sensor bandwidth: a zero-phase Gaussian low-pass of the pressure traces, on the GPU.

The contract of ``nefi/instances/photoacoustic3d/sensor.py``: ``gaussian_lowpass(traces, dt,
f_cutoff)`` applies ``H(f) = exp(-(f/f_c)^2 / 2)`` along time (zero-padded to 2n, no wrap-around;
``f_cutoff`` None or <= 0 is the identity), and ``SensorBandwidth(inner, dt_obs, f_cutoff)`` is
the trace operator ``inner`` followed by that filter: same primary field, homogeneity, required
fields and output shape as ``inner``, fidelity tag ``"<inner tag>-lp<f_c>"`` (the inner tag alone
without a filter), and resolution copies that keep the tag of the operator they came from. The
filter is an ``nw_filter`` plan; its backward is the same filter (the response is real and even).
"""

from __future__ import annotations

import torch

from ...operators.base import Fields, Operator
from ...physics.filters import filter_traces
from ...utils import shape_tuple

__all__ = ["SensorBandwidth", "gaussian_lowpass"]


def gaussian_lowpass(traces: torch.Tensor, dt: float, f_cutoff: float | None) -> torch.Tensor:
    """``exp(-(f/f_c)^2 / 2)`` along the last axis (``sensor.py:24-38``)."""
    if not f_cutoff or f_cutoff <= 0:
        return traces
    return filter_traces(traces, "gaussian", (float(dt), float(f_cutoff)))


class _FilteredTraces(Operator):
    """An operator whose traces pass through a fixed linear time filter. What the solver reads
    of an operator (primary field, homogeneity, required fields, output shape) is the inner
    operator's: a linear filter changes none of them."""

    traceable = False

    def __init__(self, inner: Operator, tag: str) -> None:
        super().__init__()
        self.inner = inner
        self.primary = inner.primary
        self.homogeneity = inner.homogeneity
        self.fidelity_tag = tag
        self._copies: dict[tuple[int, ...], Operator] = {}

    def _filter(self, traces: torch.Tensor) -> torch.Tensor:  # pragma: no cover - abstract
        raise NotImplementedError

    def _around(self, inner: Operator) -> Operator:  # pragma: no cover - abstract
        """The same filter around another inner operator."""
        raise NotImplementedError

    def forward(self, fields: Fields) -> torch.Tensor:
        return self._filter(self.inner(fields))

    def at_resolution(self, shape) -> Operator:
        key = shape_tuple(shape)
        moved = self.inner.at_resolution(key)
        if moved is self.inner:
            return self
        if key not in self._copies:
            copy = self._around(moved)
            copy.fidelity_tag = self.fidelity_tag  # one forward model across resolutions
            self._copies[key] = copy
        return self._copies[key]

    def output_shape(self, shape):
        return self.inner.output_shape(shape)

    def required_fields(self) -> tuple[str, ...]:
        return self.inner.required_fields()


class SensorBandwidth(_FilteredTraces):
    """A trace operator ``(..., n_t)`` followed by the sensors' Gaussian low-pass
    (``sensor.py:41-86``); ``f_cutoff`` None or <= 0 means ideal sensors."""

    def __init__(self, inner: Operator, dt_obs: float, f_cutoff: float | None) -> None:
        cutoff = float(f_cutoff) if f_cutoff and f_cutoff > 0 else None
        inner_tag = str(getattr(inner, "fidelity_tag", type(inner).__name__))
        super().__init__(inner, inner_tag if cutoff is None else f"{inner_tag}-lp{cutoff:g}")
        self.dt_obs = float(dt_obs)
        self.f_cutoff = cutoff

    def _filter(self, traces: torch.Tensor) -> torch.Tensor:
        return gaussian_lowpass(traces, self.dt_obs, self.f_cutoff)

    def _around(self, inner: Operator) -> Operator:
        return SensorBandwidth(inner, self.dt_obs, self.f_cutoff)

    def extra_repr(self) -> str:
        return f"f_cutoff={self.f_cutoff}, dt_obs={self.dt_obs:g}"
