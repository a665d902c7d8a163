"""Finite sensor bandwidth: a zero-phase Gaussian low-pass of the pressure traces, on the GPU.

Same function and operator as ``nefi/instances/photoacoustic3d/sensor.py``:
``gaussian_lowpass`` applies ``H(f) = exp(-(f/f_c)^2 / 2)`` along time (zero-padded to 2n, no
wrap-around) and ``SensorBandwidth`` wraps a trace operator with it. The filter is a
``nw_filter`` plan; its backward is the same filter (the response is real and even).
"""

from __future__ import annotations

from collections.abc import Sequence

import torch

from ...operators.base import Fields, Operator
from ...physics.filters import filter_traces

__all__ = ["SensorBandwidth", "gaussian_lowpass"]


def gaussian_lowpass(traces: torch.Tensor, dt: float, f_cutoff: float | None) -> torch.Tensor:
    """``exp(-(f/f_c)^2 / 2)`` along the last axis (``sensor.py:24-38``); ``f_cutoff`` None or
    <= 0 returns ``traces`` itself."""
    if f_cutoff is None or f_cutoff <= 0:
        return traces
    return filter_traces(traces, "gaussian", (float(dt), float(f_cutoff)))


class SensorBandwidth(Operator):
    """A trace operator ``(..., n_t)`` followed by the sensors' low-pass (``sensor.py:41-86``)."""

    traceable = False

    def __init__(self, inner: Operator, dt_obs: float, f_cutoff: float | None) -> None:
        super().__init__()
        self.inner = inner
        self.primary = inner.primary
        self.homogeneity = inner.homogeneity
        self.dt_obs = float(dt_obs)
        self.f_cutoff = None if f_cutoff is None or f_cutoff <= 0 else float(f_cutoff)
        tag = getattr(inner, "fidelity_tag", type(inner).__name__)
        self.fidelity_tag = f"{tag}-lp{self.f_cutoff:g}" if self.f_cutoff else str(tag)
        self._res: dict[tuple[int, ...], SensorBandwidth] = {}

    def forward(self, fields: Fields) -> torch.Tensor:
        return gaussian_lowpass(self.inner(fields), self.dt_obs, self.f_cutoff)

    def at_resolution(self, shape: Sequence[int]) -> SensorBandwidth:
        shape = tuple(int(s) for s in shape)
        inner = self.inner.at_resolution(shape)
        if inner is self.inner:
            return self
        if shape not in self._res:
            op = SensorBandwidth(inner, self.dt_obs, self.f_cutoff)
            op.fidelity_tag = self.fidelity_tag
            self._res[shape] = op
        return self._res[shape]

    def output_shape(self, shape: Sequence[int]) -> tuple[int, ...] | None:
        return self.inner.output_shape(shape)

    def required_fields(self) -> tuple[str, ...]:
        return self.inner.required_fields()

    def extra_repr(self) -> str:
        return f"f_cutoff={self.f_cutoff}, dt_obs={self.dt_obs:g}"
