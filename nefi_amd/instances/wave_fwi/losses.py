"""Band-limited trace misfit for frequency-continuation FWI, on the GPU.

Same functions, arguments and values as ``nefi/instances/wave_fwi/losses.py``: ``lowpass`` is the
zero-phase low-pass with a cosine roll-off (zero-padded to ``next_fast_len(2n)``, no wrap-around)
and ``BandLimitedMSE`` the (relative) masked MSE of low-passed residuals. Here the filter is a
``nw_filter`` plan and the misfit one fused HIP pass (residual, filter, mask, both masked means)
whose backward is the adjoint source the wave operator's reverse sweep injects.
"""

from __future__ import annotations

from torch import nn

from ...physics.filters import filter_traces, misfit
from ...registry import register

__all__ = ["BandLimitedMSE", "lowpass"]


def lowpass(x, dt: float, f_max: float, taper: float = 0.3):
    """Zero-phase low-pass along the last (time) axis with a cosine roll-off
    (``losses.py:40-50``): flat up to ``(1 - taper) f_max``, zero from ``f_max`` (Hz)."""
    return filter_traces(x, "cosine", (float(dt), float(f_max), float(taper)))


@register("loss", "band_limited_mse")
class BandLimitedMSE(nn.Module):
    """(Relative) MSE between low-pass filtered predicted and observed traces
    (``losses.py:53-86``).

    ``L = mean(|lowpass(pred - obs)|^2) / mean(|lowpass(obs)|^2)`` over the observed entries
    (``relative=True``); ``f_max=None`` leaves the traces unfiltered. ``forward`` reads
    ``ctx.pred`` and ``ctx.obs`` (``.data`` and ``.mask``) like nefi's data losses.
    """

    is_data = True
    step_dependent = False

    def __init__(self, f_max: float | None, dt: float, taper: float = 0.3, relative: bool = True,
                 name: str | None = None) -> None:
        super().__init__()
        self.name = name or type(self).__name__.lower()
        self.f_max = None if f_max is None else float(f_max)
        self.dt = float(dt)
        self.taper = float(taper)
        self.relative = relative

    def forward(self, ctx):
        obs = ctx.obs
        mask = getattr(obs, "mask", None)
        if self.f_max is None:
            return misfit(ctx.pred, obs.data, mask, "unit", (), self.relative)
        return misfit(ctx.pred, obs.data, mask, "cosine", (self.dt, self.f_max, self.taper),
                      self.relative)

    def extra_repr(self) -> str:
        return f"f_max={self.f_max}, dt={self.dt}, relative={self.relative}"
