"""Trace filters on the GPU: nefi's zero-padded FFT low-passes as plans of ``libnefi_wave``.

nefi filters traces along time as ``irfft(rfft(x, N) * H, N)[..., :n]`` with a real, even
response ``H`` (``lowpass``, ``nefi/instances/wave_fwi/losses.py:29-50``: cosine-tapered box,
``N = next_fast_len(2n)``; ``gaussian_lowpass``, ``nefi/instances/photoacoustic3d/sensor.py:24-38``:
``exp(-(f/f_c)^2/2)``, ``N = 2n``). For ``N >= 2n - 1`` that is the symmetric Toeplitz map with
taps ``h = irfft(H, N)[:n]``, which is its own adjoint. Here the taps are computed once per
(response, n) on the host in float64 and handed to a filter plan (``nw_filter_create``); the HIP
kernel applies them with an in-LDS FFT on a power-of-two length, so results equal the
reference's filter to rounding while the per-call window, padding and transforms disappear.

Autograd: the filter's backward is the same filter (self-adjoint); the fused misfit of
``BandLimitedMSE`` (``MisfitFunction``) computes residual, filter, mask and both masked means in
one pass and its adjoint source in a second.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from .. import _native as nat
from ..errors import NativeError


def next_fast_len(n: int) -> int:
    """Smallest 2^a 3^b 5^c >= n: the FFT length rule ``next_fast_len(n, real=True)`` of
    ``nefi/utils/tensor.py:21-23`` (scipy's 5-smooth rule for real transforms)."""
    n = int(n)
    if n <= 6:
        return max(n, 1)
    best = None
    p5 = 1
    while p5 < 2 * n:
        p35 = p5
        while p35 < 2 * n:
            # smallest power of two that lifts p35 to >= n
            m = p35
            while m < n:
                m *= 2
            if best is None or m < best:
                best = m
            p35 *= 3
        p5 *= 5
    return int(best)


def cosine_response(n_fft: int, dt: float, f_max: float, taper: float) -> np.ndarray:
    """The window of ``losses.py:29-37``: 1 up to (1 - taper) f_max, a cosine roll-off to 0 at
    f_max, 0 above, on the ``rfftfreq(n_fft, dt)`` grid (float64)."""
    f = np.fft.rfftfreq(n_fft, d=dt)
    f1 = (1.0 - taper) * f_max
    w = np.ones_like(f)
    roll = (f > f1) & (f < f_max)
    w[roll] = 0.5 * (1.0 + np.cos(math.pi * (f[roll] - f1) / max(f_max - f1, 1e-30)))
    w[f >= f_max] = 0.0
    return w


def gaussian_response(n_fft: int, dt: float, f_cutoff: float) -> np.ndarray:
    """``exp(-(f/f_c)^2 / 2)`` on the ``rfftfreq(n_fft, dt)`` grid (``sensor.py:33-35``)."""
    f = np.fft.rfftfreq(n_fft, d=float(dt))
    return np.exp(-0.5 * (f / float(f_cutoff)) ** 2)


def taps_of(response: np.ndarray, n_fft: int, n: int) -> np.ndarray:
    """Impulse response at lags 0..n-1 of the zero-padded FFT filter ``response`` (length
    n_fft // 2 + 1): the time-domain kernel the plan convolves with."""
    if n_fft < 2 * n - 1:
        raise ValueError("the FFT filter wraps around: n_fft must be >= 2n - 1")
    return np.ascontiguousarray(np.fft.irfft(response, n=n_fft)[:n], dtype=np.float64)


class FilterPlan:
    """A ``nw_filter`` plan: fixed taps, trace length and dtype, on one device."""

    def __init__(self, taps: np.ndarray, dtype: torch.dtype, device: torch.device) -> None:
        self.n = int(taps.shape[0])
        self.dtype = dtype
        self.device = device
        lib = nat.lib()
        h = ctypes.c_void_p()
        arr = np.ascontiguousarray(taps, dtype=np.float64)
        code = nat.NW_F64 if dtype == torch.float64 else nat.NW_F32
        with torch.cuda.device(device):
            nat.check(
                lib.nw_filter_create(code, self.n, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                     ctypes.byref(h)),
                "nw_filter_create",
            )
        self.handle = h
        self.length = int(lib.nw_filter_length(h))

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        h = getattr(self, "handle", None)
        if h is not None and nat._LIB is not None:
            nat._LIB.nw_filter_destroy(h)
            self.handle = None

    def apply(self, x: torch.Tensor, x_sub: torch.Tensor | None = None,
              mask: torch.Tensor | None = None, out: torch.Tensor | None = None,
              filter: bool = True, scale: float = 1.0,
              sums: torch.Tensor | None = None) -> None:
        """One ``nw_filter_apply`` over rows of contiguous ``(rows, n)`` device buffers."""
        rows = x.numel() // self.n
        for t in (x_sub, mask, out):
            if t is not None and (t.numel() != x.numel() or not t.is_contiguous()):
                raise NativeError("filter buffers must be contiguous and of one shape")
        nat.check(
            nat.lib().nw_filter_apply(self.handle, x.data_ptr(), nat.ptr(x_sub), nat.ptr(mask),
                                      nat.ptr(out), rows, int(filter), float(scale),
                                      nat.ptr(sums), nat.stream_of(self.device)),
            "nw_filter_apply",
        )


_PLANS: dict[tuple, FilterPlan] = {}


def plan_for(kind: str, params: tuple, n: int, dtype: torch.dtype,
             device: torch.device) -> FilterPlan:
    """The cached plan of one response (``kind`` in {"cosine", "gaussian", "unit"})."""
    key = (kind, params, int(n), dtype, str(device))
    plan = _PLANS.get(key)
    if plan is None:
        if kind == "cosine":
            dt, f_max, taper = params
            n_fft = next_fast_len(2 * n)
            taps = taps_of(cosine_response(n_fft, dt, f_max, taper), n_fft, n)
        elif kind == "gaussian":
            dt, f_cutoff = params
            taps = taps_of(gaussian_response(2 * n, dt, f_cutoff), 2 * n, n)
        else:  # no filter: the plan only carries n for the fused misfit passes
            taps = np.zeros(n)
            taps[0] = 1.0
        plan = FilterPlan(taps, dtype, device)
        _PLANS[key] = plan
    return plan


def _device_rows(x: torch.Tensor, dev: torch.device) -> torch.Tensor:
    if x.dtype not in (torch.float32, torch.float64):
        x = x.float()
    return x.detach().to(dev).contiguous()


class FilterFunction(torch.autograd.Function):
    """y = filter(x) along the last axis; backward applies the same (self-adjoint) filter."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, plan: FilterPlan) -> torch.Tensor:
        xs = _device_rows(x, plan.device)
        y = torch.empty_like(xs)
        plan.apply(xs, out=y)
        ctx.plan = plan
        ctx.src = (x.device, x.dtype)
        return y.to(device=x.device, dtype=x.dtype)

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        plan = ctx.plan
        gs = _device_rows(g, plan.device)
        out = torch.empty_like(gs)
        plan.apply(gs, out=out)
        dev, dt = ctx.src
        return out.to(device=dev, dtype=dt), None


def filter_traces(x: torch.Tensor, kind: str, params: tuple) -> torch.Tensor:
    """Apply a response of ``plan_for`` along the last axis of ``x`` (any leading shape)."""
    n = x.shape[-1]
    if x.numel() == 0:
        return x.clone()
    dev = nat.compute_device(x)
    dtype = torch.float64 if x.dtype == torch.float64 else torch.float32
    plan = plan_for(kind, params, n, dtype, dev)
    return FilterFunction.apply(x, plan)


class MisfitFunction(torch.autograd.Function):
    """``masked_mean(F(pred - obs)^2) [/ masked_mean(F(obs)^2)]`` and its gradient in ``pred``.

    F is the plan's filter (or the identity); ``mask`` (or None) has pred's shape. Forward: one
    fused pass writes Z = mask * F(pred - obs) and the float64 sums of mask * F(r)^2 and mask,
    a second (relative only) the sums of mask * F(obs)^2. Backward: the adjoint source
    dL/dpred = F(Z) * 2 / (sum(mask) * den), F being self-adjoint. ``obs`` and ``mask`` are data
    and get no gradient.
    """

    @staticmethod
    def forward(ctx, pred, obs, mask, plan: FilterPlan, filtered: bool, relative: bool):
        dev = plan.device
        p = _device_rows(pred, dev)
        o = _device_rows(obs, dev).to(p.dtype)
        m = None if mask is None else _device_rows(mask, dev).to(p.dtype)
        z = torch.empty_like(p)
        sums = torch.empty(4, device=dev, dtype=torch.float64)
        plan.apply(p, x_sub=o, mask=m, out=z, filter=filtered, sums=sums[:2])
        msum = sums[1].clamp_min(1e-12)
        num = sums[0] / msum
        if relative:
            plan.apply(o, mask=m, filter=filtered, sums=sums[2:])
            den = (sums[2] / sums[3].clamp_min(1e-12)).clamp_min(1e-30)
        else:
            den = torch.ones((), device=dev, dtype=torch.float64)
        loss = num / den
        ctx.plan = plan
        ctx.filtered = filtered
        ctx.save_for_backward(z, 2.0 / (msum * den))
        ctx.src = (pred.device, pred.dtype)
        return loss.to(device=pred.device, dtype=pred.dtype)

    @staticmethod
    def backward(ctx, g):
        z, factor = ctx.saved_tensors
        out = torch.empty_like(z)
        ctx.plan.apply(z, out=out, filter=ctx.filtered)
        out.mul_((factor * g.to(factor.device, torch.float64)).to(out.dtype))
        dev, dt = ctx.src
        return out.to(device=dev, dtype=dt), None, None, None, None, None


def misfit(pred: torch.Tensor, obs: torch.Tensor, mask: torch.Tensor | None, kind: str,
           params: tuple, relative: bool) -> torch.Tensor:
    """The fused (band-limited) relative MSE; ``kind`` "unit" skips the filter."""
    n = pred.shape[-1]
    dev = nat.compute_device(pred)
    dtype = torch.float64 if pred.dtype == torch.float64 else torch.float32
    if mask is not None:
        mask = torch.broadcast_to(mask, pred.shape)
    obs = torch.broadcast_to(obs, pred.shape)
    plan = plan_for(kind, params, n, dtype, dev)
    return MisfitFunction.apply(pred, obs, mask, plan, kind != "unit", relative)
