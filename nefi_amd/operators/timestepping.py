"""Generic differentiable time integration, with the leapfrog wave step on HIP kernels.

Same names, arguments and semantics as ``nefi/operators/timestepping.py`` for its wave parts:

* ``TimeStepper`` (``:326-465``): ``state_0 = init(fields)``, ``state_{n+1} = step(state_n,
  params, dt, n)``, observations stacked along a new leading axis; ``grad_mode="checkpoint"``
  recomputes blocks of ``checkpoint_every`` steps in the backward pass. The step is any Python
  callable (bring your own physics), so the loop stays in Python.
* ``wave_step`` (``:167-214``): one leapfrog step of ``u_tt = c^2 lap(u) + s`` (optionally
  damped), boundaries periodic / Neumann / Dirichlet / absorbing (first-order Mur faces,
  ``:145-164``). Here it is one fused HIP kernel (plus one small kernel per axis for the Mur
  faces) with a hand-written adjoint (``nw_ts_*``, ``include/nefi_stepper.h``).
* ``laplacian`` (``:73-87``), ``wave_initial_state`` (``:217-238``), ``stable_dt`` (``:104-137``).

``advection_diffusion_step`` (a different PDE) is not provided.
"""

from __future__ import annotations

import ctypes
import math
from collections.abc import Callable, Mapping, Sequence
from typing import Any

import torch
import torch.nn.functional as F
from torch.utils.checkpoint import checkpoint

from .. import _native as nat
from ..errors import OperatorError
from ..registry import register
from .base import Fields, Operator

State = torch.Tensor | tuple[torch.Tensor, ...]
StepFn = Callable[[State, Mapping[str, Any], float, int], State]

_BOUNDARIES = ("periodic", "neumann", "dirichlet", "absorbing")
_BC_CODE = {"periodic": nat.NW_TS_PERIODIC, "neumann": nat.NW_TS_NEUMANN,
            "dirichlet": nat.NW_TS_DIRICHLET, "absorbing": nat.NW_TS_ABSORBING}


# ---- native plumbing -----------------------------------------------------------------------
def _grid(shape: Sequence[int], spacing: Sequence[float], boundary: str,
          dtype: torch.dtype) -> nat.NwTsGrid:
    d = len(spacing)
    if d < 1 or d > 3 or len(shape) < d:
        raise OperatorError(f"{d} spatial dims over a tensor of shape {tuple(shape)}")
    g = nat.NwTsGrid()
    g.ndim = d
    g.dtype = nat.NW_F64 if dtype == torch.float64 else nat.NW_F32
    g.boundary = _BC_CODE[boundary]
    g.batch = math.prod(shape[: len(shape) - d])
    for a in range(3):
        k = a - (3 - d)
        g.shape[a] = int(shape[len(shape) - d + k]) if k >= 0 else 1
        g.h[a] = float(spacing[k]) if k >= 0 else 1.0
    return g


def _dev(t: torch.Tensor | None, dev: torch.device, dtype: torch.dtype) -> torch.Tensor | None:
    return None if t is None else t.detach().to(device=dev, dtype=dtype).contiguous()


def _full(v: Any, like: torch.Tensor) -> torch.Tensor:
    """A float / tensor argument broadcast to ``like``'s shape (differentiable)."""
    t = torch.as_tensor(v, dtype=like.dtype, device=like.device) if not torch.is_tensor(v) else v
    return torch.broadcast_to(t.to(dtype=like.dtype, device=like.device), like.shape).contiguous()


class _LaplacianFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, spacing, boundary):
        dev = nat.compute_device(u)
        dtype = torch.float64 if u.dtype == torch.float64 else torch.float32
        x = _dev(u, dev, dtype)
        out = torch.empty_like(x)
        g = _grid(x.shape, spacing, boundary, dtype)
        nat.check(nat.lib().nw_ts_laplacian(ctypes.byref(g), x.data_ptr(), out.data_ptr(),
                                            nat.stream_of(dev)), "nw_ts_laplacian")
        ctx.cfg = (spacing, boundary, u.device, u.dtype)
        return out.to(device=u.device, dtype=u.dtype)

    @staticmethod
    def backward(ctx, g):
        spacing, boundary, dev0, dt0 = ctx.cfg
        return _LaplacianFn.apply(g, spacing, boundary).to(device=dev0, dtype=dt0), None, None


def laplacian(u: torch.Tensor, spacing: Sequence[float], boundary: str = "neumann") -> torch.Tensor:
    """Second-order Laplacian over the trailing ``len(spacing)`` dims with ghost cells per the
    boundary (periodic wrap, Neumann / absorbing replicated edge, Dirichlet zero)."""
    if boundary not in _BOUNDARIES:
        raise OperatorError(f"unknown boundary {boundary!r}; choose one of {_BOUNDARIES}")
    return _LaplacianFn.apply(u, tuple(float(h) for h in spacing), boundary)


class _WaveStepFn(torch.autograd.Function):
    """u_next of one leapfrog step from (u_prev, u, c, s, gamma), all of u's shape (s, gamma may
    be None); backward: Mur faces undone per axis in reverse order, then the interior update."""

    @staticmethod
    def forward(ctx, u_prev, u, c, s, gam, spacing, dt, boundary):
        dev = nat.compute_device(u)
        dtype = torch.float64 if u.dtype == torch.float64 else torch.float32
        up_, u_, c_, s_, g_ = (_dev(t, dev, dtype) for t in (u_prev, u, c, s, gam))
        out = torch.empty_like(u_)
        grid = _grid(u_.shape, spacing, boundary, dtype)
        nat.check(nat.lib().nw_ts_wave_step(ctypes.byref(grid), float(dt), up_.data_ptr(),
                                             u_.data_ptr(), c_.data_ptr(), nat.ptr(s_),
                                             nat.ptr(g_), out.data_ptr(), 1, nat.stream_of(dev)),
                  "nw_ts_wave_step")
        ctx.save_for_backward(up_, u_, c_, s_, g_)
        ctx.cfg = (spacing, float(dt), boundary, dev, dtype, u.device, u.dtype)
        return out.to(device=u.device, dtype=u.dtype)

    @staticmethod
    def backward(ctx, g_next):
        up_, u_, c_, s_, g_ = ctx.saved_tensors
        spacing, dt, boundary, dev, dtype, dev0, dt0 = ctx.cfg
        lib, st = nat.lib(), nat.stream_of(dev)
        grid = _grid(u_.shape, spacing, boundary, dtype)
        need_c = ctx.needs_input_grad[2]
        gw = g_next.detach().to(device=dev, dtype=dtype).contiguous().clone()
        gu = torch.zeros_like(u_)
        gup = torch.empty_like(u_)
        gc = torch.zeros_like(u_) if need_c else None
        gs = torch.empty_like(u_) if (s_ is not None and ctx.needs_input_grad[3]) else None
        gg = torch.empty_like(u_) if (g_ is not None and ctx.needs_input_grad[4]) else None
        if boundary == "absorbing":
            axes = [a for a in range(3 - len(spacing), 3) if grid.shape[a] >= 3]
            before = {}
            if need_c and axes:  # w before each axis' Mur update (read next to the faces)
                w = torch.empty_like(u_)
                nat.check(lib.nw_ts_wave_step(ctypes.byref(grid), dt, up_.data_ptr(), u_.data_ptr(),
                                              c_.data_ptr(), nat.ptr(s_), nat.ptr(g_),
                                              w.data_ptr(), 0, st), "nw_ts_wave_step")
                for a in axes:
                    before[a] = w
                    w = w.clone()
                    nat.check(lib.nw_ts_mur(ctypes.byref(grid), a, dt, u_.data_ptr(), c_.data_ptr(),
                                            w.data_ptr(), st), "nw_ts_mur")
            for a in reversed(axes):
                nat.check(lib.nw_ts_mur_adjoint(ctypes.byref(grid), a, dt, u_.data_ptr(),
                                                c_.data_ptr(), nat.ptr(before.get(a)),
                                                gw.data_ptr(), gu.data_ptr(), nat.ptr(gc), st),
                          "nw_ts_mur_adjoint")
        nat.check(lib.nw_ts_wave_step_adjoint(ctypes.byref(grid), dt, up_.data_ptr(), u_.data_ptr(),
                                              c_.data_ptr(), nat.ptr(s_), nat.ptr(g_),
                                              gw.data_ptr(), gup.data_ptr(), gu.data_ptr(),
                                              nat.ptr(gc), nat.ptr(gs), nat.ptr(gg), st),
                  "nw_ts_wave_step_adjoint")

        def back(t):
            return None if t is None else t.to(device=dev0, dtype=dt0)

        return back(gup), back(gu), back(gc), back(gs), back(gg), None, None, None


def _resolve(value: Any, params: Mapping[str, Any], n: int, name: str) -> Any:
    """A field name -> ``params[name]``; a callable -> ``value(params, n)``; else the value."""
    if isinstance(value, str):
        if value not in params:
            raise OperatorError(
                f"stepper argument {name}={value!r} refers to a field that is not available; "
                f"fields: {tuple(params)}"
            )
        return params[value]
    if callable(value) and not torch.is_tensor(value):
        return value(params, n)
    return value


def wave_step(state: State, params: Mapping[str, Any], dt: float, n: int, *,
              spacing: Sequence[float], c: Any = "c", boundary: str = "absorbing",
              source: Any = None, damping: Any = None) -> State:
    """One leapfrog step ``(u_prev, u) -> (u, u_next)`` of ``u_tt = c^2 lap(u) + s``
    (``u_next = 2u - u_prev + (c dt)^2 lap(u) + dt^2 s``; with damping ``gamma`` the centred
    ``(1 + gamma dt/2) u_next = 2u - (1 - gamma dt/2) u_prev + dt^2 (c^2 lap(u) + s)``), then
    the Mur faces for ``boundary="absorbing"``. ``c``, ``source`` and ``damping`` are floats,
    tensors (broadcast to u), field names of ``params`` or callables ``(params, n)``."""
    if boundary not in _BOUNDARIES:
        raise OperatorError(f"unknown boundary {boundary!r}; choose one of {_BOUNDARIES}")
    u_prev, u = state
    cv = _full(_resolve(c, params, n, "c"), u)
    s = None if source is None else _full(_resolve(source, params, n, "source"), u)
    g = None if damping is None else _full(_resolve(damping, params, n, "damping"), u)
    u_prev = torch.broadcast_to(u_prev.to(u.dtype), u.shape)
    spacing = tuple(float(h) for h in spacing)
    return (u, _WaveStepFn.apply(u_prev, u, cv, s, g, spacing, float(dt), boundary))


def wave_initial_state(u0: torch.Tensor, v0: torch.Tensor | None = None, *, c: Any, dt: float,
                       spacing: Sequence[float], boundary: str = "neumann"
                       ) -> tuple[torch.Tensor, torch.Tensor]:
    """Leapfrog start ``(u_-1, u_0)`` from displacement ``u0`` and velocity ``v0``:
    ``u_-1 = u0 - dt v0 + (c dt)^2 / 2 lap(u0)`` (absorbing uses Neumann ghosts)."""
    bc = "neumann" if boundary == "absorbing" else boundary
    u_prev = u0 + 0.5 * (c * dt) ** 2 * laplacian(u0, spacing, bc)
    if v0 is not None:
        u_prev = u_prev - dt * v0
    return (u_prev, u0)


def stable_dt(spacing: Sequence[float], *, c_max: float | None = None,
              velocity_max: float | Sequence[float] | None = None,
              diffusivity_max: float | None = None, cfl: float = 0.9) -> float:
    """Largest stable explicit step times ``cfl``: the leapfrog bound
    ``1 / (c_max sqrt(sum 1/h^2))`` and the upwind + explicit-Euler bound
    ``1 / (sum |v|/h + 2 D sum 1/h^2)``, whichever are given (the smaller)."""
    inv_h2 = sum(1.0 / float(h) ** 2 for h in spacing)
    limits = []
    if c_max:
        limits.append(1.0 / (float(c_max) * math.sqrt(inv_h2)))
    if velocity_max is not None or diffusivity_max:
        v = 0.0 if velocity_max is None else velocity_max
        vs = [float(v)] * len(spacing) if isinstance(v, (int, float)) else [float(x) for x in v]
        rate = sum(abs(vi) / float(h) for vi, h in zip(vs, spacing))
        rate += 2.0 * float(diffusivity_max or 0.0) * inv_h2
        if rate > 0:
            limits.append(1.0 / rate)
    if not limits:
        raise OperatorError("stable_dt needs c_max, velocity_max and/or diffusivity_max")
    return float(cfl) * min(limits)


def _upsample(fields: Fields, shape: tuple[int, ...], names: Sequence[str]) -> dict:
    """Trailing dims of the named fields brought to ``shape`` by (bi/tri)linear interpolation
    with ``align_corners=False`` (nefi/operators/function.py:45-63 with mode "linear")."""
    d = len(shape)
    out = {}
    for k, v in fields.items():
        if k in names and torch.is_tensor(v) and v.ndim >= d and tuple(v.shape[-d:]) != shape:
            lead = v.shape[:-d]
            mode = {1: "linear", 2: "bilinear", 3: "trilinear"}[d]
            y = F.interpolate(v.reshape(-1, 1, *v.shape[-d:]), size=shape, mode=mode,
                              align_corners=False)
            out[k] = y.reshape(*lead, *shape)
        else:
            out[k] = v
    return out


@register("operator", "time_stepper")
class TimeStepper(Operator):
    """Differentiable time integration of a user-written scheme.

    Args:
        step: ``step(state, params, dt, n) -> state`` (a tensor or a tuple of tensors).
        init: ``init(fields) -> state``.
        n_steps, dt: number of steps and time step.
        observe: ``observe(state, n) -> Tensor | None`` for ``n = 0..n_steps`` (the state after
            n steps); non-None results are stacked along a new leading axis. Default: the final
            state's first component.
        params: ``params(fields) -> dict`` passed to ``step`` (default: the fields).
        field / fields: primary unknown / every field the scheme reads.
        grad_mode: ``"autograd"`` (keep every step) or ``"checkpoint"`` (recompute blocks of
            ``checkpoint_every`` steps in the backward pass).
        homogeneity, native_shape, output_shape: as nefi's operators (coarse fields are
            interpolated to ``native_shape``).
    """

    traceable = False

    def __init__(self, step: StepFn, init: Callable[[Fields], State], n_steps: int, dt: float,
                 observe: Callable[[State, int], torch.Tensor | None] | None = None, *,
                 params: Callable[[Fields], Mapping[str, Any]] | None = None, field: str = "x",
                 fields: Sequence[str] | None = None, grad_mode: str = "autograd",
                 checkpoint_every: int = 16, homogeneity: float | None = None,
                 native_shape: Sequence[int] | None = None,
                 output_shape: Sequence[int] | None = None) -> None:
        super().__init__()
        if grad_mode not in ("autograd", "checkpoint"):
            raise OperatorError(f"grad_mode must be 'autograd' or 'checkpoint', got {grad_mode!r}")
        if n_steps < 0 or checkpoint_every < 1:
            raise OperatorError("n_steps must be >= 0 and checkpoint_every >= 1")
        self.step_fn, self.init_fn = step, init
        self.n_steps, self.dt = int(n_steps), float(dt)
        self.observe_fn, self.params_fn = observe, params
        self.primary = field
        self._fields = tuple(fields) if fields is not None else (field,)
        self.grad_mode = grad_mode
        self.checkpoint_every = int(checkpoint_every)
        self.homogeneity = homogeneity
        self.native_shape = None if native_shape is None else tuple(int(s) for s in native_shape)
        self._output_shape = None if output_shape is None else tuple(int(s) for s in output_shape)

    def _seen(self, state: tuple, single: bool, n: int):
        if self.observe_fn is None:
            return None
        return self.observe_fn(state[0] if single else state, n)

    def _advance(self, n0: int, k: int, single: bool, params, *state):
        """k steps from step index n0; returns the new state followed by its observations."""
        cur = tuple(state)
        seen = []
        for n in range(n0, n0 + k):
            new = self.step_fn(cur[0] if single else cur, params, self.dt, n)
            cur = (new,) if torch.is_tensor(new) else tuple(new)
            o = self._seen(cur, single, n + 1)
            if o is not None:
                seen.append(o)
        return (*cur, *seen)

    def forward(self, fields: Fields) -> torch.Tensor:
        if self.native_shape is not None:
            fields = _upsample(fields, self.native_shape, self._fields)
        missing = [k for k in self._fields if k not in fields]
        if missing:
            raise OperatorError(f"TimeStepper needs fields {missing}; available: {tuple(fields)}")
        params = dict(self.params_fn(fields)) if self.params_fn is not None else dict(fields)
        s0 = self.init_fn(fields)
        single = torch.is_tensor(s0)
        state = (s0,) if single else tuple(s0)
        seen = []
        o = self._seen(state, single, 0)
        if o is not None:
            seen.append(o)
        blocks = self.grad_mode == "checkpoint" and torch.is_grad_enabled()
        n = 0
        while n < self.n_steps:
            k = min(self.checkpoint_every if blocks else self.n_steps, self.n_steps - n)
            if blocks:
                out = checkpoint(self._advance, n, k, single, params, *state, use_reentrant=False)
            else:
                out = self._advance(n, k, single, params, *state)
            state, new = tuple(out[: len(state)]), list(out[len(state):])
            seen.extend(new)
            n += k
        if self.observe_fn is None:
            return state[0]
        if not seen:
            raise OperatorError("observe() returned None for every step; nothing was measured")
        shapes = {tuple(t.shape) for t in seen}
        if len(shapes) > 1:
            raise OperatorError(
                f"observe() returned tensors of different shapes {sorted(shapes)}; return the "
                "same shape at every observed step (or None to skip a step)"
            )
        return torch.stack(seen)

    def at_resolution(self, shape):
        return self

    def output_shape(self, shape):
        return self._output_shape

    def required_fields(self):
        return self._fields

    def extra_repr(self) -> str:
        return (f"n_steps={self.n_steps}, dt={self.dt:.3g}, grad_mode={self.grad_mode}, "
                f"checkpoint_every={self.checkpoint_every}, fields={self._fields}")


__all__ = ["TimeStepper", "laplacian", "stable_dt", "wave_initial_state", "wave_step"]
