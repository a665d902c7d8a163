"""Acoustic wave leapfrog propagator on MI355X: the ``nefi.physics.wave`` API over HIP kernels.

Same public names, signatures, attributes and errors as ``nefi/physics/wave.py`` (exports at
wave.py:80-95), so nefi's instances, solver, losses and curricula run on it unchanged. What
changes is underneath:

* the whole time loop, the fused step (Laplacian, damping, source, receiver gather) and its
  adjoint run in ``libnefi_wave.so`` (hand-written HIP for gfx950), driven through the C ABI of
  ``include/nefi_wave.h``;
* gradients come from a hand-written reverse sweep behind ``torch.autograd.Function`` with
  snapshots every K steps and block recomputation, instead of autograd's taped graph;
* CPU inputs (nefi's data generators pass CPU float64 fields) are staged to the GPU and the
  result copied back; there is no CPU implementation, and without a GPU the operators raise.

Model, discretisation and absorbing layers are nefi's (wave.py:1-60): leapfrog in time,
centred order-2/4 Laplacian with zero values outside the padded grid, replicate-padded sound
speed, Ricker sources injected with transposed multilinear weights / cell volume, receivers
sampled by multilinear interpolation at ``t_j = j·dt_obs``.
"""

from __future__ import annotations

import logging
import math
import os
from collections.abc import Callable, Sequence

import torch
import torch.nn.functional as F

from .. import _native as nat
from ..domain import Domain
from ..errors import ConfigError, OperatorError, ShapeError
from ..operators.base import Fields, Operator
from ..registry import register
from ..utils import resample, shape_tuple
from . import autotune
from ._engine import DevicePlan, FwiFunction, HostPlan, IcFunction, Propagator
from .timestep import auto_checkpoint_every, stable_dt_wave

log = logging.getLogger("nefi")

ABSORBING = ("pml", "sponge", "cerjan", "none")
# "checkpoint" and "autograd" are nefi's names (wave.py:81); both select the hand-written
# adjoint here, with a snapshot every ceil(sqrt(n)) steps or at every step. "adjoint" is the
# name nefi's heat solver uses for its hand adjoint (operators/pde/adjoint.py:52).
GRAD_MODES = ("checkpoint", "autograd", "adjoint")

__all__ = [
    "ABSORBING",
    "GRAD_MODES",
    "WaveGrid",
    "WaveInitialConditionOperator",
    "WaveOperator",
    "laplacian",
    "laplacian_kernel",
    "ricker",
    "stable_dt",
    "time_reversal",
    "wave_energy",
]

_ABS_CODE = {"none": nat.NW_NONE, "sponge": nat.NW_SPONGE, "cerjan": nat.NW_CERJAN,
             "pml": nat.NW_PML}
# per-axis second-difference coefficients (centre, ±1, ±2) of wave.py:121
_COEFFS = {2: (-2.0, 1.0), 4: (-2.5, 4.0 / 3.0, -1.0 / 12.0)}


# --------------------------------------------------------------------------------------------
# wavelet, stability, stencils (host formulas; wave.py:101-168)
# --------------------------------------------------------------------------------------------
def ricker(t: torch.Tensor, f0: float, t0: float | None = None) -> torch.Tensor:
    """Ricker wavelet ``(1 − 2a²)·exp(−a²)`` with ``a = π f0 (t − t0)``, ``t0 = 1.2/f0`` by default."""
    shift = 1.2 / f0 if t0 is None else t0
    a2 = (math.pi * f0 * (t - shift)) ** 2
    return (1.0 - 2.0 * a2) * torch.exp(-a2)


def stable_dt(c_max: float, spacing: Sequence[float], order: int = 4, courant: float = 0.9) -> float:
    """CFL-limited leapfrog step (:func:`nefi_amd.physics.timestep.stable_dt_wave`)."""
    return stable_dt_wave(c_max, spacing, order=order, courant=courant)


def laplacian_kernel(spacing: Sequence[float], order: int, device=None, dtype=None) -> torch.Tensor:
    """The Laplacian stencil as a ``(1, 1, k, …, k)`` array (cross of ``2r+1`` taps per axis)."""
    if order not in _COEFFS:
        raise ConfigError(f"Laplacian order must be 2 or 4, got {order}")
    d = len(spacing)
    if d not in (1, 2, 3):
        raise ShapeError(f"wave stencils support 1-3 spatial dims, got {d}")
    co = _COEFFS[order]
    r = len(co) - 1
    k = torch.zeros((2 * r + 1,) * d, dtype=torch.float64)
    for ax, h in enumerate(spacing):
        for off in range(-r, r + 1):
            at = [r] * d
            at[ax] = r + off
            k[tuple(at)] += co[abs(off)] / float(h) ** 2
    return k.to(device=device, dtype=dtype or torch.get_default_dtype()).view(1, 1, *k.shape)


_STENCIL_PLANS: dict = {}


def _stencil_plan(shape: tuple[int, ...], spacing, order: int, device, dtype) -> DevicePlan:
    """A source-free, receiver-free plan on ``shape`` used only for ``nw_laplacian``."""
    key = (shape, tuple(float(h) for h in spacing), order, str(device), dtype)
    plan = _STENCIL_PLANS.get(key)
    if plan is None:
        d = len(shape)
        lead = (1,) * (3 - d)
        host = HostPlan(
            ndim=d, order=order, absorbing=nat.NW_NONE, dims=lead + tuple(shape),
            n_abs=(0, 0, 0), field_dims=lead + tuple(shape),
            h=(1.0,) * (3 - d) + tuple(float(h) for h in spacing), dt=1.0, n_steps=0, n_t=1,
            substeps=1, n_shots=1,
        )
        plan = DevicePlan(host, device, dtype)
        _STENCIL_PLANS[key] = plan
    return plan


def laplacian(
    p: torch.Tensor,
    spacing: Sequence[float],
    order: int = 2,
    kernel: torch.Tensor | None = None,
) -> torch.Tensor:
    """Centred finite-difference Laplacian over the trailing ``len(spacing)`` dims (HIP kernel).

    Values outside the grid are zero, as in wave.py:144-168. ``kernel`` is accepted for
    signature compatibility; the stencil is always the one of ``order`` and ``spacing``.
    """
    d = len(spacing)
    if order not in _COEFFS:
        raise ConfigError(f"Laplacian order must be 2 or 4, got {order}")
    if d not in (1, 2, 3) or p.ndim < d:
        raise ShapeError(f"cannot apply a {d}-D Laplacian to shape {tuple(p.shape)}")
    if kernel is not None and kernel.shape[-1] != 2 * len(_COEFFS[order]) - 1:
        raise ConfigError("kernel size does not match the Laplacian order")
    dev = nat.compute_device(p)
    work = torch.float64 if p.dtype == torch.float64 else torch.float32
    spatial = tuple(p.shape[-d:])
    x = p.detach().to(device=dev, dtype=work).contiguous()
    count = max(1, x.numel() // max(1, math.prod(spatial)))
    plan = _stencil_plan(spatial, spacing, order, dev, work)
    out = torch.empty_like(x)
    chunk = max(1, 65535 // plan.host.dims[0])
    flat_in = x.view(count, -1)
    flat_out = out.view(count, -1)
    for b0 in range(0, count, chunk):
        b1 = min(count, b0 + chunk)
        nat.check(
            nat.lib().nw_laplacian(plan.handle, flat_in[b0].data_ptr(), flat_out[b0].data_ptr(),
                                   b1 - b0, plan.stream()),
            "nw_laplacian",
        )
    return out.to(device=p.device, dtype=p.dtype)


def wave_energy(
    p_prev: torch.Tensor,
    p: torch.Tensor,
    c: torch.Tensor,
    spacing: Sequence[float],
    dt: float,
    order: int = 2,
) -> torch.Tensor:
    """Discrete leapfrog energy ``½ΔV Σ[(p − p⁻)²/(c²dt²) − p·Δ_h p⁻]`` (wave.py:171-189)."""
    d = len(spacing)
    dv = math.prod(float(h) for h in spacing)
    kin = (p - p_prev) ** 2 / (c**2 * dt**2)
    pot = -p * laplacian(p_prev, spacing, order)
    return 0.5 * dv * (kin + pot).flatten(-d).sum(-1)


# --------------------------------------------------------------------------------------------
# padded grid geometry (wave.py:195-318)
# --------------------------------------------------------------------------------------------
class WaveGrid:
    """Field grid padded by ``n_abs = ceil(W/h − 1e-9)`` absorbing cells on every side."""

    def __init__(self, domain: Domain, absorb_width: float = 0.0) -> None:
        self.domain = domain
        self.spacing = domain.spacing()
        self.d = domain.ndim
        w = float(absorb_width)
        self.n_abs = tuple(int(math.ceil(w / h - 1e-9)) if w > 0 else 0 for h in self.spacing)
        self.shape = tuple(domain.shape)
        self.padded_shape = tuple(n + 2 * a for n, a in zip(self.shape, self.n_abs))
        self.lo_pad = tuple(lo - a * h for (lo, _), a, h in
                            zip(domain.extent, self.n_abs, self.spacing))

    @property
    def cell_volume(self) -> float:
        return math.prod(self.spacing)

    def pad(self, x: torch.Tensor, mode: str = "replicate") -> torch.Tensor:
        """Pad the trailing dims: ``"replicate"`` (material maps) or ``"zeros"``."""
        if not any(self.n_abs):
            return x
        if mode == "zeros":
            widths: list[int] = []
            for a in reversed(self.n_abs):
                widths += [a, a]
            return F.pad(x, widths)
        for ax, a in enumerate(self.n_abs):
            if a == 0:
                continue
            dim = x.ndim - self.d + ax
            n = x.shape[dim]
            idx = torch.clamp(torch.arange(-a, n + a, device=x.device), 0, n - 1)
            x = x.index_select(dim, idx)
        return x

    def crop(self, x: torch.Tensor) -> torch.Tensor:
        sl = [slice(None)] * (x.ndim - self.d)
        sl += [slice(a, a + n) for a, n in zip(self.n_abs, self.shape)]
        return x[tuple(sl)]

    def layer_depth(self, ax: int, faces: bool = False) -> torch.Tensor:
        """Relative depth ξ ∈ [0, 1] into the layer along ``ax`` at centres or faces (float64)."""
        a, n = self.n_abs[ax], self.shape[ax]
        count = n + 2 * a + (1 if faces else 0)
        pos = torch.arange(count, dtype=torch.float64) + (0.0 if faces else 0.5)
        if a == 0:
            return torch.zeros(count, dtype=torch.float64)
        return torch.clamp(torch.maximum(a - pos, pos - (a + n)), min=0.0) / a

    def depth(self) -> torch.Tensor:
        out = []
        for ax in range(self.d):
            view = [1] * self.d
            view[ax] = -1
            out.append(self.layer_depth(ax).view(view).expand(self.padded_shape))
        return torch.stack(out)

    def point_weights(self, points) -> tuple[torch.Tensor, torch.Tensor]:
        """Multilinear stencils ``(index, weight)`` of shape ``(n, 2^d)`` into the flat padded grid."""
        pts = torch.as_tensor(points, dtype=torch.float64)
        if pts.ndim != 2 or pts.shape[1] != self.d:
            raise ShapeError(f"points must have shape (n, {self.d}), got {tuple(pts.shape)}")
        n = pts.shape[0]
        flat = torch.zeros(n, 1, dtype=torch.long)
        wts = torch.ones(n, 1, dtype=torch.float64)
        for ax in range(self.d):
            h, lo, m = self.spacing[ax], self.lo_pad[ax], self.padded_shape[ax]
            u = (pts[:, ax] - lo) / h - 0.5
            if bool((u < -1e-9).any()) or bool((u > m - 1 + 1e-9).any()):
                raise ConfigError(
                    f"point outside the simulation grid along axis {ax}: coordinates "
                    f"{pts[:, ax].tolist()} vs grid [{lo + 0.5 * h}, {lo + (m - 0.5) * h}]"
                )
            u = u.clamp(0.0, m - 1)
            i0 = torch.floor(u).clamp(max=max(m - 2, 0)).long()
            t = u - i0
            pair_i = torch.stack([i0, (i0 + 1).clamp(max=m - 1)], dim=1)
            pair_w = torch.stack([1.0 - t, t], dim=1)
            stride = math.prod(self.padded_shape[ax + 1 :])
            flat = (flat[:, :, None] + stride * pair_i[:, None, :]).reshape(n, -1)
            wts = (wts[:, :, None] * pair_w[:, None, :]).reshape(n, -1)
        return flat, wts

    def nearest_index(self, points) -> torch.Tensor:
        pts = torch.as_tensor(points, dtype=torch.float64)
        flat = torch.zeros(pts.shape[0], dtype=torch.long)
        for ax in range(self.d):
            h, lo, m = self.spacing[ax], self.lo_pad[ax], self.padded_shape[ax]
            i = torch.floor((pts[:, ax] - lo) / h).long().clamp(0, m - 1)
            flat = flat * m + i
        return flat

    def source_density(self, points) -> torch.Tensor:
        """Dense discrete δ-functions ``(n, *padded)`` (weights / cell volume). Host helper."""
        idx, w = self.point_weights(points)
        out = torch.zeros(idx.shape[0], math.prod(self.padded_shape), dtype=torch.float64)
        out.scatter_add_(1, idx, w / self.cell_volume)
        return out.view(idx.shape[0], *self.padded_shape)


def _as_points(x, d: int, name: str) -> torch.Tensor:
    t = torch.as_tensor(x, dtype=torch.float64)
    if t.ndim == 1 and d == 1:
        t = t[:, None]
    if t.ndim != 2 or t.shape[1] != d:
        raise ShapeError(f"{name} must have shape (n, {d}), got {tuple(t.shape)}")
    return t


def _merge_taps(idx: torch.Tensor, val: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Sum coinciding taps of each row (the scatter-add of wave.py:317), keep 2^d columns."""
    out_i = torch.empty_like(idx)
    out_v = torch.zeros_like(val)
    for r in range(idx.shape[0]):
        acc: dict[int, float] = {}
        order: list[int] = []
        for i, v in zip(idx[r].tolist(), val[r].tolist()):
            if i not in acc:
                acc[i] = 0.0
                order.append(i)
            acc[i] += v
        cols = order + [order[0]] * (idx.shape[1] - len(order))
        vals = [acc[i] for i in order] + [0.0] * (idx.shape[1] - len(order))
        out_i[r] = torch.tensor(cols)
        out_v[r] = torch.tensor(vals, dtype=val.dtype)
    # a repeated index with value 0 must not be matched twice by the kernel: push it out of range
    for r in range(idx.shape[0]):
        seen = set()
        for k in range(idx.shape[1]):
            i = int(out_i[r, k])
            if i in seen:
                out_i[r, k] = -1
            seen.add(i)
    return out_i, out_v


# --------------------------------------------------------------------------------------------
# shared leapfrog machinery (wave.py:345-576)
# --------------------------------------------------------------------------------------------
class _LeapfrogBase(Operator):
    """Grid, time axis and absorbing-layer bookkeeping shared by the wave operators."""

    fidelity_tag = "wave-leapfrog"
    traceable = False

    def __init__(
        self,
        domain: Domain,
        receivers,
        *,
        n_t: int,
        dt_obs: float,
        c_max: float,
        order: int = 4,
        absorbing: str = "pml",
        absorb_width: float | None = None,
        absorb_R: float = 1e-3,
        cerjan_gamma: float = 0.3,
        courant: float = 0.9,
        grad_mode: str = "checkpoint",
        checkpoint_every: int | None = None,
    ) -> None:
        super().__init__()
        if absorbing not in ABSORBING:
            raise ConfigError(f"absorbing must be one of {ABSORBING}, got {absorbing!r}")
        if absorbing == "pml" and domain.ndim > 2:
            raise ConfigError("absorbing='pml' supports 1-D/2-D grids; use 'sponge' in 3-D")
        if grad_mode not in GRAD_MODES:
            raise ConfigError(f"grad_mode must be one of {GRAD_MODES}, got {grad_mode!r}")
        if n_t < 1 or dt_obs <= 0:
            raise ConfigError("n_t must be >= 1 and dt_obs > 0")
        if not 0.0 < absorb_R < 1.0:
            raise ConfigError(f"absorb_R must be in (0, 1), got {absorb_R}")
        if order not in _COEFFS:
            raise ConfigError(f"Laplacian order must be 2 or 4, got {order}")
        self.domain = domain
        self.receivers = _as_points(receivers, domain.ndim, "receivers")
        self.n_t, self.dt_obs = int(n_t), float(dt_obs)
        self.c_max = float(c_max)
        self.order = int(order)
        self.absorbing = absorbing
        if absorb_width is None:
            absorb_width = 10.0 * max(domain.spacing()) if absorbing != "none" else 0.0
        self.absorb_width = float(absorb_width) if absorbing != "none" else 0.0
        self.absorb_R = float(absorb_R)
        self.cerjan_gamma = float(cerjan_gamma)
        self.courant = float(courant)
        self.grad_mode = grad_mode
        self.checkpoint_every = checkpoint_every
        self.grid = WaveGrid(domain, self.absorb_width)
        dt_max = stable_dt(self.c_max, self.grid.spacing, self.order, self.courant)
        self.substeps = max(1, int(math.ceil(self.dt_obs / dt_max - 1e-12)))
        self.dt = self.dt_obs / self.substeps
        self.n_steps = (self.n_t - 1) * self.substeps
        self._cache: dict = {}
        self._res: dict = {}
        # receiver stencils are geometry, fixed for the operator's life
        self._rec = self.grid.point_weights(self.receivers)

    @property
    def use_pml(self) -> bool:
        return self.absorbing == "pml" and any(self.grid.n_abs)

    def _kwargs(self) -> dict:
        return {
            "n_t": self.n_t,
            "dt_obs": self.dt_obs,
            "c_max": self.c_max,
            "order": self.order,
            "absorbing": self.absorbing,
            "absorb_width": self.absorb_width,
            "absorb_R": self.absorb_R,
            "cerjan_gamma": self.cerjan_gamma,
            "courant": self.courant,
            "grad_mode": self.grad_mode,
            "checkpoint_every": self.checkpoint_every,
        }

    @property
    def times(self) -> torch.Tensor:
        return torch.arange(self.n_t, dtype=torch.float64) * self.dt_obs

    @property
    def record(self) -> list[int]:
        return [j * self.substeps for j in range(self.n_t)]

    # ---- absorbing profiles (wave.py:431-456) ------------------------------------------------
    def _profile_max(self, ax: int, factor: float) -> float:
        a, h = self.grid.n_abs[ax], self.grid.spacing[ax]
        return factor * self.c_max * math.log(1.0 / self.absorb_R) / (a * h)

    def damping_profile(self) -> torch.Tensor:
        """Sponge rate ``σ = Σ_a σ_max,a ξ_a²``, ``σ_max,a = 3 c_max ln(1/R)/W_a`` (padded, float64)."""
        g = self.grid
        xi = g.depth()
        sig = torch.zeros(g.padded_shape, dtype=torch.float64)
        for ax, a in enumerate(g.n_abs):
            if a:
                sig = sig + self._profile_max(ax, 3.0) * xi[ax] ** 2
        return sig

    def pml_profile(self, ax: int, faces: bool = False) -> torch.Tensor:
        """PML rate ``ζ = ζ_max ξ²`` along ``ax``, ``ζ_max = 3 c_max ln(1/R)/(2W)`` (1-D, float64)."""
        if self.grid.n_abs[ax] == 0:
            return self.grid.layer_depth(ax, faces) * 0.0
        return self._profile_max(ax, 1.5) * self.grid.layer_depth(ax, faces) ** 2

    def _axis_rate(self, ax: int, factor: float) -> torch.Tensor:
        """One axis' term of the sponge/PML rate at cell centres (1-D, float64)."""
        xi = self.grid.layer_depth(ax)
        if self.grid.n_abs[ax] == 0:
            return xi * 0.0
        return self._profile_max(ax, factor) * xi**2

    # ---- host plan (replaces _static, wave.py:459-504) ---------------------------------------
    def _abs_code(self) -> int:
        if self.absorbing == "pml":
            return nat.NW_PML if self.use_pml else nat.NW_NONE
        if self.absorbing == "none" or not any(self.grid.n_abs):
            return nat.NW_NONE
        return _ABS_CODE[self.absorbing]

    def _host_plan(self, n_shots: int) -> HostPlan:
        g = self.grid
        d = g.d
        lead = 3 - d
        code = self._abs_code()
        hp = HostPlan(
            ndim=d, order=self.order, absorbing=code,
            dims=(1,) * lead + tuple(g.padded_shape),
            n_abs=(0,) * lead + tuple(g.n_abs),
            field_dims=(1,) * lead + tuple(g.shape),
            h=(1.0,) * lead + tuple(float(h) for h in g.spacing),
            dt=self.dt, n_steps=self.n_steps, n_t=self.n_t, substeps=self.substeps,
            n_shots=n_shots,
            # wavefield rows start on 128-byte cache lines (32 elements of float32)
            pitch=-(-g.padded_shape[-1] // 32) * 32,
        )
        dt = self.dt
        for ax in range(d):
            k = lead + ax
            if code == nat.NW_SPONGE:
                hp.damp[k] = 0.5 * dt * self._axis_rate(ax, 3.0)
            elif code == nat.NW_CERJAN:
                hp.damp[k] = torch.exp(-((self.cerjan_gamma * g.layer_depth(ax)) ** 2))
            elif code == nat.NW_PML:
                zc = self.pml_profile(ax)
                zf = self.pml_profile(ax, faces=True)
                den = 1.0 + 0.5 * dt * zf
                hp.damp[k] = 0.5 * dt * zc
                hp.pml_zc[k] = zc
                hp.pml_zf[k] = zf
                hp.pml_alpha[k] = (1.0 - 0.5 * dt * zf) / den
                hp.pml_kappa[k] = dt / (den * g.spacing[ax])
        return hp

    def _receiver_tables(self, hp: HostPlan) -> None:
        """Every shot samples the same receivers (wave.py:589)."""
        idx, w = self._rec
        R, S = idx.shape[0], hp.n_shots
        hp.rec_idx, hp.rec_w = idx, w
        hp.rec_first = torch.zeros(S, dtype=torch.int64)
        hp.rec_count = torch.full((S,), R, dtype=torch.int64)
        hp.trace_first = torch.arange(S, dtype=torch.int64) * R
        hp.n_trace_rows = S * R

    def _check_speed(self, c: torch.Tensor) -> None:
        cmax = float(c.detach().max())
        if cmax > self.c_max * (1.0 + 1e-6):
            raise OperatorError(
                f"wave speed max {cmax:.4g} exceeds the c_max={self.c_max:.4g} used for the CFL "
                "time step; raise c_max or bound the field (e.g. a Bounded head)"
            )
        if float(c.detach().min()) <= 0:
            raise OperatorError("wave speed must be positive everywhere")

    def _propagator(self, key, make_host: Callable[[], HostPlan], device, dtype) -> Propagator:
        full = (key, str(device), dtype)
        prop = self._cache.get(full)
        if prop is None:
            prop = Propagator(DevicePlan(make_host(), device, dtype), self.grad_mode,
                              self.checkpoint_every)
            prop.plan.tune_choice = autotune.tune(prop.plan)
            # each nw_forward / nw_adjoint call becomes one hipGraph launch after its first use
            if os.environ.get("NEFI_AMD_GRAPH", "1") != "0":
                prop.plan.set_option(nat.OPT_GRAPH, 1)
            # scheduling of blocked forward launches: 1 work queue for calls without keep
            # buffers (library default), 2 for all of them, 0 one workgroup per tile
            queue = os.environ.get("NEFI_AMD_WORK_QUEUE")
            if queue is not None and prop.plan.option(nat.OPT_WORK_QUEUE) != int(queue):
                prop.plan.set_option(nat.OPT_WORK_QUEUE, int(queue))
            self._cache[full] = prop
        prop.grad_mode = self.grad_mode  # instances switch it after construction
        prop.checkpoint_every = self.checkpoint_every
        return prop

    def _run_observed(self, prop: Propagator, m, state, psi, first_step, observe, record,
                      out_device, out_dtype):
        """Advance the state between record steps and hand each recorded state to ``observe``."""
        P = prop.plan
        h = P.host
        n_steps = h.n_steps
        rec = sorted({int(n) for n in (record if record is not None else range(1, n_steps + 1))})
        bad = [n for n in rec if n < 0 or n > n_steps]
        if bad:
            raise ConfigError(f"record indices {bad} outside [0, {n_steps}]")
        taper = None
        if h.absorbing == nat.NW_CERJAN:
            tap = None
            for ax in range(self.grid.d):
                v = torch.exp(-((self.cerjan_gamma * self.grid.layer_depth(ax)) ** 2))
                view = [1] * self.grid.d
                view[ax] = -1
                v = v.view(view)
                tap = v if tap is None else torch.minimum(tap, v)
            taper = tap.to(device=P.device, dtype=P.dtype)

        def snapshot(n):
            prev = h.wavefield_view(state[0], h.n_shots)
            if taper is not None and n > first_step:
                prev = prev * taper
            out = [prev, h.wavefield_view(state[1], h.n_shots)]
            if P.pml:
                for a, axis in enumerate(h.used_axes):
                    # one shot's face grid: extent + 1 along `axis`, rows P.face_pitch[a] apart
                    # (the faces of the last axis are stored at a pitch past their D + 1)
                    fs = list(self.grid.padded_shape)
                    fs[axis - (3 - self.grid.d)] += 1
                    stored = fs[:-1] + [P.face_pitch[a]]
                    n_face = h.n_shots * math.prod(stored)  # buffers are rounded up past this
                    out.append(psi[a][:n_face].view(h.n_shots, *stored)[..., : fs[-1]])
            return tuple(t.to(device=out_device, dtype=out_dtype) for t in out)

        obs = []
        n = 0
        with P.eager():
            for r in rec:
                if r > n:
                    P.forward(m, state, psi, n, r, first_step=first_step)
                    n = r
                o = observe(snapshot(n), n)
                if o is not None:
                    obs.append(o)
            if n < n_steps:
                P.forward(m, state, psi, n, n_steps, first_step=first_step)
        return snapshot(n_steps), obs

    def extra_repr(self) -> str:
        g = self.grid
        return (
            f"shape={g.shape}, padded={g.padded_shape}, n_t={self.n_t}, dt_obs={self.dt_obs:.4g}, "
            f"dt={self.dt:.4g} (x{self.substeps}), order={self.order}, absorbing={self.absorbing}"
        )


def _work_dtype(t: torch.Tensor) -> torch.dtype:
    return torch.float64 if t.dtype == torch.float64 else torch.float32


# --------------------------------------------------------------------------------------------
# FWI operator (wave.py:582-750)
# --------------------------------------------------------------------------------------------
@register("operator", "wave")
class WaveOperator(_LeapfrogBase):
    """Acoustic FWI forward model: sound speed ``c(x)`` → traces ``(n_sources, n_receivers, n_t)``.

    Arguments and attributes as nefi's ``WaveOperator`` (wave.py:610-655). The traces and the
    gradient with respect to ``c`` are computed by ``libnefi_wave.so``.
    """

    def __init__(
        self,
        domain: Domain,
        sources,
        receivers,
        *,
        n_t: int,
        dt_obs: float,
        f0: float,
        c_max: float,
        field: str = "c",
        t0: float | None = None,
        amplitude: float = 1.0,
        order: int = 4,
        absorbing: str = "pml",
        absorb_width: float | None = None,
        absorb_R: float = 1e-3,
        cerjan_gamma: float = 0.3,
        courant: float = 0.9,
        grad_mode: str = "checkpoint",
        checkpoint_every: int | None = None,
        source_batch: int | None = None,
    ) -> None:
        super().__init__(
            domain, receivers, n_t=n_t, dt_obs=dt_obs, c_max=c_max, order=order,
            absorbing=absorbing, absorb_width=absorb_width, absorb_R=absorb_R,
            cerjan_gamma=cerjan_gamma, courant=courant, grad_mode=grad_mode,
            checkpoint_every=checkpoint_every,
        )
        self.primary = field
        self.homogeneity = None
        self.sources = _as_points(sources, domain.ndim, "sources")
        self.f0 = float(f0)
        self.t0 = 1.2 / self.f0 if t0 is None else float(t0)
        self.amplitude = float(amplitude)
        self.source_batch = source_batch
        self.fidelity_tag = f"wave-leapfrog-o{self.order}-{self.absorbing}"
        self._src = self.grid.point_weights(self.sources)

    @property
    def n_sources(self) -> int:
        return int(self.sources.shape[0])

    @property
    def n_receivers(self) -> int:
        return int(self.receivers.shape[0])

    def wavelet(self) -> torch.Tensor:
        """Source time function at ``t_n = n·dt`` for ``n < n_steps`` (float64)."""
        t = torch.arange(self.n_steps, dtype=torch.float64) * self.dt
        return self.amplitude * ricker(t, self.f0, self.t0)

    def _shot_plan(self, s0: int, s1: int) -> HostPlan:
        hp = self._host_plan(s1 - s0)
        idx, w = self._src
        val = w[s0:s1] / self.grid.cell_volume
        mi, mv = _merge_taps(idx[s0:s1], val)
        hp.src_idx = mi
        hp.src_val = mv * self.dt**2
        hp.wavelet = self.wavelet()
        self._receiver_tables(hp)
        return hp

    def at_resolution(self, shape):
        shape = shape_tuple(shape)
        if shape == tuple(self.domain.shape):
            return self
        if shape not in self._res:
            op = WaveOperator(
                self.domain.at(shape), self.sources, self.receivers, f0=self.f0, t0=self.t0,
                amplitude=self.amplitude, field=self.primary, source_batch=self.source_batch,
                **self._kwargs(),
            )
            op.fidelity_tag = self.fidelity_tag
            self._res[shape] = op
        return self._res[shape]

    def output_shape(self, shape):
        return (self.n_sources, self.n_receivers, self.n_t)

    def _prepare(self, c: torch.Tensor, sources: slice | None):
        if tuple(c.shape) != tuple(self.domain.shape):
            raise ShapeError(f"c has shape {tuple(c.shape)}, operator expects {self.domain.shape}")
        self._check_speed(c)
        s0, s1, _ = (sources or slice(None)).indices(self.n_sources)
        dev = nat.compute_device(c)
        work = _work_dtype(c)
        prop = self._propagator(("shots", s0, s1), lambda: self._shot_plan(s0, s1), dev, work)
        return prop, s0, s1, dev, work

    def _traces(self, c: torch.Tensor, sources: slice | None) -> torch.Tensor:
        prop, s0, s1, dev, work = self._prepare(c, sources)
        x = c.to(device=dev, dtype=work)
        flat = FwiFunction.apply(x, prop)
        out = flat.view(s1 - s0, self.n_receivers, self.n_t)
        return out.to(device=c.device, dtype=c.dtype)

    def simulate(
        self,
        c: torch.Tensor,
        observe: Callable | None = None,
        record: Sequence[int] | None = None,
        sources: slice | None = None,
    ) -> tuple[tuple[torch.Tensor, ...], list[torch.Tensor]]:
        """Run the leapfrog for sound speed ``c`` and return ``(state, obs)`` (wave.py:699-740).

        With the default ``observe`` the observations are the receiver samples at the record
        steps, differentiable with respect to ``c``. A custom ``observe(state, n)`` receives the
        padded state ``(p^{n-1}, p^n[, ψ…])`` at each step of ``record``; that path carries no
        gradient. The returned final state is detached.
        """
        if observe is None:
            tr = self._traces(c, sources)
            prop, *_ = self._prepare(c, sources)
            return (), list(tr.unbind(-1))
        if c.requires_grad and torch.is_grad_enabled():
            raise ConfigError("simulate with a custom observe carries no gradient; detach c")
        prop, s0, s1, dev, work = self._prepare(c, sources)
        P = prop.plan
        m, _ = P.model(c.detach().to(device=dev, dtype=work).contiguous(), False)
        state, psi = P.new_state()
        return self._run_observed(prop, m, state, psi, 0, observe, record, c.device, c.dtype)

    def forward(self, fields: Fields) -> torch.Tensor:
        c = self.get_field(fields)
        n = self.n_sources
        chunk = self.source_batch or n
        outs = [self._traces(c, slice(s0, min(n, s0 + chunk))) for s0 in range(0, n, chunk)]
        return outs[0] if len(outs) == 1 else torch.cat(outs, dim=0)


# --------------------------------------------------------------------------------------------
# photoacoustic (initial-value) operator (wave.py:756-872)
# --------------------------------------------------------------------------------------------
@register("operator", "wave_initial_condition")
class WaveInitialConditionOperator(_LeapfrogBase):
    """Photoacoustic forward model: initial pressure ``p0`` (zero initial velocity) → traces.

    ``p⁰ = pad_zeros(p0)``, ``p⁻¹ = p⁰ + ½ m Δ_h p⁰`` (wave.py:849-850); linear in ``p0``
    (``homogeneity = 1``); output ``(…, n_receivers, n_t)`` keeping leading batch dims.
    """

    def __init__(
        self,
        domain: Domain,
        receivers,
        *,
        n_t: int,
        dt_obs: float,
        c: float | torch.Tensor = 1.0,
        c_max: float | None = None,
        field: str = "p0",
        order: int = 4,
        absorbing: str = "pml",
        absorb_width: float | None = None,
        absorb_R: float = 1e-3,
        cerjan_gamma: float = 0.3,
        courant: float = 0.9,
        grad_mode: str = "checkpoint",
        checkpoint_every: int | None = None,
    ) -> None:
        c_t = torch.as_tensor(c, dtype=torch.float64)
        if c_t.ndim not in (0, domain.ndim):
            raise ShapeError("c must be a scalar or a map on the domain grid")
        if c_t.ndim > 0 and tuple(c_t.shape) != tuple(domain.shape):
            c_t = resample(c_t, domain.shape)
        c_max = float(c_t.max()) if c_max is None else float(c_max)
        super().__init__(
            domain, receivers, n_t=n_t, dt_obs=dt_obs, c_max=c_max, order=order,
            absorbing=absorbing, absorb_width=absorb_width, absorb_R=absorb_R,
            cerjan_gamma=cerjan_gamma, courant=courant, grad_mode=grad_mode,
            checkpoint_every=checkpoint_every,
        )
        self.primary = field
        self.homogeneity = 1.0
        self.c_known = c_t
        self.fidelity_tag = f"wave-leapfrog-ic-o{self.order}-{self.absorbing}"
        self._models: dict = {}

    def at_resolution(self, shape):
        shape = shape_tuple(shape)
        if shape == tuple(self.domain.shape):
            return self
        if shape not in self._res:
            c = self.c_known
            if c.ndim > 0:
                c = resample(c, shape)
            op = WaveInitialConditionOperator(
                self.domain.at(shape), self.receivers, c=c, field=self.primary, **self._kwargs()
            )
            op.fidelity_tag = self.fidelity_tag
            self._res[shape] = op
        return self._res[shape]

    def output_shape(self, shape):
        return (int(self.receivers.shape[0]), self.n_t)

    def _batch_plan(self, batch: int) -> HostPlan:
        hp = self._host_plan(batch)
        self._receiver_tables(hp)
        return hp

    def _setup(self, p0: torch.Tensor):
        d = self.grid.d
        if tuple(p0.shape[p0.ndim - d :]) != tuple(self.domain.shape):
            raise ShapeError(
                f"p0 has shape {tuple(p0.shape)}, operator expects (..., {self.domain.shape})"
            )
        lead = tuple(p0.shape[: p0.ndim - d])
        batch = math.prod(lead) if lead else 1
        dev = nat.compute_device(p0)
        work = _work_dtype(p0)
        prop = self._propagator(("batch", batch), lambda: self._batch_plan(batch), dev, work)
        key = (str(dev), work)
        if key not in self._models:
            c = self.c_known
            if c.ndim == 0:
                c = c.expand(self.domain.shape)
            cd = c.to(device=dev, dtype=work).contiguous()
            self._models[key] = prop.plan.model(cd, True)
        m, ma = self._models[key]
        return prop, lead, batch, dev, work, m, ma

    def simulate(self, p0: torch.Tensor, observe=None, record=None):
        """Leapfrog run from ``p0``; see :meth:`WaveOperator.simulate`."""
        if observe is None:
            y = self.forward({self.primary: p0})
            return (), list(y.unbind(-1))
        if p0.requires_grad and torch.is_grad_enabled():
            raise ConfigError("simulate with a custom observe carries no gradient; detach p0")
        prop, lead, batch, dev, work, m, ma = self._setup(p0)
        P = prop.plan
        state, psi = P.new_state()
        x = p0.detach().to(device=dev, dtype=work).contiguous()
        nat.check(
            nat.lib().nw_ic_state(P.handle, m.data_ptr(), x.data_ptr(), state[0].data_ptr(),
                                  state[1].data_ptr(), P.stream()),
            "nw_ic_state",
        )
        return self._run_observed(prop, m, state, psi, 0, observe, record, p0.device, p0.dtype)

    def forward(self, fields: Fields) -> torch.Tensor:
        p0 = self.get_field(fields)
        prop, lead, batch, dev, work, m, ma = self._setup(p0)
        x = p0.to(device=dev, dtype=work)
        flat = IcFunction.apply(x, prop, m, ma)
        R = int(self.receivers.shape[0])
        out = flat.view(*lead, R, self.n_t) if lead else flat.view(R, self.n_t)
        return out.to(device=p0.device, dtype=p0.dtype)


# --------------------------------------------------------------------------------------------
# classical PAT baseline (wave.py:878-968)
# --------------------------------------------------------------------------------------------
@torch.no_grad()
def time_reversal(
    traces: torch.Tensor,
    domain: Domain,
    receivers,
    *,
    dt_obs: float,
    c: float | torch.Tensor = 1.0,
    mode: str = "dirichlet",
    order: int = 4,
    absorbing: str = "pml",
    absorb_width: float | None = None,
    absorb_R: float = 1e-3,
    courant: float = 0.9,
) -> torch.Tensor:
    """Time-reversal reconstruction of ``p0`` from receiver traces ``(n_receivers, n_t)``.

    ``"dirichlet"``: propagate from rest backwards in time, clamping the receivers' nearest
    cells to the linearly interpolated reversed traces after every step (``nw_clamp``).
    ``"adjoint"``: ``Aᵀ d`` through the hand-written reverse sweep.
    """
    if mode not in ("dirichlet", "adjoint"):
        raise ConfigError(f"mode must be 'dirichlet' or 'adjoint', got {mode!r}")
    tr = torch.as_tensor(traces)
    if tr.ndim == 3 and tr.shape[0] == 1:
        tr = tr[0]
    if tr.ndim != 2:
        raise ShapeError(f"traces must be (n_receivers, n_t), got {tuple(tr.shape)}")
    dtype = tr.dtype if tr.is_floating_point() else torch.float32
    tr = tr.to(dtype)
    op = WaveInitialConditionOperator(
        domain, receivers, n_t=tr.shape[-1], dt_obs=dt_obs, c=c, order=order,
        absorbing=absorbing, absorb_width=absorb_width, absorb_R=absorb_R, courant=courant,
        grad_mode="autograd",
    )
    if mode == "adjoint":
        with torch.enable_grad():
            p0 = torch.zeros(domain.shape, device=tr.device, dtype=dtype, requires_grad=True)
            (grad,) = torch.autograd.grad((op({"p0": p0}) * tr).sum(), p0)
        return grad

    g = op.grid
    sub = op.substeps
    n_t = tr.shape[-1]
    probe = torch.zeros(domain.shape, device=tr.device, dtype=dtype)
    prop, lead, batch, dev, work, m, ma = op._setup(probe)
    P = prop.plan
    trd = tr.to(device=dev, dtype=work)
    cells = P.host.pitched(g.nearest_index(op.receivers)).to(device=dev, dtype=torch.int32)

    def reversed_data(n: int) -> torch.Tensor:
        u = (n_t - 1) - n / sub
        j0 = min(max(int(math.floor(u)), 0), n_t - 1)
        j1 = min(j0 + 1, n_t - 1)
        t = u - j0
        return ((1.0 - t) * trd[:, j0] + t * trd[:, j1]).contiguous()

    def clamp(buf: torch.Tensor, n: int) -> None:
        vals = reversed_data(n)
        nat.check(nat.lib().nw_clamp(P.handle, buf.data_ptr(), cells.data_ptr(), vals.data_ptr(),
                                     int(cells.numel()), P.stream()), "nw_clamp")

    state, psi = P.new_state()
    clamp(state[1], 0)
    with P.eager():
        for n in range(op.n_steps):
            P.forward(m, state, psi, n, n + 1)
            clamp(state[1], n + 1)
    out = g.crop(P.host.wavefield_view(state[1], 1)[0])
    return out.to(device=tr.device, dtype=dtype).clone()
