"""Time-loop helpers of the wave path (``nefi/physics/timestep.py:73-225``).

``run_timestepping`` is the generic explicit integrator of the reference API: it drives a
user-supplied Python ``step_fn``. The wave operators of this package do not use it (their loop
runs inside the HIP library); it stays for callers that integrate their own step functions.
The CFL helpers give the same closed forms as the reference:
``dt_max = courant · 2 / (c_max · sqrt(λ_max))`` with ``λ_max = Σ_a ρ_order / h_a²``,
``ρ_2 = 4`` and ``ρ_4 = 16/3``.
"""

from __future__ import annotations

import math
from collections.abc import Callable, Iterable, Sequence

import torch
from torch.utils.checkpoint import checkpoint

from ..errors import ConfigError

State = tuple[torch.Tensor, ...]

__all__ = [
    "State",
    "auto_checkpoint_every",
    "laplacian_eig_max",
    "run_timestepping",
    "stable_dt_diffusion",
    "stable_dt_wave",
]

# max |symbol|·h² of the 1-D centred second-difference stencils (order 2 and 4)
_RHO = {2: 4.0, 4: 16.0 / 3.0}


def auto_checkpoint_every(n_steps: int) -> int:
    """``ceil(sqrt(n_steps))``: the block length minimising ``n/k + k`` stored states."""
    return max(1, math.ceil(math.sqrt(max(1, int(n_steps)))))


def laplacian_eig_max(spacing: Sequence[float], order: int = 2) -> float:
    if order not in _RHO:
        raise ConfigError(f"Laplacian order must be 2 or 4, got {order}")
    return sum(_RHO[order] / float(h) ** 2 for h in spacing)


def stable_dt_wave(c_max: float, spacing: Sequence[float], order: int = 2,
                   courant: float = 0.9) -> float:
    """Largest stable leapfrog step ``courant · 2 / (c_max sqrt(λ_max))``."""
    if not 0.0 < courant <= 1.0:
        raise ConfigError(f"courant safety factor must be in (0, 1], got {courant}")
    if c_max <= 0:
        raise ConfigError(f"c_max must be positive, got {c_max}")
    return float(courant) * 2.0 / (float(c_max) * math.sqrt(laplacian_eig_max(spacing, order)))


def stable_dt_diffusion(d_max: float, spacing: Sequence[float], rate_max: float = 0.0,
                        safety: float = 0.9) -> float:
    """Largest stable explicit-Euler step of ``u_t = D Δu − r u`` (second-order Laplacian)."""
    if not 0.0 < safety <= 1.0:
        raise ConfigError(f"safety factor must be in (0, 1], got {safety}")
    lam = float(d_max) * laplacian_eig_max(spacing, 2) + max(0.0, float(rate_max))
    if lam <= 0:
        raise ConfigError("stable_dt_diffusion needs a positive diffusivity or rate")
    return float(safety) * 2.0 / lam


def _tuple(x) -> State:
    return (x,) if torch.is_tensor(x) else tuple(x)


def run_timestepping(
    step_fn: Callable,
    state0,
    params,
    n_steps: int,
    dt: float,
    observe: Callable | None = None,
    *,
    checkpoint_every: int | None = None,
    record: Iterable[int] | None = None,
) -> tuple[State, list[torch.Tensor]]:
    """Integrate ``state_{n+1} = step_fn(state_n, params, n, dt)`` for ``n_steps`` steps.

    ``observe(state, n)`` is called for every ``n`` in ``record`` (default ``1..n_steps``), with
    ``n = 0`` meaning the initial state; ``None`` results are dropped. With ``checkpoint_every``
    set and gradients recorded, blocks of that many steps are recomputed in the backward pass.
    Returns ``(final_state, observations)``.
    """
    n_steps = int(n_steps)
    if n_steps < 0:
        raise ConfigError(f"n_steps must be >= 0, got {n_steps}")
    state = _tuple(state0)
    prm = () if params is None else _tuple(params)
    if record is None:
        wanted = set(range(1, n_steps + 1))
    else:
        wanted = {int(n) for n in record}
        outside = sorted(n for n in wanted if n < 0 or n > n_steps)
        if outside:
            raise ConfigError(f"record indices {outside} outside [0, {n_steps}]")

    def look(st: State, n: int, into: list) -> None:
        if observe is not None and n in wanted:
            o = observe(st, n)
            if o is not None:
                into.append(o)

    obs: list[torch.Tensor] = []
    look(state, 0, obs)
    grad_on = torch.is_grad_enabled() and any(t.requires_grad for t in (*prm, *state))
    if not (checkpoint_every and grad_on) or int(checkpoint_every) >= n_steps:
        for n in range(n_steps):
            state = _tuple(step_fn(state, prm, n, dt))
            look(state, n + 1, obs)
        return state, obs

    k = max(1, int(checkpoint_every))
    width = len(state)
    for n0 in range(0, n_steps, k):
        n1 = min(n_steps, n0 + k)

        def segment(*args, n0=n0, n1=n1):
            st = tuple(args[:width])
            pp = tuple(args[width:])
            seen: list[torch.Tensor] = []
            for n in range(n0, n1):
                st = _tuple(step_fn(st, pp, n, dt))
                look(st, n + 1, seen)
            return (*st, *seen)

        out = checkpoint(segment, *state, *prm, use_reentrant=False)
        state = tuple(out[:width])
        obs.extend(out[width:])
    return state, obs
