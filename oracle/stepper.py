"""FP64 oracle of nefi's generic leapfrog wave stepper (test infrastructure).

Only the tests use this module, as the checker; the product is
``nefi_amd.operators.timestepping`` over the HIP library.

Restated in numpy float64 from ``nefi/operators/timestepping.py``: the ghost-cell padding
(``_pad1``, :60-70), the second-order ``laplacian`` (:73-87), the Mur faces (``_mur``,
:145-164), ``wave_step`` (:198-214), ``wave_initial_state`` (:217-238) and the time loop of
``TimeStepper.forward`` without gradients (:405-441), in the same operation order.

Pinned by (tests/test_stepper.py): the reference's own stepper test
(``tests/test_operator_zoo.py:196-232``: the left sensor of an absorbing 1-D run peaks at
p0/2, d'Alembert); exact discrete solutions at Courant number 1 in 1-D, where the leapfrog
is exact for travelling waves: a pulse leaves through absorbing (Mur, k = 0) faces without a
trace, wraps around a periodic grid unchanged, and reflects off Neumann / Dirichlet walls as
the even / odd image about the ghost cell; symmetry of the Laplacian matrix; the 2-D Laplacian
of a quadratic (exact).
"""

from __future__ import annotations

import numpy as np

BOUNDARIES = ("periodic", "neumann", "dirichlet", "absorbing")


def _pad1(u: np.ndarray, axis: int, boundary: str) -> np.ndarray:
    first = np.take(u, [0], axis=axis)
    last = np.take(u, [u.shape[axis] - 1], axis=axis)
    if boundary == "periodic":
        return np.concatenate([last, u, first], axis=axis)
    if boundary in ("neumann", "absorbing"):
        return np.concatenate([first, u, last], axis=axis)
    if boundary == "dirichlet":
        return np.concatenate([np.zeros_like(first), u, np.zeros_like(last)], axis=axis)
    raise ValueError(boundary)


def laplacian(u: np.ndarray, spacing, boundary: str = "neumann") -> np.ndarray:
    d = len(spacing)
    out = np.zeros_like(u, dtype=np.float64)
    for i, h in enumerate(spacing):
        ax = u.ndim - d + i
        n = u.shape[ax]
        up = _pad1(u, ax, boundary)
        lo = np.take(up, range(0, n), axis=ax)
        hi = np.take(up, range(2, n + 2), axis=ax)
        out = out + (lo - 2.0 * u + hi) / float(h) ** 2
    return out


def _slab(x, axis, i):
    return np.take(x, [i], axis=axis)


def mur(u_next, u, c, dt, spacing):
    d = len(spacing)
    for i, h in enumerate(spacing):
        ax = u.ndim - d + i
        n = u.shape[ax]
        if n < 3:
            continue
        c_lo, c_hi = _slab(c, ax, 0), _slab(c, ax, n - 1)
        k_lo = (c_lo * dt - h) / (c_lo * dt + h)
        k_hi = (c_hi * dt - h) / (c_hi * dt + h)
        lo = _slab(u, ax, 1) + k_lo * (_slab(u_next, ax, 1) - _slab(u, ax, 0))
        hi = _slab(u, ax, n - 2) + k_hi * (_slab(u_next, ax, n - 2) - _slab(u, ax, n - 1))
        u_next = np.concatenate([lo, np.take(u_next, range(1, n - 1), axis=ax), hi], axis=ax)
    return u_next


def wave_step(u_prev, u, dt, spacing, c, boundary="absorbing", source=None, damping=None):
    """(u_prev, u) -> u_next; c, source, damping are arrays broadcastable to u (or None)."""
    c = np.broadcast_to(np.asarray(c, np.float64), u.shape)
    lap = laplacian(u, spacing, "neumann" if boundary == "absorbing" else boundary)
    acc = c**2 * lap
    if source is not None:
        acc = acc + source
    if damping is not None:
        g = damping
        u_next = (2.0 * u - (1.0 - 0.5 * g * dt) * u_prev + dt**2 * acc) / (1.0 + 0.5 * g * dt)
    else:
        u_next = 2.0 * u - u_prev + dt**2 * acc
    if boundary == "absorbing":
        u_next = mur(u_next, u, c, dt, spacing)
    return u_next


def wave_initial_state(u0, v0=None, *, c, dt, spacing, boundary="neumann"):
    bc = "neumann" if boundary == "absorbing" else boundary
    u_prev = u0 + 0.5 * (c * dt) ** 2 * laplacian(u0, spacing, bc)
    if v0 is not None:
        u_prev = u_prev - dt * v0
    return u_prev, u0


def run(u_prev, u, n_steps, dt, spacing, c, boundary="absorbing", source=None, damping=None,
        observe=None):
    """The time loop: observations of the state after n = 0..n_steps steps, stacked."""
    obs = []
    if observe is not None:
        obs.append(observe(u, 0))
    for n in range(n_steps):
        u_prev, u = u, wave_step(u_prev, u, dt, spacing, c, boundary, source, damping)
        if observe is not None:
            obs.append(observe(u, n + 1))
    return np.stack(obs) if observe is not None else u
