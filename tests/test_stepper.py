"""The generic wave stepper's oracle pinned without the GPU, and the stepper's host logic.

The oracle (``oracle/stepper.py``) restates nefi's ``wave_step`` / ``laplacian`` / Mur faces; it
is checked against the reference's own stepper test and against exact discrete solutions: at
Courant number 1 the 1-D leapfrog carries travelling waves exactly, so a pulse leaves through
Mur faces without a trace, wraps around a periodic grid unchanged and reflects off Neumann /
Dirichlet walls as the even / odd image about the ghost cell. ``TimeStepper``'s loop (Python,
the step is user code) runs here with a plain torch step; the C ABI's argument checks run
without a GPU.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
import torch

from nefi_amd import _native as nat
from nefi_amd.domain import Domain
from nefi_amd.errors import OperatorError
from nefi_amd.operators.timestepping import TimeStepper, stable_dt
from oracle import stepper as os_


def _pulse(x):
    return np.exp(-0.5 * ((x - 0.0) / 2.5) ** 2)


def test_reference_stepper_test_passes_on_the_oracle():
    # tests/test_operator_zoo.py:196-232: left sensor of the absorbing 1-D run peaks at p0/2
    n = 32
    h = 1.0 / n
    dt = stable_dt((h,), c_max=1.0, cfl=0.5)
    t = (np.arange(n) + 0.5) / n
    p0 = np.exp(-0.5 * ((t - 0.35) / 0.06) ** 2) + 0.6 * np.exp(-0.5 * ((t - 0.7) / 0.08) ** 2)
    up, u = os_.wave_initial_state(p0, c=1.0, dt=dt, spacing=(h,))
    y = os_.run(up, u, 2 * n, dt, (h,), 1.0, "absorbing", observe=lambda s, k: s[[0, -1]])
    assert y.shape == (65, 2)
    assert abs(y[:, 0].max() - 0.5) < 0.03


def _courant1(boundary, direction, n=64, steps=120, x0=20.0):
    i = np.arange(n, dtype=float)
    if direction > 0:  # f(i - t)
        u, up = _pulse(i - x0), _pulse(i + 1 - x0)
    else:  # g(i + t)
        u, up = _pulse(i - x0), _pulse(i - 1 - x0)
    states = [u]
    for _ in range(steps):
        up, u = u, os_.wave_step(up, u, 1.0, (1.0,), 1.0, boundary)
        states.append(u)
    return np.array(states)


def test_mur_faces_absorb_travelling_pulses_exactly_at_courant_one():
    for direction in (1, -1):
        s = _courant1("absorbing", direction)
        assert np.abs(s[-1]).max() < 1e-12, direction  # the pulse left the grid
        k = np.arange(s.shape[0])
        want = _pulse(np.arange(64)[None, :] - 20.0 - direction * k[:, None])
        assert np.abs(s - want).max() < 1e-12  # and was a pure travelling wave until then


def test_periodic_grid_carries_a_pulse_around_unchanged():
    s = _courant1("periodic", 1, n=64, steps=64, x0=30.0)
    assert np.abs(s[64] - s[0]).max() < 1e-12


@pytest.mark.parametrize("boundary,sign,ghost", [("neumann", 1.0, -1.0), ("dirichlet", -1.0, -2.0)])
def test_walls_reflect_as_even_or_odd_images(boundary, sign, ghost):
    # left-moving g(i + t) on i >= 0 with the image sign * g(ghost - i + t): even about -1/2
    # (replicated edge), odd about -1 (zero ghost)
    i = np.arange(64)[None, :].astype(float)
    k = np.arange(-1, 41)[:, None].astype(float)
    want = _pulse(i + k - 12.0) + sign * _pulse(ghost - i + k - 12.0)  # levels -1 .. 40
    up, u = want[0], want[1]
    states = [u]
    for _ in range(40):
        up, u = u, os_.wave_step(up, u, 1.0, (1.0,), 1.0, boundary)
        states.append(u)
    assert np.abs(np.array(states) - want[1:]).max() < 1e-12


@pytest.mark.parametrize("boundary", os_.BOUNDARIES)
def test_oracle_laplacian_is_symmetric_and_exact_on_quadratics(boundary):
    shape, h = (5, 6), (0.5, 0.25)
    N = math.prod(shape)
    A = np.zeros((N, N))
    for j in range(N):
        e = np.zeros(N)
        e[j] = 1.0
        A[:, j] = os_.laplacian(e.reshape(shape), h, boundary).ravel()
    assert np.allclose(A, A.T, atol=1e-12)
    x = np.arange(shape[0])[:, None] * h[0]
    y = np.arange(shape[1])[None, :] * h[1]
    lap = os_.laplacian(x**2 + 3.0 * y**2 + 0 * x * y, h, boundary)
    assert np.allclose(lap[1:-1, 1:-1], 8.0)


def test_stable_dt_bounds():
    assert stable_dt((0.1, 0.2), c_max=3.0, cfl=1.0) == pytest.approx(
        1.0 / (3.0 * math.sqrt(1 / 0.01 + 1 / 0.04)))
    n = 64
    assert stable_dt((1.0 / n,), velocity_max=1.0, diffusivity_max=1e-3, cfl=0.9) == \
        pytest.approx(0.9 / (n + 2e-3 * n**2))
    with pytest.raises(OperatorError):
        stable_dt((0.1,))


def _torch_step(state, params, dt, n):
    """A plain torch leapfrog step (periodic), standing in for user code on the CPU."""
    up, u = state
    lap = (torch.roll(u, 1, -1) - 2 * u + torch.roll(u, -1, -1)) / 0.1**2
    return (u, 2 * u - up + (dt * params["c"]) ** 2 * lap)


def test_time_stepper_loop_blocks_and_errors_on_the_cpu():
    c = torch.full((40,), 1.0, dtype=torch.float64)
    w = torch.randn(31, 3, dtype=torch.float64, generator=torch.Generator().manual_seed(0))
    grads, outs = {}, {}
    for mode in ("autograd", "checkpoint"):
        op = TimeStepper(_torch_step, init=lambda f: (f["p0"], f["p0"]), n_steps=30, dt=0.05,
                         observe=lambda s, k: s[1][[0, 7, 39]], field="p0", fields=("p0", "c"),
                         grad_mode=mode, checkpoint_every=7)
        x = torch.linspace(0, 1, 40, dtype=torch.float64).sin().requires_grad_(True)
        y = op({"p0": x, "c": c})
        (g,) = torch.autograd.grad((y * w).sum(), x)
        grads[mode], outs[mode] = g, y.detach()
    assert outs["autograd"].shape == (31, 3)
    assert torch.allclose(outs["autograd"], outs["checkpoint"], atol=1e-14)
    assert torch.allclose(grads["autograd"], grads["checkpoint"], atol=1e-12)
    final = TimeStepper(_torch_step, init=lambda f: (f["p0"], f["p0"]), n_steps=3, dt=0.05,
                        field="p0", fields=("p0", "c"))
    assert final({"p0": torch.zeros(40), "c": c}).shape == (40,)
    with pytest.raises(OperatorError, match="None for every step"):
        TimeStepper(lambda s, p, dt, k: s, init=lambda f: f["x"], n_steps=3, dt=0.1,
                    observe=lambda s, k: None)({"x": torch.zeros(3)})
    with pytest.raises(OperatorError, match="needs fields"):
        final({"p0": torch.zeros(40)})
    with pytest.raises(OperatorError):
        TimeStepper(_torch_step, init=lambda f: f["x"], n_steps=3, dt=0.1, grad_mode="adjoint")
    with pytest.raises(OperatorError, match="different shapes"):
        TimeStepper(lambda s, p, dt, k: s, init=lambda f: f["x"], n_steps=2, dt=0.1,
                    observe=lambda s, k: s[: k + 1])({"x": torch.zeros(3)})
    up = TimeStepper(lambda s, p, dt, k: s, init=lambda f: f["u0"], n_steps=1, dt=0.1,
                     field="u0", native_shape=(32,))
    assert up({"u0": torch.rand(16)}).shape == (32,)
    assert Domain.unit((8,)).spacing() == (0.125,)


def test_stepper_abi_rejects_bad_grids_without_a_gpu():
    lib = nat.lib()
    g = nat.NwTsGrid()
    g.ndim, g.dtype, g.boundary, g.batch = 4, 0, 0, 1
    assert lib.nw_ts_laplacian(ctypes.byref(g), None, None, None) == 1
    assert b"ndim" in lib.nw_last_error()
    g.ndim = 1
    g.shape[0], g.shape[1], g.shape[2] = 1, 1, 8
    g.h[2] = 0.0
    assert lib.nw_ts_laplacian(ctypes.byref(g), None, None, None) == 1
    g.h[2] = 0.1
    g.boundary = 9
    assert lib.nw_ts_laplacian(ctypes.byref(g), None, None, None) == 1
    g.boundary = nat.NW_TS_ABSORBING
    g.shape[1] = 2  # an unused leading extent other than 1
    assert lib.nw_ts_mur(ctypes.byref(g), 2, 0.1, None, None, None, None) == 1
    g.shape[1] = 1
    assert lib.nw_ts_mur(ctypes.byref(g), 1, 0.1, 1, 1, 1, None) == 1  # not a grid axis
    assert lib.nw_ts_wave_step(ctypes.byref(g), 0.1, None, None, None, None, None, None, 1,
                               None) == 1
