"""The HIP wave stepper against the FP64 oracle, and its hand-written adjoint.

``wave_step`` / ``laplacian`` (``nw_ts_*``) against ``oracle/stepper.py`` over 1-D / 2-D / 3-D
grids with batches, every boundary, heterogeneous speed, sources and damping (float64 to
rounding, float32 within 1e-6); the reference's own ``TimeStepper`` test (checkpoint = autograd,
d'Alembert sensor peak) on the HIP step; the adjoint's gradients in the initial state, speed,
damping and source against autograd through a float64 torch restatement of the reference's
step and against central differences; exact absorption at Courant number 1.
"""

from __future__ import annotations

from functools import partial

import numpy as np
import pytest
import torch

from _support import rel
from nefi_amd.domain import Domain
from nefi_amd.operators.timestepping import (TimeStepper, laplacian, stable_dt,
                                             wave_initial_state, wave_step)
from oracle import stepper as os_

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
CASES = [((3, 50), (0.1,)), ((2, 20, 17), (0.1, 0.12)), ((2, 9, 8, 7), (0.1, 0.11, 0.09))]


@pytest.mark.parametrize("boundary", os_.BOUNDARIES)
@pytest.mark.parametrize("shape,spacing", CASES)
def test_laplacian_and_step_match_oracle(shape, spacing, boundary):
    print(f"[case stepper {shape} {boundary}]", flush=True)
    rng = np.random.default_rng(len(shape) * 10 + len(boundary))
    u, up = rng.standard_normal((2, *shape))
    c = 1.0 + 0.3 * rng.random(shape)
    s = rng.standard_normal(shape)
    gam = 2.0 * rng.random(shape)
    dt = 0.8 * stable_dt(spacing, c_max=1.3)
    lap_want = os_.laplacian(u, spacing, boundary)
    for dtype, tol in ((torch.float64, 1e-13), (torch.float32, 1e-6)):
        T = partial(torch.tensor, dtype=dtype, device=DEV)
        lap = laplacian(T(u), spacing, boundary)
        assert rel(lap.double().cpu().numpy(), lap_want) < tol
        for kind in ("plain", "source", "damped"):
            kw = {"source": s} if kind == "source" else ({"damping": gam} if kind == "damped" else {})
            want = os_.wave_step(up, u, dt, spacing, c, boundary, **kw)
            uu = T(u)
            out = wave_step((T(up), uu), {"c": T(c)}, dt, 0, spacing=spacing, c="c",
                            boundary=boundary, **{k: T(v) for k, v in kw.items()})
            assert out[0] is uu
            e = rel(out[1].double().cpu().numpy(), want)
            print(f"{shape} {boundary} {kind} {dtype}: rel {e:.1e}", flush=True)
            assert e < tol


def _reference_setup(mode, n=32, k=8):
    dom = Domain.unit((n,))
    h = dom.spacing()
    dt = stable_dt(h, c_max=1.0, cfl=0.5)
    op = TimeStepper(
        step=partial(wave_step, spacing=h, c=1.0, boundary="absorbing"),
        init=lambda f: wave_initial_state(f["p0"], c=1.0, dt=dt, spacing=h),
        n_steps=2 * n, dt=dt, observe=lambda s, i: s[1][[0, -1]], field="p0", grad_mode=mode,
        checkpoint_every=k, homogeneity=1.0)
    t = dom.physical_coords(dtype=torch.float64)[..., 0].to(DEV)
    p0 = torch.exp(-0.5 * ((t - 0.35) / 0.06) ** 2) + 0.6 * torch.exp(-0.5 * ((t - 0.7) / 0.08) ** 2)
    return op, p0


def test_reference_time_stepper_test_on_the_hip_step():
    # tests/test_operator_zoo.py:218-232
    grads, outs = {}, {}
    w = torch.randn(65, 2, dtype=torch.float64, generator=torch.Generator().manual_seed(0)).to(DEV)
    for mode in ("autograd", "checkpoint"):
        op, p0 = _reference_setup(mode)
        x = p0.clone().requires_grad_(True)
        y = op({"p0": x})
        (g,) = torch.autograd.grad((y * w).sum(), x)
        grads[mode], outs[mode] = g, y.detach()
    assert outs["autograd"].shape == (65, 2)
    assert torch.allclose(outs["autograd"], outs["checkpoint"], atol=1e-12)
    err = float((grads["autograd"] - grads["checkpoint"]).abs().max())
    assert err < 1e-5 * float(grads["autograd"].abs().max())
    assert float(outs["autograd"][:, 0].max()) == pytest.approx(0.5, abs=0.03)
    # the same run on the oracle
    op, p0 = _reference_setup("autograd")
    h = (1.0 / 32,)
    dt = stable_dt(h, c_max=1.0, cfl=0.5)
    up, u = os_.wave_initial_state(p0.cpu().numpy(), c=1.0, dt=dt, spacing=h)
    want = os_.run(up, u, 64, dt, h, 1.0, "absorbing", observe=lambda s, k: s[[0, -1]])
    assert rel(outs["autograd"].cpu().numpy(), want) < 1e-13


def _torch_reference_step(state, params, dt, n, spacing, boundary):
    """The reference's wave_step in plain torch (float64): autograd gives its gradient."""
    up, u = state
    d = len(spacing)
    bc = "neumann" if boundary == "absorbing" else boundary
    lap = torch.zeros_like(u)
    for i, h in enumerate(spacing):
        dim = u.ndim - d + i
        nn = u.shape[dim]
        first, last = u.narrow(dim, 0, 1), u.narrow(dim, nn - 1, 1)
        if bc == "periodic":
            pad = torch.cat([last, u, first], dim)
        elif bc == "neumann":
            pad = torch.cat([first, u, last], dim)
        else:
            pad = torch.cat([torch.zeros_like(first), u, torch.zeros_like(last)], dim)
        lap = lap + (pad.narrow(dim, 0, nn) - 2.0 * u + pad.narrow(dim, 2, nn)) / h**2
    c, s, g = params["c"], params["s"], params["g"]
    acc = c**2 * lap + s
    un = (2.0 * u - (1.0 - 0.5 * g * dt) * up + dt**2 * acc) / (1.0 + 0.5 * g * dt)
    if boundary == "absorbing":
        for i, h in enumerate(spacing):
            dim = u.ndim - d + i
            nn = u.shape[dim]
            c_lo, c_hi = c.narrow(dim, 0, 1), c.narrow(dim, nn - 1, 1)
            k_lo = (c_lo * dt - h) / (c_lo * dt + h)
            k_hi = (c_hi * dt - h) / (c_hi * dt + h)
            lo = u.narrow(dim, 1, 1) + k_lo * (un.narrow(dim, 1, 1) - u.narrow(dim, 0, 1))
            hi = u.narrow(dim, nn - 2, 1) + k_hi * (un.narrow(dim, nn - 2, 1) - u.narrow(dim, nn - 1, 1))
            un = torch.cat([lo, un.narrow(dim, 1, nn - 2), hi], dim)
    return (u, un)


@pytest.mark.parametrize("boundary", os_.BOUNDARIES)
@pytest.mark.parametrize("shape,spacing", [((24, 19), (0.1, 0.12)), ((2, 7, 6, 5), (0.1, 0.1, 0.1))])
def test_adjoint_matches_autograd_of_the_reference_step(shape, spacing, boundary):
    print(f"[case stepper adjoint {shape} {boundary}]", flush=True)
    rng = np.random.default_rng(7)
    T = partial(torch.tensor, dtype=torch.float64, device=DEV)
    fields = {"p0": T(rng.standard_normal(shape)), "c": T(1.0 + 0.3 * rng.random(shape)),
              "s": T(0.1 * rng.standard_normal(shape)), "g": T(1.5 * rng.random(shape))}
    dt = 0.7 * stable_dt(spacing, c_max=1.3)
    n_steps = 25
    w = T(rng.standard_normal((n_steps + 1, *shape)))
    results = {}
    for name, step in (("hip", partial(wave_step, spacing=spacing, c="c", boundary=boundary,
                                       source="s", damping="g")),
                       ("ref", partial(_torch_reference_step, spacing=spacing,
                                       boundary=boundary))):
        op = TimeStepper(step, init=lambda f: (f["p0"], f["p0"]), n_steps=n_steps, dt=dt,
                         observe=lambda s, k: s[1], field="p0", fields=("p0", "c", "s", "g"),
                         grad_mode="checkpoint", checkpoint_every=6)
        x = {k: v.clone().requires_grad_(True) for k, v in fields.items()}
        y = op(x)
        gs = torch.autograd.grad((y * w).sum(), [x[k] for k in ("p0", "c", "s", "g")])
        results[name] = (y.detach(), gs)
    assert rel(results["hip"][0].cpu().numpy(), results["ref"][0].cpu().numpy()) < 1e-13
    for k, gh, gr in zip(("p0", "c", "s", "g"), results["hip"][1], results["ref"][1]):
        e = rel(gh.cpu().numpy(), gr.cpu().numpy())
        print(f"{shape} {boundary} d/d{k}: rel {e:.1e}", flush=True)
        assert e < 1e-11, k
    # central differences of the HIP objective along one direction per field
    op = TimeStepper(partial(wave_step, spacing=spacing, c="c", boundary=boundary, source="s",
                             damping="g"), init=lambda f: (f["p0"], f["p0"]), n_steps=n_steps,
                     dt=dt, observe=lambda s, k: s[1], field="p0", fields=("p0", "c", "s", "g"))
    with torch.no_grad():
        for j, k in enumerate(("p0", "c", "s", "g")):
            e = T(rng.standard_normal(shape))
            eps = 1e-6

            def J(delta):
                f = dict(fields)
                f[k] = fields[k] + delta * e
                return float((op(f) * w).sum())

            fd = (J(eps) - J(-eps)) / (2 * eps)
            ad = float((results["hip"][1][j] * e).sum())
            assert abs(fd - ad) <= 1e-6 * max(abs(fd), 1e-8), (k, fd, ad)


def test_mur_faces_absorb_exactly_at_courant_one_on_the_gpu():
    i = torch.arange(64, dtype=torch.float64, device=DEV)
    pulse = lambda x: torch.exp(-0.5 * (x / 2.5) ** 2)  # noqa: E731
    up, u = pulse(i + 1 - 20.0), pulse(i - 20.0)
    state = (up, u)
    for n in range(120):
        state = wave_step(state, {}, 1.0, n, spacing=(1.0,), c=1.0, boundary="absorbing")
    assert float(state[1].abs().max()) < 1e-12
