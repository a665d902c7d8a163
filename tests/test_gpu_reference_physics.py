"""The reference's own physics tests, run on the HIP operators.

Each test restates one of nefi's wave tests (tests/test_physics_wave.py, and the 3-D
initial-value checks of tests/test_photoacoustic3d.py) with the same inputs and the same
bounds, calling ``nefi_amd.physics.wave`` instead of ``nefi.physics.wave``:

* the leapfrog's CFL limit;
* the pulse speed;
* energy conservation without absorption and its decay with a PML or sponge;
* absorbing-layer reflection against a run in a 9× larger domain;
* checkpoint = autograd = central differences;
* multiresolution and source batching;
* linearity of the initial-value operator and time-reversal refocusing;
* the exact adjoint of the 3-D initial-value operator.

Inputs are CPU float64 tensors, as in the reference's tests: the operators compute on the GPU
and hand results back on the input's device.
"""

from __future__ import annotations

import math

import pytest
import torch

from nefi_amd.domain import Domain
from nefi_amd.physics.timestep import stable_dt_diffusion, stable_dt_wave
from nefi_amd.physics.wave import (WaveInitialConditionOperator, WaveOperator, laplacian,
                                   time_reversal, wave_energy)

pytestmark = pytest.mark.gpu
F64 = torch.float64


def _peak_time(y: torch.Tensor, dt: float) -> float:
    """Sub-sample peak time by parabolic interpolation through the largest sample."""
    i = int(torch.argmax(y))
    a, b, c = (float(v) for v in y[i - 1 : i + 2])
    return (i + 0.5 * (a - c) / (a - 2 * b + c)) * dt


def test_cfl_helpers_and_leapfrog_stability_limit():
    # tests/test_physics_wave.py:58-73
    print("[case CFL limit]", flush=True)
    h = (0.1, 0.1)
    assert math.isclose(stable_dt_wave(1.0, h, order=2, courant=1.0), 0.1 / math.sqrt(2))
    assert math.isclose(stable_dt_wave(1.0, h, order=4, courant=1.0), 0.1 * math.sqrt(3 / 8))
    assert math.isclose(stable_dt_diffusion(1.0, h, safety=1.0), 0.1**2 / 4)
    p0 = torch.randn(40, 40, dtype=F64, generator=torch.Generator().manual_seed(0)).cuda()
    for factor, bounded in ((0.99, True), (1.05, False)):
        dt = factor * stable_dt_wave(1.0, h, order=4, courant=1.0)
        prev, cur = p0.clone(), p0.clone()
        for _ in range(400):
            prev, cur = cur, 2 * cur - prev + dt**2 * laplacian(cur, h, order=4)
        print(f"dt = {factor} x limit: max |p| after 400 steps {float(cur.abs().max()):.3e}",
              flush=True)
        assert (float(cur.abs().max()) < 1e3) == bounded


@pytest.mark.parametrize("order", [2, 4])
def test_pulse_travels_at_speed_c_1d(order):
    # tests/test_physics_wave.py:79-97
    print(f"[case pulse speed o{order}]", flush=True)
    dom = Domain((400,), ((0.0, 4.0),))
    c0 = 2.0
    op = WaveOperator(dom, [[0.5]], [[1.5], [3.0]], n_t=400, dt_obs=0.004, f0=5.0, c_max=c0,
                      order=order, absorbing="pml", absorb_width=0.5)
    tr = op({"c": torch.full((400,), c0, dtype=F64)})[0]
    speed = 1.5 / (_peak_time(tr[1], op.dt_obs) - _peak_time(tr[0], op.dt_obs))
    print(f"order {order}: speed {speed:.5f} (c = {c0})", flush=True)
    assert abs(speed - c0) / c0 < 0.03


def test_energy_conserved_without_absorption_and_decays_with_it():
    # tests/test_physics_wave.py:100-124
    print("[case energy]", flush=True)
    dom = Domain((40, 40), ((0.0, 1.0), (0.0, 1.0)))
    c = torch.full((40, 40), 1.0, dtype=F64)
    c[20:, :] = 1.5
    kw = dict(n_t=150, dt_obs=0.01, f0=8.0, c_max=1.5, order=4, grad_mode="autograd")
    rec = list(range(80, 1 + 149 * 2, 20))

    def energy(op, crop):
        cc = crop(op.grid.pad(c, "replicate"))

        def obs(state, n):
            return wave_energy(crop(state[0]), crop(state[1]), cc.to(state[0].device),
                               op.grid.spacing, op.dt, 4)

        return torch.stack(op.simulate(c, observe=obs, record=rec)[1]).squeeze(-1)

    closed = WaveOperator(dom, [[0.5, 0.5]], [[0.2, 0.2]], absorbing="none", **kw)
    e_closed = energy(closed, lambda x: x)
    drift = float((e_closed.max() - e_closed.min()) / e_closed.mean())
    print(f"closed box: relative energy drift {drift:.2e}", flush=True)
    assert drift < 1e-9
    for absorbing in ("pml", "sponge"):
        op = WaveOperator(dom, [[0.5, 0.5]], [[0.2, 0.2]], absorbing=absorbing,
                          absorb_width=0.3, **kw)
        e = energy(op, op.grid.crop)
        left = float(e[-1] / e.max())
        print(f"{absorbing}: energy left in the domain {left:.3e}", flush=True)
        assert left < 0.05


@pytest.mark.parametrize("absorbing,width,tol",
                         [("pml", 0.2, 0.01), ("sponge", 0.6, 0.05), ("cerjan", 0.6, 0.05)])
def test_absorbing_boundary_reflection_small(absorbing, width, tol):
    # tests/test_physics_wave.py:127-142
    print(f"[case reflection {absorbing}]", flush=True)
    c0, f0 = 2.0, 5.0
    small = Domain((60, 60), ((0.0, 1.2), (0.0, 1.2)))
    big = Domain((180, 180), ((-1.2, 2.4), (-1.2, 2.4)))
    kw = dict(n_t=220, dt_obs=0.004, f0=f0, c_max=c0, grad_mode="autograd")
    src, rec = [[0.6, 0.8]], [[0.6, 1.0]]
    ref = WaveOperator(big, src, rec, absorbing="none", **kw)
    ref_tr = ref({"c": torch.full(big.shape, c0, dtype=F64)})[0, 0]
    op = WaveOperator(small, src, rec, absorbing=absorbing, absorb_width=width, **kw)
    tr = op({"c": torch.full(small.shape, c0, dtype=F64)})[0, 0]
    refl = float((tr - ref_tr).abs().max() / ref_tr.abs().max())
    print(f"{absorbing}: reflected amplitude {refl:.4f} (bound {tol})", flush=True)
    assert refl < tol


def test_checkpoint_gradients_equal_autograd_and_finite_differences():
    # tests/test_physics_wave.py:145-165
    print("[case gradients: modes and central differences]", flush=True)
    gen = torch.Generator().manual_seed(0)
    dom = Domain((16, 16), ((0.0, 1.0), (0.0, 1.0)))
    c = (2.0 + 0.2 * torch.randn(16, 16, dtype=F64, generator=gen)).clamp(1.6, 2.4)
    kw = dict(n_t=40, dt_obs=0.01, f0=5.0, c_max=2.5, absorb_width=0.2)
    src, rec = [[0.2, 0.3], [0.2, 0.7]], [[0.8, y] for y in (0.2, 0.4, 0.6, 0.8)]
    w = torch.randn(2, 4, 40, dtype=F64, generator=gen)
    grads = []
    for mode, ck in (("autograd", None), ("checkpoint", 5), ("checkpoint", None)):
        op = WaveOperator(dom, src, rec, grad_mode=mode, checkpoint_every=ck, **kw)
        x = c.clone().requires_grad_(True)
        (g,) = torch.autograd.grad((op({"c": x}) * w).sum(), x)
        grads.append(g)
    for g in grads[1:]:
        assert float((g - grads[0]).norm() / grads[0].norm()) < 1e-6
    op = WaveOperator(dom, src, rec, grad_mode="autograd", **kw)
    e = torch.randn(16, 16, dtype=F64, generator=gen)
    eps = 1e-6
    fd = float((op({"c": c + eps * e}) * w).sum()) - float((op({"c": c - eps * e}) * w).sum())
    fd /= 2 * eps
    ad = float((grads[0] * e).sum())
    print(f"central difference {fd:.12e}, adjoint {ad:.12e}", flush=True)
    assert abs(fd - ad) < 1e-6 * max(1.0, abs(fd))


def test_wave_operator_multiresolution_and_source_batching():
    # tests/test_physics_wave.py:168-189
    print("[case multiresolution, source batches]", flush=True)
    dom = Domain((32, 32), ((0.0, 1.0), (0.0, 1.0)))
    src = [[0.1, 0.3], [0.1, 0.5], [0.1, 0.7]]
    rec = [[0.9, y] for y in torch.linspace(0.1, 0.9, 6).tolist()]
    op = WaveOperator(dom, src, rec, n_t=60, dt_obs=0.016, f0=3.0, c_max=2.5, absorb_width=0.25)
    x = dom.physical_coords(dtype=F64)
    c = 2.0 + 0.2 * torch.exp(-((x - 0.5) ** 2).sum(-1) / 0.02)
    y = op({"c": c})
    assert y.shape == (3, 6, 60) == op.output_shape((32, 32))
    coarse = op.at_resolution((16, 16))
    assert coarse is op.at_resolution((16, 16)) and coarse.grid.shape == (16, 16)
    assert coarse.fidelity_tag == op.fidelity_tag
    xc = dom.at((16, 16)).physical_coords(dtype=F64)
    yc = coarse({"c": 2.0 + 0.2 * torch.exp(-((xc - 0.5) ** 2).sum(-1) / 0.02)})
    assert yc.shape == y.shape
    err = float((yc - y).norm() / y.norm())
    print(f"2x coarser grid: relative trace difference {err:.3f}", flush=True)
    assert err < 0.2
    batched = WaveOperator(dom, src, rec, n_t=60, dt_obs=0.016, f0=3.0, c_max=2.5,
                           absorb_width=0.25, source_batch=2)
    assert torch.allclose(batched({"c": c}), y, atol=1e-12)
    with pytest.raises(Exception, match="exceeds"):
        op({"c": torch.full((32, 32), 3.0, dtype=F64)})


def test_initial_condition_operator_is_linear_and_time_reversal_refocuses():
    # tests/test_physics_wave.py:192-214
    print("[case initial-value linearity, time reversal]", flush=True)
    gen = torch.Generator().manual_seed(0)
    n = 48
    dom = Domain((n, n), ((-1.0, 1.0), (-1.0, 1.0)))
    th = 2 * math.pi * torch.arange(64, dtype=F64) / 64
    rec = torch.stack([0.9 * torch.cos(th), 0.9 * torch.sin(th)], 1)
    op = WaveInitialConditionOperator(dom, rec, n_t=200, dt_obs=0.01, c=1.0, absorb_width=0.25)
    assert op.homogeneity == 1.0
    x = dom.physical_coords(dtype=F64)
    src = torch.tensor([0.3, -0.2], dtype=F64)
    p0 = torch.exp(-((x - src) ** 2).sum(-1) / (2 * 0.05**2))
    q0 = torch.randn(n, n, dtype=F64, generator=gen)
    y1, y2 = op({"p0": p0}), op({"p0": q0})
    y3 = op({"p0": 2 * p0 - 3 * q0})
    assert y1.shape == (64, 200)
    lin = float((y3 - (2 * y1 - 3 * y2)).norm() / y3.norm())
    print(f"linearity defect {lin:.2e}", flush=True)
    assert lin < 1e-10
    for mode in ("dirichlet", "adjoint"):
        img = time_reversal(y1, dom, rec, dt_obs=0.01, c=1.0, mode=mode)
        i = int(torch.argmax(img))
        peak = x.reshape(-1, 2)[i]
        corr = float(torch.corrcoef(torch.stack([img.flatten().cpu(), p0.flatten()]))[0, 1])
        off = float((peak - src).norm())
        print(f"{mode}: peak {off:.4f} from the source, correlation {corr:.3f}", flush=True)
        assert off < 3 * dom.spacing()[0]
        assert corr > 0.5


def test_3d_initial_value_operator_is_linear_with_an_exact_adjoint():
    # tests/test_photoacoustic3d.py:49-59 on the operator the instance wraps: a 12 x 12 x 8 grid
    # over 3 x 3 x 2, 16 sensors on the top face, sponge layer, float64
    print("[case 3-D initial-value linearity and adjoint]", flush=True)
    dom = Domain((12, 12, 8), ((0.0, 3.0), (0.0, 3.0), (0.0, 2.0)))
    sensors = [[0.375 + 0.75 * i, 0.375 + 0.75 * j, 0.0] for i in range(4) for j in range(4)]
    op = WaveInitialConditionOperator(dom, sensors, n_t=13, dt_obs=0.25, c=1.0, order=2,
                                      absorbing="sponge", absorb_width=0.5)
    gen = torch.Generator().manual_seed(0)
    a = torch.rand(12, 12, 8, dtype=F64, generator=gen)
    b = torch.rand(12, 12, 8, dtype=F64, generator=gen)
    ya, yb = op({"p0": a}), op({"p0": b})
    assert torch.allclose(op({"p0": 2 * a - 0.5 * b}), 2 * ya - 0.5 * yb, atol=1e-10)
    y = torch.randn(ya.shape, dtype=F64, generator=gen)
    xz = torch.zeros(12, 12, 8, dtype=F64, requires_grad=True)
    (adj,) = torch.autograd.grad((op({"p0": xz}) * y).sum(), xz)
    lhs, rhs = float((ya * y).sum()), float((a * adj).sum())
    print(f"<A a, y> = {lhs:.12e}, <a, A^T y> = {rhs:.12e}", flush=True)
    assert lhs == pytest.approx(rhs, rel=1e-9)
