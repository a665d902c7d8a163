"""The fast 3-D float32 kernels (2.5-D tiles) against the generic 3-D kernels and the oracle.

The tiled kernels run the forward steps of 3-D float32 plans without a PML and the reverse
steps without an imaging term (the initial-value operator's sweep). They must give the generic
kernels' traces and gradients, and the FP64 oracle's to float32 accuracy, for every absorbing
layer they cover and both orders: on grids whose width spans two 256-column chunks without
filling the second, whose rows and planes are not multiples of the tile's, with batches small
enough for the 4-row tiles and large enough for the 8-row tiles. The FWI operator runs the
whole gradient on them: tiled forward and recompute sweeps that keep Laplacians, and tiled
reverse steps that image from them.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from _support import rel
from nefi_amd import _native as nat
from nefi_amd.domain import Domain
from nefi_amd.physics.wave import WaveInitialConditionOperator, WaveOperator
from oracle.leapfrog import Problem

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
H = 1e-4


def _ic_case(absorbing, order, shape, n_t=36):
    rng = np.random.default_rng(11)
    ext = tuple((0.0, n * H) for n in shape)
    det = [[rng.uniform(0.1, 0.9) * n * H for n in shape] for _ in range(9)]
    c = 1500.0 * (1.0 + 0.05 * rng.random(shape))
    kw = dict(n_t=n_t, dt_obs=0.3 * H / 1600.0, c_max=1600.0, order=order, absorbing=absorbing,
              absorb_width=3 * H if absorbing != "none" else None)
    return Domain(shape, ext), det, c, kw


def _ic_run(op, p0, w, fast):
    op({"p0": p0.detach()})  # the plan exists after the first call
    plan = next(iter(op._cache.values())).plan
    plan.set_option(nat.OPT_FAST_3D, int(fast))
    assert plan.option(nat.OPT_FAST_3D) == int(fast)
    x = p0.detach().clone().requires_grad_(True)
    y = op({"p0": x})
    (g,) = torch.autograd.grad((y * w).sum(), x)
    return y.detach().cpu().double().numpy(), g.cpu().double().numpy()


@pytest.mark.parametrize("batch", [3, 120])
@pytest.mark.parametrize("absorbing,order", [("sponge", 2), ("sponge", 4), ("cerjan", 4),
                                             ("cerjan", 2), ("none", 2), ("none", 4)])
def test_fast_3d_initial_value_operator_matches_generic(absorbing, order, batch):
    # padded widths 268 (two 256-column chunks), 19 rows, 17 planes
    shape = (11, 13, 262) if absorbing != "none" else (17, 19, 268)
    dom, det, c, kw = _ic_case(absorbing, order, shape)
    op = WaveInitialConditionOperator(dom, det, c=torch.tensor(c), **kw)
    print(f"[case fast 3-D IC {absorbing} o{order} batch {batch} padded {op.grid.padded_shape}]",
          flush=True)
    assert op.grid.padded_shape[2] == 268
    gen = torch.Generator(DEV).manual_seed(order)
    p0 = torch.randn(batch, *shape, device=DEV, generator=gen)
    w = torch.randn(batch, len(det), kw["n_t"], device=DEV, generator=gen)
    y_f, g_f = _ic_run(op, p0, w, True)
    y_g, g_g = _ic_run(op, p0, w, False)
    ey, eg = rel(y_f, y_g), rel(g_f, g_g)
    print(f"{absorbing} o{order} batch {batch}: fast vs generic traces {ey:.2e}, "
          f"gradient {eg:.2e}", flush=True)
    assert ey < 1e-6 and eg < 1e-5


@pytest.mark.parametrize("absorbing", ["sponge", "cerjan"])
def test_fast_3d_initial_value_operator_against_oracle(absorbing):
    shape = (9, 14, 34)  # padded (15, 20, 40)
    dom, det, c, kw = _ic_case(absorbing, 4, shape, n_t=30)
    op = WaveInitialConditionOperator(dom, det, c=torch.tensor(c), **kw)
    print(f"[case fast 3-D IC vs oracle {absorbing} padded {op.grid.padded_shape}]", flush=True)
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal((2, *shape))
    w = rng.standard_normal((2, len(det), kw["n_t"]))
    y, g = _ic_run(op, torch.tensor(p0, dtype=torch.float32, device=DEV),
                   torch.tensor(w, dtype=torch.float32, device=DEV), True)
    P = Problem(shape, dom.extent, det, **kw)
    yo, go = P.ic_traces(c, p0), P.ic_gradient(c, w)
    ey, eg = rel(y, yo), rel(g, go)
    print(f"{absorbing}: fast 3-D vs FP64 oracle traces {ey:.2e}, gradient {eg:.2e}", flush=True)
    assert ey < 1e-5 and eg < 1e-5


@pytest.mark.parametrize("absorbing,order", [("sponge", 4), ("cerjan", 2)])
def test_fast_3d_forward_in_the_fwi_gradient(absorbing, order):
    """Tiled forward and recompute sweeps (checkpoint writes included; the keep sweep writes the
    Laplacian of each kept level) and tiled reverse steps imaging from those kept Laplacians:
    the same traces and speed gradient as the generic kernels (which keep p^n and image from
    its stencil), and the oracle's to float32 accuracy."""
    shape = (12, 10, 262)
    rng = np.random.default_rng(8)
    ext = tuple((0.0, n * 10.0) for n in shape)
    src = [[rng.uniform(0.2, 0.8) * n * 10.0 for n in shape] for _ in range(3)]
    rec = [[rng.uniform(0.1, 0.9) * n * 10.0 for n in shape] for _ in range(7)]
    c = 1500.0 + 300.0 * rng.random(shape)
    kw = dict(n_t=40, dt_obs=0.002, f0=15.0, c_max=1800.0, order=order, absorbing=absorbing,
              absorb_width=30.0)
    op = WaveOperator(Domain(shape, ext), src, rec, grad_mode="checkpoint", checkpoint_every=7,
                      **kw)
    print(f"[case fast 3-D FWI {absorbing} o{order} padded {op.grid.padded_shape}]", flush=True)
    w = torch.randn(len(src), len(rec), kw["n_t"], device=DEV,
                    generator=torch.Generator(DEV).manual_seed(1))
    out = {}
    for fast in (True, False):
        x = torch.tensor(c, dtype=torch.float32, device=DEV, requires_grad=True)
        op({"c": x.detach()})
        plan = next(iter(op._cache.values())).plan
        plan.set_option(nat.OPT_FAST_3D, int(fast))
        assert plan.option(nat.OPT_KEEP_LAP) == int(fast)  # the tiled plan keeps Laplacians
        y = op({"c": x})
        (g,) = torch.autograd.grad((y * w).sum(), x)
        out[fast] = (y.detach().cpu().double().numpy(), g.cpu().double().numpy())
    ey, eg = rel(out[True][0], out[False][0]), rel(out[True][1], out[False][1])
    P = Problem(shape, ext, rec, sources=src, **kw)
    yo = P.fwi_traces(c)
    go = P.fwi_gradient(c, w.cpu().double().numpy())
    eo, ego = rel(out[True][0], yo), rel(out[True][1], go)
    print(f"{absorbing} o{order}: fast vs generic traces {ey:.2e}, gradient {eg:.2e}; "
          f"vs FP64 oracle traces {eo:.2e}, gradient {ego:.2e}", flush=True)
    assert ey < 1e-6 and eg < 1e-5
    assert eo < 1e-5 and ego < 1e-4


def test_fast_3d_option_is_refused_where_it_does_not_apply():
    # a padded width that is not a multiple of 4 (9 + 2 * 3), and a float64 plan
    for width, dtype in ((9, torch.float32), (10, torch.float64)):
        shape = (8, 8, width)
        dom = Domain(shape, tuple((0.0, n * H) for n in shape))
        op = WaveInitialConditionOperator(dom, [[4 * H, 4 * H, 5 * H]], n_t=5,
                                          dt_obs=0.3 * H / 1600, c=1500.0, c_max=1600.0, order=2,
                                          absorbing="sponge", absorb_width=3 * H)
        op({"p0": torch.zeros(shape, device=DEV, dtype=dtype)})
        plan = next(iter(op._cache.values())).plan
        assert plan.option(nat.OPT_FAST_3D) == 0
        with pytest.raises(Exception, match="fast 3-D"):
            plan.set_option(nat.OPT_FAST_3D, 1)


def test_fast_3d_step_gathers_uneven_receivers():
    """The tiled step gathers its own input's receivers in workgroups after its tiles: shots
    with 3, 40 and 70 receivers (up to 560 taps, three gather workgroups per shot) give the
    separate gather kernel's traces."""
    from nefi_amd.physics.batched import ShotBatchWaveOperator

    shape = (10, 12, 34)  # padded (16, 18, 40)
    rng = np.random.default_rng(13)
    ext = tuple((0.0, n * 10.0) for n in shape)
    geo = [([[rng.uniform(0.2, 0.8) * n * 10.0 for n in shape]],
            [[rng.uniform(0.05, 0.95) * n * 10.0 for n in shape] for _ in range(r)])
           for r in (3, 40, 70)]
    c = 1500.0 + 300.0 * rng.random(shape)
    op = ShotBatchWaveOperator(Domain(shape, ext), geo, n_t=30, dt_obs=0.002, f0=15.0,
                               c_max=1800.0, order=2, absorbing="sponge", absorb_width=30.0)
    x = torch.tensor(c, dtype=torch.float32, device=DEV)
    out = {}
    for fast in (True, False):
        op({"c": x})
        next(iter(op._cache.values())).plan.set_option(nat.OPT_FAST_3D, int(fast))
        out[fast] = op({"c": x}).cpu().double().numpy()
    e = rel(out[True], out[False])
    print(f"uneven receivers, fused vs separate gather: traces {e:.2e}", flush=True)
    assert out[True].shape[0] == 113 and e < 1e-6
