"""The LDS-tiled 2-D float64 forward step against the generic float64 kernel and the oracle.

The tiled step (nefi_wave_2d64.hip) runs every forward step of 2-D float64 plans without a PML:
the data generator's path. It must give the generic kernel's traces and gradients to
float64 rounding, and the FP64 oracle's, for every absorbing layer it covers and both orders:
on grids whose width is not a multiple of the 128-column tile and whose height is not a
multiple of its 16 rows, with several shots, uneven receivers and sources near the edges.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from _support import rel
from nefi_amd import _native as nat
from nefi_amd.domain import Domain
from nefi_amd.physics.batched import ShotBatchWaveOperator
from oracle.leapfrog import Problem

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _survey(absorbing, order, n=(45, 290), shots=4, n_t=50):
    rng = np.random.default_rng(21)
    h = 10.0
    ext = ((0.0, n[0] * h), (0.0, n[1] * h))
    geo = []
    for s in range(shots):
        src = [[rng.uniform(0.02, 0.98) * ext[0][1], rng.uniform(0.01, 0.99) * ext[1][1]]]
        rec = [[rng.uniform(h, ext[0][1] - h), rng.uniform(h, ext[1][1] - h)]
               for _ in range(int(rng.integers(3, 25)))]
        geo.append((src, rec))
    c = 1500.0 + 300.0 * rng.random(n)
    kw = dict(n_t=n_t, dt_obs=0.002, f0=15.0, c_max=1800.0, order=order, absorbing=absorbing,
              absorb_width=80.0 if absorbing != "none" else None)
    return Domain(n, ext), geo, c, kw


def _run(op, c, w, fast):
    x = torch.tensor(c, dtype=torch.float64, device=DEV, requires_grad=True)
    op({"c": x.detach()})  # the plan exists after the first call
    plan = next(iter(op._cache.values())).plan
    plan.set_option(nat.OPT_FAST_F64, int(fast))
    assert plan.option(nat.OPT_FAST_F64) == int(fast)
    y = op({"c": x})
    (g,) = torch.autograd.grad((y * w).sum(), x)
    return y.detach().cpu().numpy(), g.cpu().numpy()


@pytest.mark.parametrize("absorbing,order", [("sponge", 4), ("sponge", 2), ("cerjan", 4),
                                             ("cerjan", 2), ("none", 4), ("none", 2)])
def test_tiled_float64_step_matches_generic_and_oracle(absorbing, order):
    dom, geo, c, kw = _survey(absorbing, order)
    op = ShotBatchWaveOperator(dom, geo, grad_mode="checkpoint", checkpoint_every=11, **kw)
    print(f"[case tiled f64 {absorbing} o{order} padded {op.grid.padded_shape}]", flush=True)
    rows = sum(len(r) for _, r in geo)
    w = np.random.default_rng(order).standard_normal((rows, kw["n_t"]))
    wt = torch.tensor(w, device=DEV)
    y_f, g_f = _run(op, c, wt, True)
    y_g, g_g = _run(op, c, wt, False)
    ey, eg = rel(y_f, y_g), rel(g_f, g_g)
    ys, gs, r0 = [], np.zeros(dom.shape), 0
    for src, rec in geo:
        P = Problem(dom.shape, dom.extent, rec, sources=src, **kw)
        ys.append(P.fwi_traces(c)[0])
        gs += P.fwi_gradient(c, w[r0 : r0 + len(rec)][None])
        r0 += len(rec)
    eo, ego = rel(y_f, np.concatenate(ys, 0)), rel(g_f, gs)
    print(f"{absorbing} o{order}: tiled vs generic traces {ey:.1e}, gradient {eg:.1e}; vs FP64 "
          f"oracle traces {eo:.1e}, gradient {ego:.1e}", flush=True)
    assert ey < 1e-13 and eg < 1e-12
    assert eo < 1e-11 and ego < 1e-10


def test_tiled_float64_option_is_refused_where_it_does_not_apply():
    for n, absorbing in (((20, 31), "sponge"), ((20, 30), "pml")):  # padded width 47; PML
        dom = Domain(n, ((0.0, n[0] * 10.0), (0.0, n[1] * 10.0)))
        op = ShotBatchWaveOperator(dom, [([[100.0, 100.0]], [[50.0, 50.0]])], n_t=5,
                                   dt_obs=0.002, f0=15.0, c_max=1800.0, order=2,
                                   absorbing=absorbing, absorb_width=80.0)
        op({"c": torch.full(n, 1500.0, dtype=torch.float64, device=DEV)})
        plan = next(iter(op._cache.values())).plan
        assert plan.option(nat.OPT_FAST_F64) == 0
        with pytest.raises(Exception, match="float64"):
            plan.set_option(nat.OPT_FAST_F64, 1)


@pytest.mark.parametrize("absorbing,order", [("sponge", 2), ("sponge", 4), ("cerjan", 2),
                                             ("none", 4)])
def test_tiled_float64_3d_step_matches_generic_and_oracle(absorbing, order):
    """The 2.5-D tiled float64 step of 3-D plans (the photoacoustic data generator's path):
    width 270 padded (three 128-column chunks, the last partly filled), 17 rows and 13 planes
    (not multiples of the 4-row tile and the 8-plane chunk), a batch of 3 fields."""
    from nefi_amd.physics.wave import WaveInitialConditionOperator

    H = 1e-4
    shape = (7, 11, 264) if absorbing != "none" else (13, 17, 270)
    ext = tuple((0.0, n * H) for n in shape)
    rng = np.random.default_rng(31)
    det = [[rng.uniform(0.1, 0.9) * n * H for n in shape] for _ in range(11)]
    c = 1500.0 * (1.0 + 0.05 * rng.random(shape))
    kw = dict(n_t=30, dt_obs=0.3 * H / 1600.0, c_max=1600.0, order=order, absorbing=absorbing,
              absorb_width=3 * H if absorbing != "none" else None)
    op = WaveInitialConditionOperator(Domain(shape, ext), det, c=torch.tensor(c), **kw)
    print(f"[case tiled f64 3-D {absorbing} o{order} padded {op.grid.padded_shape}]", flush=True)
    assert op.grid.padded_shape[2] == 270
    p0 = rng.standard_normal((3, *shape))
    x = torch.tensor(p0, device=DEV)
    out = {}
    for fast in (True, False):
        with torch.no_grad():
            op({"p0": x})
            plan = next(iter(op._cache.values())).plan
            plan.set_option(nat.OPT_FAST_F64, int(fast))
            assert plan.option(nat.OPT_FAST_F64) == int(fast)
            out[fast] = op({"p0": x}).cpu().numpy()
    e = rel(out[True], out[False])
    P = Problem(shape, ext, det, **kw)
    eo = rel(out[True], P.ic_traces(c, p0))
    print(f"{absorbing} o{order}: tiled vs generic traces {e:.1e}, vs FP64 oracle {eo:.1e}",
          flush=True)
    assert e < 1e-13 and eo < 1e-11
