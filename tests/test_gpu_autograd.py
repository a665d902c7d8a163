"""The autograd contract of the wave operators and the graph cache under many blocks.

nefi differentiates one forward several times: its VJP diagnostics call
``autograd.grad(..., retain_graph=True)`` again and again (nefi/diagnostics/_common.py:211-229)
and GradNorm balancing takes one gradient per loss term (nefi/solve/balancing.py:102); the
reference's non-reentrant checkpointing recomputes on every backward
(nefi/physics/timestep.py:165). Every backward through one forward must give the same gradient.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from nefi_amd import _native as nat
from nefi_amd.domain import Domain
from nefi_amd.physics.batched import ShotBatchWaveOperator
from nefi_amd.physics.wave import WaveOperator

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _survey(absorbing="sponge", n=(70, 300), shots=5, n_t=80, every=13):
    rng = np.random.default_rng(3)
    h = 10.0
    ext = ((0.0, n[0] * h), (0.0, n[1] * h))
    geo = []
    for _ in range(shots):
        src = [[rng.uniform(0.05, 0.95) * ext[0][1], rng.uniform(0.02, 0.3) * ext[1][1]]]
        r = int(rng.integers(3, 40))
        rec = [[rng.uniform(h, ext[0][1] - h), rng.uniform(h, ext[1][1] - h)] for _ in range(r)]
        geo.append((src, rec))
    c = 1500.0 + 300.0 * rng.random(n)
    kw = dict(n_t=n_t, dt_obs=0.002, f0=15.0, c_max=1800.0, order=4, absorbing=absorbing,
              absorb_width=80.0, grad_mode="checkpoint", checkpoint_every=every)
    return Domain(n, ext), geo, c, kw


@pytest.mark.parametrize("absorbing", ["sponge", "pml"])
@pytest.mark.parametrize("overlap", ["0", "1"])
def test_repeated_backward_through_one_forward_is_bit_identical(absorbing, overlap,
                                                                monkeypatch):
    """grad twice with retain_graph on one forward (the first reverse sweep rotates the kept
    blocks), then a third time after another forward reused the workspace: three equal
    gradients, equal to a fresh forward + backward; never zeros."""
    monkeypatch.setenv("NEFI_AMD_OVERLAP", overlap)
    print(f"[case repeated backward {absorbing} overlap {overlap}]", flush=True)
    dom, geo, c, kw = _survey(absorbing)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, generator=torch.Generator(DEV).manual_seed(9))
    x = torch.tensor(c, dtype=torch.float32, device=DEV, requires_grad=True)
    (g_ref,) = torch.autograd.grad((op({"c": x}) * w).sum(), x)
    y = op({"c": x})
    loss = (y * w).sum()
    (g1,) = torch.autograd.grad(loss, x, retain_graph=True)
    (g2,) = torch.autograd.grad(loss, x, retain_graph=True)
    y_other = op({"c": x * 0.99})  # reuses (and overwrites) the workspace
    (g3,) = torch.autograd.grad(loss, x, retain_graph=True)
    (g4,) = torch.autograd.grad((y_other * w).sum(), x)
    torch.cuda.synchronize()
    kept = sum(1 for k in next(iter(op._cache.values())).ws.bufs if k.startswith("keep"))
    print(f"kept blocks {kept}; |g| {float(g1.norm()):.3e}; equal: "
          f"{torch.equal(g1, g_ref)} {torch.equal(g2, g_ref)} {torch.equal(g3, g_ref)}",
          flush=True)
    assert float(g1.norm()) > 0
    for g in (g1, g2, g3):
        assert torch.equal(g, g_ref)
    assert not torch.equal(g4, g_ref) and float(g4.norm()) > 0


def test_repeated_backward_of_the_shared_receiver_operator():
    """The same through WaveOperator's source batches (one plan per batch): backward via
    .backward() twice with retain_graph accumulates exactly twice the gradient."""
    dom, geo, c, kw = _survey("sponge", shots=3)
    src = [s[0] for s, _ in geo]
    rec = geo[0][1]
    op = WaveOperator(dom, src, rec, source_batch=2, **kw)
    w = torch.randn(3, len(rec), kw["n_t"], device=DEV,
                    generator=torch.Generator(DEV).manual_seed(4))
    x = torch.tensor(c, dtype=torch.float32, device=DEV, requires_grad=True)
    loss = (op({"c": x}) * w).sum()
    loss.backward(retain_graph=True)
    g1 = x.grad.clone()
    loss.backward()
    assert float(g1.norm()) > 0
    assert torch.equal(x.grad, 2 * g1)


def test_graph_cache_beyond_its_capacity_matches_eager():
    """A sweep of 200 snapshot blocks replays more distinct calls than the plan keeps graphs
    (256): the least recently launched graph is destroyed once its launch completed. The graph
    count stays at the cap and every replayed gradient equals the eager one; at 40 blocks
    every call stays cached and the count is stable from call to call."""
    for n_t, every, expect_full in ((121, 3, False), (201, 1, True)):
        dom, geo, c, kw = _survey("sponge", n=(40, 88), shots=3, n_t=n_t, every=every)
        op = ShotBatchWaveOperator(dom, geo, **kw)
        rows = sum(len(r) for _, r in geo)
        w = torch.randn(rows, n_t, device=DEV, generator=torch.Generator(DEV).manual_seed(5))
        x = torch.tensor(c, dtype=torch.float32, device=DEV, requires_grad=True)
        op({"c": x})
        plan = next(iter(op._cache.values())).plan
        plan.set_option(nat.OPT_GRAPH, 0)
        (g_e,) = torch.autograd.grad((op({"c": x}) * w).sum(), x)
        plan.set_option(nat.OPT_GRAPH, 1)
        counts = []
        for rep in range(3):
            (g,) = torch.autograd.grad((op({"c": x}) * w).sum(), x)
            torch.cuda.synchronize()
            counts.append(plan.option(nat.OPT_GRAPH_COUNT))
            assert torch.equal(g, g_e), (n_t, every, rep)
        print(f"{-(-(n_t - 1) // every)} blocks: graphs held {counts}", flush=True)
        if expect_full:
            assert counts == [256, 256, 256]
        else:
            assert counts[1] == counts[2] and 0 < counts[2] < 256
