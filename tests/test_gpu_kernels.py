"""The fast 2-D kernels against the generic kernels and the FP64 oracle, over tile shapes.

The fast forward and reverse kernels (nefi_wave_2d.hip) must give the generic kernels'
results for every tile choice the autotuner may make (rows per wavefront strip, shots per
workgroup), on grids whose width is not a multiple of the 256-column wavefront span and whose
height is not a multiple of the strip, with several shots, uneven receivers and sources near
the grid edge.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from _support import rel
from nefi_amd.domain import Domain
from nefi_amd.physics.batched import ShotBatchWaveOperator
from oracle.leapfrog import Problem

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _survey(absorbing="sponge", order=4, n=(70, 300), shots=5, n_t=60):
    rng = np.random.default_rng(3)
    h = 10.0
    ext = ((0.0, n[0] * h), (0.0, n[1] * h))
    geo = []
    for s in range(shots):
        src = [[rng.uniform(0.05, 0.95) * ext[0][1], rng.uniform(0.02, 0.3) * ext[1][1]]]
        r = int(rng.integers(3, 40))
        # inside the outermost cell centres, so the unpadded "none" grid holds them too
        rec = [[rng.uniform(h, ext[0][1] - h), rng.uniform(h, ext[1][1] - h)] for _ in range(r)]
        geo.append((src, rec))
    c = 1500.0 + 300.0 * rng.random(n)
    kw = dict(n_t=n_t, dt_obs=0.002, f0=15.0, c_max=1800.0, order=order, absorbing=absorbing,
              absorb_width=80.0 if absorbing != "none" else None, grad_mode="checkpoint",
              checkpoint_every=13)
    return Domain(n, ext), geo, c, kw


def _grad(op, c, w, tiles=None, lds=(0, 0), xcd=1, depth=1, keep_lap=1, depth_adj=1, queue=1):
    from nefi_amd import _native as nat

    x = torch.tensor(c, dtype=torch.float32, device=DEV, requires_grad=True)
    y = op({"c": x})
    if tiles is not None:
        # the plan exists after the first call; set tiles and run again
        prop = next(iter(op._cache.values()))
        for kind, (rows, spw) in enumerate(tiles):
            prop.plan.set_tiles(kind, rows, spw)
        prop.plan.set_option(nat.OPT_LDS_TILES_STEP, lds[0])
        prop.plan.set_option(nat.OPT_LDS_TILES_ADJOINT, lds[1])
        prop.plan.set_option(nat.OPT_XCD_MAP, xcd)
        prop.plan.set_option(nat.OPT_TIME_BLOCK, depth)
        prop.plan.set_option(nat.OPT_KEEP_LAP, keep_lap)
        prop.plan.set_option(nat.OPT_TIME_BLOCK_ADJOINT, depth_adj)
        prop.plan.set_option(nat.OPT_WORK_QUEUE, queue)
        y = op({"c": x})
    (g,) = torch.autograd.grad((y * w).sum(), x)
    return y.detach().cpu().double().numpy(), g.cpu().double().numpy()


@pytest.mark.parametrize("absorbing,order", [("sponge", 4), ("cerjan", 4), ("none", 2),
                                             ("sponge", 2), ("pml", 4), ("pml", 2)])
def test_fast_kernels_match_generic_over_tiles(absorbing, order):
    dom, geo, c, kw = _survey(absorbing, order)
    print(f"[case fast-vs-generic {absorbing} o{order}]", flush=True)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, dtype=torch.float32,
                    generator=torch.Generator(DEV).manual_seed(0))
    y0, g0 = _grad(op, c, w, tiles=[(0, 1), (0, 1)])
    for tiles, lds, xcd in (([(16, 1), (16, 1)], (0, 0), 1), ([(4, 2), (8, 3)], (0, 0), 0),
                            ([(32, 5), (64, 5)], (0, 0), 1), ([(8, 1), (4, 4)], (0, 0), 1),
                            ([(8, 2), (8, 3)], (1, 1), 1), ([(16, 1), (32, 5)], (1, 1), 0),
                            ([(32, 4), (16, 2)], (1, 1), 1)):
        y, g = _grad(op, c, w, tiles=tiles, lds=lds, xcd=xcd)
        ey, eg = rel(y, y0), rel(g, g0)
        print(f"{absorbing} o{order} tiles {tiles} lds {lds} xcd {xcd}: traces {ey:.2e}, "
              f"gradient {eg:.2e}", flush=True)
        assert ey < 1e-6 and eg < 1e-5


@pytest.mark.parametrize("absorbing,order", [("sponge", 4), ("cerjan", 4), ("none", 2),
                                             ("sponge", 2)])
def test_time_blocking_is_bit_identical(absorbing, order):
    """Temporally blocked forward launches (depth 2-4) give exactly the traces and gradient of
    one launch per step: uneven receivers sampled at intermediate levels, checkpoints every 13
    steps (blocks cut short, down to one step), recompute sweeps keeping every level. The grid
    is three column tiles wide, so tiles inside the layer-free box (stepped without damping or
    bounds checks) and tiles at the grid edge both occur."""
    from nefi_amd import _native as nat

    dom, geo, c, kw = _survey(absorbing, order, n=(70, 600))
    print(f"[case time-blocking {absorbing} o{order}]", flush=True)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, dtype=torch.float32,
                    generator=torch.Generator(DEV).manual_seed(6))
    for ty, spw, depth, xcd in ((16, 1, 2, 1), (8, 3, 3, 0), (32, 2, 4, 1), (16, 5, 4, 1),
                                (8, 1, 4, 0), (32, 4, 3, 1)):
        tiles = [(ty, spw), (16, 2)]
        y0, g0 = _grad(op, c, w, tiles=tiles, lds=(1, 1), xcd=xcd, depth=1)
        nat.profile_enable(True)
        y, g = _grad(op, c, w, tiles=tiles, lds=(1, 1), xcd=xcd, depth=depth)
        blocked = nat.profile_units(nat.K_STEPS)
        nat.profile_enable(False)
        ey, eg = rel(y, y0), rel(g, g0)
        print(f"{absorbing} o{order} rows {ty} spw {spw} depth {depth} xcd {xcd}: "
              f"{blocked} blocked steps, traces {ey:.1e}, gradient {eg:.1e}", flush=True)
        assert blocked > 0
        assert np.array_equal(y, y0) and np.array_equal(g, g0)


@pytest.mark.parametrize("absorbing,order,n", [("sponge", 4, (70, 600)), ("cerjan", 4, (90, 84)),
                                               ("none", 2, (54, 612)), ("sponge", 4, (70, 328)),
                                               ("sponge", 2, (38, 200))])
def test_work_queue_is_bit_identical(absorbing, order, n):
    """Blocked forward launches as a work queue (persistent workgroups taking tiles, and single
    shots of the last tiles, from per-XCD queues) give exactly the traces and gradient of one
    workgroup per tile: quiet, boundary and split-column tiles, shot groups that do not divide
    the shots, every blocked tile height and depth. Each gradient runs several queue launches
    back to back, so the counters' reset between launches is exercised too."""
    from nefi_amd import _native as nat

    dom, geo, c, kw = _survey(absorbing, order, n=n, shots=7)
    print(f"[case work queue {absorbing} o{order} {n}]", flush=True)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, dtype=torch.float32,
                    generator=torch.Generator(DEV).manual_seed(31))
    for ty, spw, depth in ((32, 4, 4), (16, 3, 3), (8, 2, 2), (32, 7, 2), (16, 1, 4)):
        tiles = [(ty, spw), (16, 2)]
        y0, g0 = _grad(op, c, w, tiles=tiles, lds=(1, 1), depth=depth, queue=0)
        prop = next(iter(op._cache.values()))
        assert prop.plan.option(nat.OPT_WORK_QUEUE_ITEMS) == 0
        nat.profile_enable(True)
        y, g = _grad(op, c, w, tiles=tiles, lds=(1, 1), depth=depth, queue=2)  # keep sweeps too
        blocked = nat.profile_units(nat.K_STEPS)
        nat.profile_enable(False)
        items = prop.plan.option(nat.OPT_WORK_QUEUE_ITEMS)
        print(f"queue {absorbing} o{order} {n} rows {ty} spw {spw} depth {depth}: {items} items, "
              f"{blocked} blocked steps, traces {rel(y, y0):.1e}, gradient {rel(g, g0):.1e}",
              flush=True)
        assert items > 0 and blocked > 0
        assert np.array_equal(y, y0) and np.array_equal(g, g0)


def test_time_blocking_with_2048_receiver_taps_per_shot():
    """Shots at the tap limit of a blocked launch (512 receivers x 4 taps = 2048, as in the
    64-shot benchmark survey) sample every intermediate level exactly."""
    from nefi_amd import _native as nat

    n, h = (48, 600), 10.0
    dom = Domain(n, ((0.0, n[0] * h), (0.0, n[1] * h)))
    geo = []
    for s in range(3):
        src = [[(s + 1) * 0.25 * n[0] * h, 120.0 + 10.0 * s]]
        r = (512, 300, 511)[s]
        rec = [[(2.0 + 0.3 * (i % 3)) * h, (1.0 + (n[1] - 2.0) * (i + 0.5) / r) * h]
               for i in range(r)]
        geo.append((src, rec))
    c = 1500.0 + 200.0 * np.random.default_rng(11).random(n)
    kw = dict(n_t=40, dt_obs=0.002, f0=15.0, c_max=1700.0, order=4, absorbing="sponge",
              absorb_width=80.0, grad_mode="checkpoint", checkpoint_every=13)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, generator=torch.Generator(DEV).manual_seed(12))
    for ty, depth in ((32, 4), (16, 3), (8, 2), (32, 2)):
        y0, g0 = _grad(op, c, w, tiles=[(ty, 2), (16, 2)], lds=(1, 1), depth=1)
        y, g = _grad(op, c, w, tiles=[(ty, 2), (16, 2)], lds=(1, 1), depth=depth)
        print(f"2048 taps rows {ty} depth {depth}: traces {rel(y, y0):.1e}, "
              f"gradient {rel(g, g0):.1e}", flush=True)
        assert np.array_equal(y, y0) and np.array_equal(g, g0)
    assert nat.OPT_TIME_BLOCK == 6


@pytest.mark.parametrize("absorbing,order,n", [("sponge", 4, (70, 328)), ("cerjan", 4, (90, 84)),
                                               ("none", 2, (54, 612)), ("sponge", 2, (38, 200))])
def test_time_blocking_split_column_is_bit_identical(absorbing, order, n):
    """Grids whose last columns fit in half a wavefront row: blocked launches step them as
    split workgroups (two half-width tiles, row tiles 2j and 2j + 1), with an odd number of
    row tiles for some heights, so the last split workgroup's second tile lies below the grid.
    Sources and receivers fall in the split column too. Traces and gradient equal those of one
    launch per step exactly. Padded widths: 344 and 100 (order 4: one or no full column, then
    96-104 split columns), 612 (order 2: two full columns of 248, then 116 split) and 216
    (order 2: 216 columns do not fit in half a row, so one full column, as before)."""
    from nefi_amd import _native as nat

    rng = np.random.default_rng(21)
    h = 10.0
    ext = ((0.0, n[0] * h), (0.0, n[1] * h))
    geo = []
    for s in range(4):
        # source columns over the whole width, the last shots' in the rightmost 40 cells
        col = rng.uniform(0.05, 0.95) if s < 2 else 1.0 - rng.uniform(5.0, 40.0) / n[1]
        src = [[rng.uniform(0.1, 0.9) * ext[0][1], col * ext[1][1]]]
        r = int(rng.integers(5, 60))
        rec = [[rng.uniform(h, ext[0][1] - h), rng.uniform(h, ext[1][1] - h)] for _ in range(r)]
        geo.append((src, rec))
    c = 1500.0 + 300.0 * rng.random(n)
    kw = dict(n_t=60, dt_obs=0.002, f0=15.0, c_max=1800.0, order=order, absorbing=absorbing,
              absorb_width=80.0 if absorbing != "none" else None, grad_mode="checkpoint",
              checkpoint_every=13)
    print(f"[case split-column {absorbing} o{order} {n}]", flush=True)
    op = ShotBatchWaveOperator(Domain(n, ext), geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, dtype=torch.float32,
                    generator=torch.Generator(DEV).manual_seed(22))
    # the LDS reverse step splits the columns past the last full 256 the same way (344, 100
    # and 612 columns): its gradient against the generic kernels'
    yg, gg = _grad(op, c, w, tiles=[(0, 1), (0, 1)])
    for rows_adj, spw_adj in ((16, 2), (8, 3), (32, 1)):
        y, g = _grad(op, c, w, tiles=[(16, 1), (rows_adj, spw_adj)], lds=(1, 1))
        print(f"split reverse rows {rows_adj} spw {spw_adj}: traces {rel(y, yg):.1e}, "
              f"gradient {rel(g, gg):.1e}", flush=True)
        assert rel(y, yg) < 1e-6 and rel(g, gg) < 1e-5
    for ty, spw, depth, xcd in ((32, 4, 4, 1), (16, 1, 4, 0), (8, 3, 3, 1), (32, 2, 2, 1),
                                (16, 4, 2, 0)):
        tiles = [(ty, spw), (16, 2)]
        y0, g0 = _grad(op, c, w, tiles=tiles, lds=(1, 1), xcd=xcd, depth=1)
        nat.profile_enable(True)
        y, g = _grad(op, c, w, tiles=tiles, lds=(1, 1), xcd=xcd, depth=depth)
        blocked = nat.profile_units(nat.K_STEPS)
        nat.profile_enable(False)
        print(f"split {absorbing} o{order} {n} rows {ty} spw {spw} depth {depth} xcd {xcd}: "
              f"{blocked} blocked steps, traces {rel(y, y0):.1e}, gradient {rel(g, g0):.1e}",
              flush=True)
        assert blocked > 0
        assert np.array_equal(y, y0) and np.array_equal(g, g0)


@pytest.mark.parametrize("absorbing,order", [("sponge", 4), ("cerjan", 4), ("none", 2)])
def test_kept_laplacian_matches_kept_wavefield(absorbing, order):
    """Keep buffers holding the Laplacian of p^n (written by the forward kernels, read pointwise
    by the reverse step) give exactly the gradient of keep buffers holding p^n (the reverse step
    applying the stencil itself): register march and LDS kernels, blocked and one-step sweeps."""
    from nefi_amd import _native as nat

    dom, geo, c, kw = _survey(absorbing, order, n=(70, 600))
    print(f"[case kept-laplacian {absorbing} o{order}]", flush=True)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, dtype=torch.float32,
                    generator=torch.Generator(DEV).manual_seed(7))
    for tiles, lds, depth in (([(16, 2), (16, 2)], (1, 1), 1), ([(8, 1), (8, 3)], (0, 0), 1),
                              ([(32, 4), (8, 2)], (1, 1), 4), ([(16, 2), (16, 5)], (1, 0), 2)):
        y0, g0 = _grad(op, c, w, tiles=tiles, lds=lds, depth=depth, keep_lap=0)
        y, g = _grad(op, c, w, tiles=tiles, lds=lds, depth=depth, keep_lap=1)
        plan = next(iter(op._cache.values())).plan
        used = plan.option(nat.OPT_KEEP_LAP)
        print(f"{absorbing} o{order} tiles {tiles} lds {lds} depth {depth}: kept laplacian "
              f"{used}, traces {rel(y, y0):.1e}, gradient {rel(g, g0):.1e}", flush=True)
        assert used == 1
        assert np.array_equal(y, y0) and np.array_equal(g, g0)


@pytest.mark.parametrize("order", [4, 2])
def test_pml_kept_laplacian_matches_kept_wavefield_and_face_fields(order):
    """PML: keep buffers holding the forward step's Laplacian (face divergences included) give
    exactly the gradient of keep buffers holding p^n and both face-field levels, which the
    reverse step otherwise reads to rebuild that Laplacian; kept Laplacians are the PML
    default."""
    from nefi_amd import _native as nat

    dom, geo, c, kw = _survey("pml", order, n=(70, 600))
    print(f"[case pml kept-laplacian o{order}]", flush=True)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, dtype=torch.float32,
                    generator=torch.Generator(DEV).manual_seed(17))
    x = torch.tensor(c, dtype=torch.float32, device=DEV)
    op({"c": x})
    assert next(iter(op._cache.values())).plan.option(nat.OPT_KEEP_LAP) == 1  # the default
    for tiles, lds in (([(16, 2), (16, 2)], (1, 1)), ([(8, 1), (8, 3)], (1, 1)),
                       ([(32, 5), (4, 5)], (1, 0))):
        y0, g0 = _grad(op, c, w, tiles=tiles, lds=lds, keep_lap=0)
        y, g = _grad(op, c, w, tiles=tiles, lds=lds, keep_lap=1)
        used = next(iter(op._cache.values())).plan.option(nat.OPT_KEEP_LAP)
        print(f"pml o{order} tiles {tiles}: kept laplacian {used}, traces {rel(y, y0):.1e}, "
              f"gradient {rel(g, g0):.1e}", flush=True)
        assert used == 1
        assert np.array_equal(y, y0) and np.array_equal(g, g0)


@pytest.mark.parametrize("absorbing,order", [("sponge", 4), ("cerjan", 4), ("none", 2),
                                             ("sponge", 2)])
def test_blocked_reverse_matches_single_steps(absorbing, order):
    """Temporally blocked reverse launches (depth 2-4: residual injection at intermediate
    levels, imaging from kept Laplacians, quiet and boundary tiles) give the gradient of one
    reverse launch per step; only the float64 imaging sums are added in another order."""
    from nefi_amd import _native as nat

    dom, geo, c, kw = _survey(absorbing, order, n=(70, 600))
    print(f"[case blocked-reverse {absorbing} o{order}]", flush=True)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, dtype=torch.float32,
                    generator=torch.Generator(DEV).manual_seed(8))
    for tiles, depth, depth_adj in (([(16, 2), (16, 2)], 1, 2), ([(32, 4), (8, 3)], 4, 3),
                                    ([(16, 1), (16, 5)], 2, 4), ([(8, 2), (8, 1)], 3, 4)):
        y0, g0 = _grad(op, c, w, tiles=tiles, lds=(1, 1), depth=depth, depth_adj=1)
        nat.profile_enable(True)
        y, g = _grad(op, c, w, tiles=tiles, lds=(1, 1), depth=depth, depth_adj=depth_adj)
        blocked = nat.profile_units(nat.K_ADJOINT_STEPS)
        nat.profile_enable(False)
        eg = rel(g, g0)
        print(f"{absorbing} o{order} tiles {tiles} depth {depth}/{depth_adj}: {blocked} blocked "
              f"reverse steps, traces {rel(y, y0):.1e}, gradient {eg:.1e}", flush=True)
        assert blocked > 0
        assert np.array_equal(y, y0) and eg < 1e-6


def test_blocked_reverse_initial_condition_2d():
    """The initial-value operator's reverse sweep (residual injected at every recorded step, no
    imaging sum) is bit-identical blocked and stepped one by one."""
    from nefi_amd import _native as nat
    from nefi_amd.physics.wave import WaveInitialConditionOperator

    n, h = (64, 320), 10.0
    dom = Domain(n, ((0.0, n[0] * h), (0.0, n[1] * h)))
    rec = [[h * 2.5, h * (4 + 3 * i)] for i in range(90)]
    op = WaveInitialConditionOperator(dom, rec, n_t=70, dt_obs=0.002, c=1600.0, order=4,
                                      absorbing="sponge", absorb_width=80.0)
    rng = np.random.default_rng(9)
    p0 = torch.tensor(rng.standard_normal(n), dtype=torch.float32, device=DEV,
                      requires_grad=True)
    w = torch.randn(len(rec), 70, device=DEV, generator=torch.Generator(DEV).manual_seed(10))
    grads = {}
    for depth_adj in (1, 2, 3, 4):
        op({"p0": p0})  # the plan exists after the first call
        prop = next(iter(op._cache.values()))
        prop.plan.set_tiles(1, 16, 1)
        prop.plan.set_option(nat.OPT_LDS_TILES_ADJOINT, 1)
        prop.plan.set_option(nat.OPT_TIME_BLOCK_ADJOINT, depth_adj)
        nat.profile_enable(True)
        (g,) = torch.autograd.grad((op({"p0": p0}) * w).sum(), p0)
        blocked = nat.profile_units(nat.K_ADJOINT_STEPS)
        nat.profile_enable(False)
        grads[depth_adj] = g.cpu().numpy()
        print(f"IC 2-D reverse depth {depth_adj}: {blocked} blocked steps, gradient vs single "
              f"steps {rel(grads[depth_adj], grads[1]):.1e}", flush=True)
        assert (blocked > 0) == (depth_adj > 1)
        assert np.array_equal(grads[depth_adj], grads[1])


@pytest.mark.parametrize("absorbing", ["sponge", "pml"])
def test_fast_path_against_oracle_uneven_shots(absorbing):
    print(f"[case oracle uneven shots {absorbing}]", flush=True)
    dom, geo, c, kw = _survey(absorbing, 4, n=(40, 88), shots=3, n_t=50)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = np.random.default_rng(4).standard_normal((rows, kw["n_t"]))
    y, g = _grad(op, c, torch.tensor(w, dtype=torch.float32, device=DEV))
    prop = next(iter(op._cache.values()))
    assert prop.plan.tiles(0)[0] > 0 and prop.plan.tiles(1)[0] > 0  # fast kernels in use
    pk = {k: v for k, v in kw.items() if k not in ("grad_mode", "checkpoint_every")}
    ys, gs = [], np.zeros(dom.shape)
    r0 = 0
    for src, rec in geo:
        P = Problem(dom.shape, dom.extent, rec, sources=src, **pk)
        ys.append(P.fwi_traces(c)[0])
        gs += P.fwi_gradient(c, w[r0 : r0 + len(rec)][None])
        r0 += len(rec)
    yo = np.concatenate(ys, 0)
    ey, eg = rel(y, yo), rel(g, gs)
    print(f"{absorbing} uneven shots vs FP64 oracle: traces {ey:.2e}, gradient {eg:.2e}",
          flush=True)
    assert ey < 1e-5 and eg < 1e-4


@pytest.mark.parametrize("order", [4, 2])
def test_pml_quiet_tiles_match_generic_and_oracle(order):
    """PML tiles inside the layer-free box read no face fields (the reverse tiles write none):
    on 576 padded columns whole 256-column tiles lie inside the box, beside tiles that cross the
    layer. Traces and gradients equal the generic kernels' (kept Laplacians and p^n tiles) and
    the FP64 oracle's to float32 accuracy."""
    print(f"[case PML quiet tiles o{order}]", flush=True)
    dom, geo, c, kw = _survey("pml", order, n=(72, 560), shots=3, n_t=60)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = np.random.default_rng(6).standard_normal((rows, kw["n_t"]))
    wt = torch.tensor(w, dtype=torch.float32, device=DEV)
    y0, g0 = _grad(op, c, wt, tiles=[(0, 1), (0, 1)])
    for tiles, keep_lap in (([(16, 1), (16, 1)], 1), ([(8, 2), (8, 3)], 1),
                            ([(16, 3), (8, 1)], 0)):
        y, g = _grad(op, c, wt, tiles=tiles, lds=(1, 1), keep_lap=keep_lap)
        ey, eg = rel(y, y0), rel(g, g0)
        print(f"PML o{order} tiles {tiles} kept Laplacians {keep_lap}: traces {ey:.2e}, "
              f"gradient {eg:.2e} vs generic", flush=True)
        assert ey < 1e-6 and eg < 1e-5
    pk = {k: v for k, v in kw.items() if k not in ("grad_mode", "checkpoint_every")}
    ys, gs, r0 = [], np.zeros(dom.shape), 0
    for src, rec in geo:
        P = Problem(dom.shape, dom.extent, rec, sources=src, **pk)
        ys.append(P.fwi_traces(c)[0])
        gs += P.fwi_gradient(c, w[r0 : r0 + len(rec)][None])
        r0 += len(rec)
    eo, ego = rel(y, np.concatenate(ys, 0)), rel(g, gs)
    print(f"PML o{order} quiet tiles vs FP64 oracle: traces {eo:.2e}, gradient {ego:.2e}",
          flush=True)
    assert eo < 1e-5 and ego < 1e-4


@pytest.mark.parametrize("absorbing", ["sponge", "pml"])
def test_hip_graph_replays_match_eager(absorbing):
    """Every call captured once, then replayed: each replay (with eager work and syncs between
    them) must reproduce the eager traces and gradient exactly (PML: the face fields' buffer
    pairs swap like the wavefields')."""
    from nefi_amd import _native as nat

    print(f"[case graph replays {absorbing}]", flush=True)
    dom, geo, c, kw = _survey(absorbing, 4, n=(40, 88), shots=3, n_t=50)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, generator=torch.Generator(DEV).manual_seed(5))
    x = torch.tensor(c, dtype=torch.float32, device=DEV, requires_grad=True)
    op({"c": x})
    plan = next(iter(op._cache.values())).plan
    plan.set_option(nat.OPT_GRAPH, 0)
    y_e = op({"c": x})
    (g_e,) = torch.autograd.grad((y_e * w).sum(), x)
    plan.set_option(nat.OPT_GRAPH, 1)
    counts = []
    for rep in range(4):
        y = op({"c": x})
        (g,) = torch.autograd.grad((y * w).sum(), x)
        torch.cuda.synchronize()
        _ = torch.randn(1000, device=DEV).sum().item()  # unrelated eager work between replays
        counts.append(plan.option(nat.OPT_GRAPH_COUNT))
        print(f"graph replay {rep}: graphs held {counts[-1]}", flush=True)
        assert torch.equal(y, y_e) and torch.equal(g, g_e), rep
    assert counts[-1] > 0


def test_autograd_comparator_matches_hip_gradient():
    """bench.py's plain-autograd path (the reference's aten op sequence) and the HIP reverse
    sweep give the same gradient: they time the same computation."""
    import bench

    dom, geo, c, kw = _survey("sponge", 4, n=(40, 88), shots=3, n_t=50)
    kw["checkpoint_every"] = 200
    geo = [(s, r) for s, r in geo]
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    res = torch.randn(rows, kw["n_t"], device=DEV, dtype=torch.float64,
                      generator=torch.Generator(DEV).manual_seed(2))
    x = torch.tensor(c, dtype=torch.float64, device=DEV, requires_grad=True)
    (g_hip,) = torch.autograd.grad((op({"c": x}) * res).sum(), x)
    g_ref = bench.autograd_leapfrog(op, x.detach(), 3, op.n_steps, res, block=17)
    e = rel(g_ref.cpu().numpy(), g_hip.cpu().numpy())
    print(f"autograd comparator vs HIP gradient (float64): {e:.2e}", flush=True)
    assert e < 1e-10


@pytest.mark.parametrize("absorbing", ["sponge", "pml"])
def test_overlapped_reverse_sweep_is_bit_identical(absorbing, monkeypatch):
    """The reverse sweep with earlier blocks recomputed on a second stream while the forward
    sweep keeps the last 2, 3 or 4 blocks (as many rotating keep buffers) gives exactly the
    gradient of the one-stream sweep, eager and replayed from graphs (7 blocks of 13 steps)."""
    from nefi_amd import _native as nat

    print(f"[case overlapped reverse sweep {absorbing}]", flush=True)
    dom, geo, c, kw = _survey(absorbing, 4, n=(70, 300), shots=5, n_t=80)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, generator=torch.Generator(DEV).manual_seed(9))
    x = torch.tensor(c, dtype=torch.float32, device=DEV, requires_grad=True)
    grads = []
    for flag, kept in (("0", "4"), ("1", "2"), ("1", "3"), ("1", "4"), ("1", "4")):
        monkeypatch.setenv("NEFI_AMD_OVERLAP", flag)
        monkeypatch.setenv("NEFI_AMD_KEEP_BLOCKS", kept)
        (g,) = torch.autograd.grad((op({"c": x}) * w).sum(), x)
        torch.cuda.synchronize()
        grads.append(g)
        print(f"overlap {flag} kept blocks {kept}: equal {torch.equal(g, grads[0])}", flush=True)
    prop = next(iter(op._cache.values()))
    assert "keep4" in prop.ws.bufs and prop.plan.option(nat.OPT_GRAPH) in (0, 1)
    for g in grads[1:]:
        assert torch.equal(g, grads[0])


def test_overlapped_sweep_with_blocked_reverse_is_bit_identical(monkeypatch):
    """Blocked reverse launches (depth 2, kept Laplacians) rotate through wavefields of their
    own, so the overlapped sweep (blocked recompute on a second stream) runs beside them and
    gives exactly the one-stream gradient."""
    print("[case overlapped sweep, blocked reverse]", flush=True)
    dom, geo, c, kw = _survey("sponge", 4, n=(70, 300), shots=5, n_t=80)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    rows = sum(len(r) for _, r in geo)
    w = torch.randn(rows, kw["n_t"], device=DEV, generator=torch.Generator(DEV).manual_seed(13))
    tiles = [(32, 2), (16, 2)]
    out = []
    for flag, kept in (("0", "4"), ("1", "2"), ("1", "4")):
        monkeypatch.setenv("NEFI_AMD_OVERLAP", flag)
        monkeypatch.setenv("NEFI_AMD_KEEP_BLOCKS", kept)
        out.append(_grad(op, c, w, tiles=tiles, lds=(1, 1), depth=4, keep_lap=1, depth_adj=2))
        print(f"overlap {flag} kept blocks {kept}: traces {rel(out[-1][0], out[0][0]):.1e}, "
              f"gradient {rel(out[-1][1], out[0][1]):.1e}", flush=True)
    for y, g in out[1:]:
        assert np.array_equal(y, out[0][0]) and np.array_equal(g, out[0][1])


def test_misaligned_buffers_are_refused():
    """The tiled kernels move 16-byte vectors: nw_forward / nw_adjoint refuse a buffer that is
    not 16-byte aligned with an error instead of launching."""
    from nefi_amd import _native as nat

    dom, geo, c, kw = _survey("sponge", 4, n=(40, 88), shots=2, n_t=10)
    op = ShotBatchWaveOperator(dom, geo, **kw)
    op({"c": torch.tensor(c, dtype=torch.float32, device=DEV)})
    plan = next(iter(op._cache.values())).plan
    m = plan.zeros(plan.host.cells + 4)
    state, _ = plan.new_state()
    sp = nat.ptr_array([state[0].data_ptr(), state[1].data_ptr()])
    rc = nat.lib().nw_forward(plan.handle, m.data_ptr() + 4, sp, None, 0, 1, None, None, 0, None,
                              0, 0, plan.stream())
    assert rc != 0 and b"16-byte" in nat.lib().nw_last_error()
    lp = nat.ptr_array([state[0].data_ptr() + 8, state[1].data_ptr()])
    rc = nat.lib().nw_adjoint(plan.handle, m.data_ptr(), m.data_ptr(), lp, None, 0, 1, None, 0,
                              None, 0, None, 0, plan.stream())
    assert rc != 0 and b"16-byte" in nat.lib().nw_last_error()
    plan.forward(m[: plan.host.cells], state, [], 0, 1)  # aligned buffers still run
