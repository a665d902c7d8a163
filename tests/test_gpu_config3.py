"""The benchmark workload itself against the FP64 oracle, on the production path.

Config 3: 1024² field grid (+20-cell sponge = 1064² padded), dx = 10 m, 64 shots at evenly
spaced x with source depths U[2, 40] cells and U[64, 512] receivers per shot along the top row,
4000 steps at CFL 0.4, snapshot every 200 steps (the survey ``bench.py`` times). The operator
runs exactly as the benchmark runs it: shipped tiles, hipGraphs captured and then replayed, the
temporally blocked forward sweep keeping its last blocks, and the reverse sweep overlapping the
recomputation of earlier blocks on a second stream.

The residual dJ/dtraces is nonzero on two shots only (shot 0, and the shot with the most
receivers), so the gradient of the whole 64-shot batch equals the sum of those two shots'
gradients, which the FP64 oracle computes with its own checkpointed reverse sweep (snapshots
every 200 steps, one block of states in host memory at a time). The oracle runs on a host thread
while the GPU computes.

The same two shots also run through the float64 HIP path and through the reference's own op
sequence (the conv2d leapfrog of nefi/physics/wave.py:529-552, restated in bench.py) in float32
and float64. Bars:

* float64 (HIP and the reference sequence) against the oracle: 1e-10 traces, 1e-9 gradient;
* float32 traces within 1e-5 relative L2 of the reference's float32 traces (the stated
  tolerance), and against the oracle within twice the reference float32's own error: after
  4000 float32 steps the reference's traces are themselves ~1.8e-5 from FP64;
* the float32 gradient within 1e-4 of the oracle's.
"""

from __future__ import annotations

import threading
import time

import numpy as np
import pytest
import torch

import bench
from _support import rel
from nefi_amd import _native as nat
from nefi_amd.domain import Domain
from nefi_amd.physics.batched import ShotBatchWaveOperator
from oracle.leapfrog import Problem

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.mark.timeout(600)
def test_config3_production_gradient_matches_fp64_oracle():
    wl = bench.survey(1024, 64, 4001)
    n, L = wl["n"], wl["L"]
    counts = [int(k) for k in wl["counts"]]
    picked = [0, int(np.argmax(counts))]
    starts = np.concatenate([[0], np.cumsum(counts)])
    print(f"[case config 3: 64 shots, 1064^2, 4000 steps, K=200; oracle shots {picked} with "
          f"{[counts[s] for s in picked]} receivers]", flush=True)
    rng = np.random.default_rng(11)
    res = np.zeros((int(starts[-1]), 4001))
    for s in picked:
        res[starts[s] : starts[s + 1]] = rng.standard_normal((counts[s], 4001))

    # FP64 oracle of the two shots, on a host thread (the C calls release the GIL)
    oracle_out: dict = {}

    def run_oracle():
        t0 = time.perf_counter()
        g = np.zeros((n, n))
        for s in picked:
            src, rec = wl["shots"][s]
            P = Problem((n, n), ((0.0, L), (0.0, L)), rec, sources=src, **wl["kw"])
            gs, tr = P.fwi_gradient(wl["c"], res[starts[s] : starts[s + 1]][None], every=200,
                                    with_traces=True)
            g += gs
            oracle_out[s] = tr[0]
        oracle_out["grad"] = g
        oracle_out["seconds"] = time.perf_counter() - t0

    worker = threading.Thread(target=run_oracle, daemon=True)
    worker.start()

    torch.cuda.empty_cache()  # the kept blocks are sized from free device memory
    dom = Domain((n, n), ((0.0, L), (0.0, L)))
    op = ShotBatchWaveOperator(dom, wl["shots"], grad_mode="checkpoint", checkpoint_every=200,
                               **wl["kw"])
    x = torch.tensor(wl["c"], dtype=torch.float32, device=DEV, requires_grad=True)
    r = torch.tensor(res, dtype=torch.float32, device=DEV)
    outs = []
    for call in range(2):  # the first call captures the graphs, the second replays them
        y = op({"c": x})
        (g,) = torch.autograd.grad((y * r).sum(), x)
        torch.cuda.synchronize()
        outs.append((y.detach().double().cpu().numpy(), g.double().cpu().numpy()))
        print(f"GPU gradient call {call} done", flush=True)
    prop = next(iter(op._cache.values()))
    plan = prop.plan
    kept = sum(1 for name in prop.ws.bufs if name.startswith("keep"))
    print(f"production path: snapshot every {prop.block}, {kept} kept blocks, graphs "
          f"{plan.option(nat.OPT_GRAPH_COUNT)}, tiles {plan.tiles(0)} / {plan.tiles(1)}, "
          f"depth {plan.option(nat.OPT_TIME_BLOCK)}", flush=True)
    assert prop.block == 200
    assert plan.option(nat.OPT_GRAPH) == 1 and plan.option(nat.OPT_GRAPH_COUNT) > 0
    assert plan.tiles(0)[0] > 0 and plan.tiles(1)[0] > 0  # the fast kernels
    # graph replays reproduce the capturing call's traces and gradient exactly
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    y, g = outs[1]
    del prop, plan
    op._cache.clear()  # release the production workspace (~245 GB of kept blocks)
    torch.cuda.empty_cache()

    # the same two shots through the float64 HIP path, and through the reference's own op
    # sequence (conv2d leapfrog, restated in bench.py) in float32 and float64
    sub = ShotBatchWaveOperator(dom, [wl["shots"][s] for s in picked], grad_mode="checkpoint",
                                checkpoint_every=200, **wl["kw"])
    rows = np.concatenate([np.arange(starts[s], starts[s + 1]) for s in picked])
    x64 = torch.tensor(wl["c"], dtype=torch.float64, device=DEV, requires_grad=True)
    r64 = torch.tensor(res[rows], dtype=torch.float64, device=DEV)
    y64 = sub({"c": x64})
    (g64,) = torch.autograd.grad((y64 * r64).sum(), x64)
    ref = {}
    for dt in (torch.float32, torch.float64):
        tr, gr = bench.reference_leapfrog(op, torch.tensor(wl["c"], dtype=dt, device=DEV), picked,
                                          op.n_steps, r64.to(dt), block=200)
        ref[dt] = (tr.double().cpu().numpy(), gr.double().cpu().numpy())
    y64 = y64.detach().cpu().numpy()
    g64 = g64.cpu().numpy()

    worker.join(timeout=540)
    assert not worker.is_alive(), "the FP64 oracle did not finish"
    print(f"FP64 oracle: 2 shots in {oracle_out['seconds']:.1f} s", flush=True)
    yo = np.concatenate([oracle_out[s] for s in picked], 0)
    go = oracle_out["grad"]
    err = {
        "traces f32 (production path)": rel(y[rows], yo),
        "traces f64 (HIP)": rel(y64, yo),
        "traces reference f32": rel(ref[torch.float32][0], yo),
        "traces reference f64": rel(ref[torch.float64][0], yo),
        "gradient f32 (production path)": rel(g, go),
        "gradient f64 (HIP)": rel(g64, go),
        "gradient reference f32": rel(ref[torch.float32][1], go),
        "gradient reference f64": rel(ref[torch.float64][1], go),
    }
    for k, v in err.items():
        print(f"{k} vs FP64 oracle: {v:.2e}", flush=True)
    for s in picked:
        print(f"shot {s}: float32 traces vs FP64 oracle "
              f"{rel(y[starts[s] : starts[s + 1]], oracle_out[s]):.2e}", flush=True)
    print(f"float32 traces vs the reference's float32 traces: "
          f"{rel(y[rows], ref[torch.float32][0]):.2e}", flush=True)
    assert err["traces f64 (HIP)"] < 1e-10 and err["gradient f64 (HIP)"] < 1e-9
    assert err["traces reference f64"] < 1e-10 and err["gradient reference f64"] < 1e-9
    assert err["gradient f32 (production path)"] < 1e-4
    assert rel(y[rows], ref[torch.float32][0]) < 1e-5
    assert err["traces f32 (production path)"] < 2 * err["traces reference f32"]
