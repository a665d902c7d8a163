"""Benchmark: fused forward + adjoint FWI gradient of the 64-shot uneven 2-D survey.

Workload (one "step" = one gradient evaluation dJ/dc: forward sweep with snapshots, block
recomputation and the reverse sweep with the imaging sum): 1024² field grid, dx = 10 m,
20-cell sponge (1064² padded), 4000 leapfrog steps at CFL 0.4, 64 shots at evenly spaced x and
depths drawn from U[2, 40] cells, per-shot receiver counts drawn from U[64, 512] along the top
row, snapshot every 200 steps. The sound speed is 2000 m/s + 1500 m/s·depth/L plus a smooth
seeded perturbation (Gaussian-filtered N(0,1), σ = 10 cells, max 100 m/s); the residual fed to
the reverse sweep is a fixed seeded random trace set. Inputs are resident on the GPU before
timing starts.

``value`` = padded cells × steps × shots / seconds per gradient (Gcell-updates/s) on the
production path: captured hipGraphs replayed (when the tuned plan enables them) and the reverse
sweep overlapping the next block's recompute on a second stream.
``roofline`` (and ``kernels_per_gradient``) come from a separate timed run in which every launch
is bracketed by HIP events, calls run eagerly and the reverse sweep runs on one stream, so each
launch has the GPU to itself. It uses the algorithmic bytes per cell-update of the dominant kernel (16 B for the
forward step: read p^n, p^{n-1}, m, write p^{n+1}; 28 B for the reverse step: 16 B adjoint
stencil + 12 B imaging) times the cell-updates of one launch (a temporally blocked forward
launch advances several steps) over its average duration, timed with HIP events around every
launch. Temporal blocking moves fewer bytes than that algorithmic count, so for the blocked
kernel ``achieved`` is an effective rate and ``traffic`` (measured HBM bytes) shows the real
one: ``measured_gbs`` / ``measured_frac`` are those bytes over the same launch time. The shot
groups of the reverse step also share the model tile and the float64 gradient slice, so its
effective ``frac`` can exceed 1 while the measured one stays below. ``gradient`` gives the whole
gradient's effective rate at the algorithmic bytes of a fwd + adj cell-update (60 B sponge,
108 B PML: forward, recompute, adjoint stencil and imaging, no blocking or sharing).
``cpu_baseline`` times the reference's CPU path (its op sequence in PyTorch on the CPU, float32,
all host threads; one thread and the FP64 oracle beside it) on a bounded sample of the same
workload on this machine's host cores.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {"step": 16.0, "steps_blocked": 16.0, "adjoint": 28.0,
                    "adjoint_blocked": 28.0}
# PML plans (--absorbing pml): the step also reads and writes both face fields (28 B); the
# reverse step reads the adjoint face fields and both psi levels of the imaging term (52 B), or
# with kept Laplacians (the PML default) one kept value instead of p^n and the psi levels (40 B)
BYTES_PER_UPDATE_PML = {"step": 28.0, "adjoint": 52.0}
BYTES_PER_UPDATE_PML_KEPT_LAP = {"step": 28.0, "adjoint": 40.0}
# a whole gradient's algorithmic bytes per fwd + adj cell-update: forward sweep, recompute sweep,
# adjoint stencil and imaging (16 + 16 + 16 + 12 B; PML 32 + 32 + 32 + 12 B)
GRADIENT_BYTES = {"sponge": 60.0, "cerjan": 60.0, "pml": 108.0}


def survey(n=1024, shots=64, n_t=4001, seed=0, absorbing="sponge"):
    """Geometry and model of the uneven 64-shot survey (host, float64)."""
    from scipy.ndimage import gaussian_filter

    dx = 10.0
    L = n * dx
    rng = np.random.default_rng(seed)
    z = (np.arange(n) + 0.5) / n
    pert = gaussian_filter(rng.standard_normal((n, n)), 10.0, mode="wrap")
    c = 2000.0 + 1500.0 * z[None, :] + 100.0 * pert / np.abs(pert).max()
    c_max = 3600.0
    depths = rng.uniform(2.0, 40.0, shots) * dx
    counts = rng.integers(64, 513, shots)
    shot_geo = []
    for s in range(shots):
        src = [[(s + 0.5) * L / shots, float(depths[s])]]
        r = int(counts[s])
        rec = [[(i + 0.5) * L / r, 0.5 * dx] for i in range(r)]
        shot_geo.append((src, rec))
    kw = dict(n_t=n_t, dt_obs=0.4 * dx / c_max, f0=15.0, c_max=c_max, order=4,
              absorbing=absorbing, absorb_width=20 * dx)
    return dict(n=n, dx=dx, L=L, c=c, shots=shot_geo, kw=kw, counts=counts)


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def cpu_baseline(wl, op, budget_s=10.0):
    """The reference's CPU path on this machine's host cores: the reference's own op sequence
    (``reference_leapfrog``: conv2d Laplacian, damping products, dense source map, gather,
    torch.utils.checkpoint every 200 steps, autograd backward) in float32 on the CPU, for shot 0
    of the workload's grid, over a bounded number of steps scaled to ~``budget_s`` seconds,
    with all host threads and with one thread. The FP64 oracle (C, OpenMP) on the same sample
    is a second entry."""
    import torch

    from oracle.leapfrog import Problem

    n = wl["n"]
    cells = op.grid.padded_shape[0] * op.grid.padded_shape[1]
    nproc = os.cpu_count() or 1
    threads_all = int(os.environ.get("OMP_NUM_THREADS", nproc))
    c = torch.tensor(wl["c"], dtype=torch.float32)
    rows = op.receiver_counts[0]
    res = torch.randn(rows, op.n_t, generator=torch.Generator().manual_seed(1))
    saved_threads = torch.get_num_threads()

    def run_ref(threads):
        torch.set_num_threads(threads)
        reference_leapfrog(op, c, [0], 2, res, block=200)  # warm-up
        t0 = time.perf_counter()
        reference_leapfrog(op, c, [0], 10, res, block=200)
        t = time.perf_counter() - t0
        steps = int(max(10, min(op.n_steps, budget_s / max(t, 1e-3) * 10)))
        t0 = time.perf_counter()
        reference_leapfrog(op, c, [0], steps, res, block=200)
        t = time.perf_counter() - t0
        return {"value": round(cells * steps / t / 1e9, 5), "unit": "Gcell-updates/s",
                "cores": threads, "kind": "port",
                "sample": f"the reference's op sequence (conv2d leapfrog, checkpoint every 200 "
                          f"steps, autograd backward) in float32 on the CPU with {threads} "
                          f"thread(s): 1 shot x {steps} steps on the {cells}-cell padded grid "
                          f"in {t:.1f} s, gradient (forward + recompute + backward)"}

    try:
        main = run_ref(threads_all)
        one = run_ref(1)
    finally:
        torch.set_num_threads(saved_threads)

    src, rec = wl["shots"][0]
    kw = dict(wl["kw"])

    def run_oracle(n_t):
        kw["n_t"] = n_t
        P = Problem((n, n), ((0.0, wl["L"]), (0.0, wl["L"])), rec, sources=src, **kw)
        r = np.random.default_rng(1).standard_normal((1, len(rec), n_t))
        t0 = time.perf_counter()
        P.fwi_gradient(wl["c"], r, every=200)
        return time.perf_counter() - t0, P.cells * P.n_steps, P.n_steps

    t, work, _ = run_oracle(6)
    steps = int(max(5, min(4000, (budget_s / max(t, 1e-3)) * 5)))
    t, work, steps = run_oracle(steps + 1)
    main["one_thread"] = one
    main["oracle_fp64"] = {
        "value": round(work / t / 1e9, 5), "unit": "Gcell-updates/s", "cores": threads_all,
        "kind": "port",
        "sample": f"FP64 oracle gradient (C, OpenMP; snapshots every 200 steps, block "
                  f"recomputation), 1 shot x {steps} steps on the same grid, {t:.1f} s"}
    main["cpu_model"] = _cpu_model()
    main["nproc"] = nproc
    return main


def reference_leapfrog(op, c, shot_ids, n_steps, residual=None, block=200):
    """The plain PyTorch path of the reference on the GPU, for the shots ``shot_ids``.

    The reference's own op sequence for a sponge layer (nefi/physics/wave.py:529-552 inside
    timestep.py:151-169): conv2d Laplacian with the (1,1,5,5) cross kernel, addcmul, a dense
    source map times the host wavelet value, the damping products, multilinear gather at every
    record step, and torch.utils.checkpoint over blocks of `block` steps (block=None: every
    step's tape kept, grad_mode="autograd"). Returns ``(traces, dJ/dc)`` for
    J = sum(residual * traces), with ``residual`` the rows of those shots; without a residual
    it runs under no_grad and returns ``(traces, None)``. ``traces`` is ``(rows, samples)``.
    """
    import torch
    import torch.nn.functional as F
    from torch.utils.checkpoint import checkpoint

    g = op.grid
    dev, dt = c.device, op.dt
    dtype = c.dtype
    shot_ids = list(shot_ids)
    shots = len(shot_ids)
    kern = torch.zeros(5, 5, dtype=torch.float64)
    for ax, h in enumerate(g.spacing):
        for off, co in ((-2, -1 / 12), (-1, 4 / 3), (0, -2.5), (1, 4 / 3), (2, -1 / 12)):
            at = [2, 2]
            at[ax] += off
            kern[tuple(at)] += co / h**2
    kern = kern.to(dev, dtype).view(1, 1, 5, 5)
    sig = op.damping_profile()
    half = 0.5 * sig * dt
    a = (1.0 / (1.0 + half)).to(dev, dtype)
    b = (1.0 - half).to(dev, dtype)
    src = (g.source_density(torch.cat([op.shot_sources[s] for s in shot_ids])) * dt**2).to(
        dev, dtype)
    wav = [float(v) for v in op.wavelet()]
    idx, w = op._rec
    counts = [op.receiver_counts[s] for s in shot_ids]
    starts = [sum(op.receiver_counts[:s]) for s in shot_ids]
    rows_sel = torch.cat([torch.arange(r0, r0 + n) for r0, n in zip(starts, counts)])
    ridx = idx[rows_sel].to(dev)
    rw = w[rows_sel].to(dev, dtype)
    shot_of = torch.repeat_interleave(torch.arange(shots), torch.tensor(counts)).to(dev)
    flat_idx = shot_of[:, None] * (g.padded_shape[0] * g.padded_shape[1]) + ridx
    record = set(op.record)

    def sample(p):
        return (p.reshape(-1)[flat_idx] * rw).sum(-1)

    def step(p_prev, p, c2dt2, n):
        lap = F.conv2d(p.reshape(-1, 1, *g.padded_shape), kern, padding=2).reshape(p.shape)
        rhs = torch.addcmul(2.0 * p - b * p_prev, c2dt2, lap)
        rhs = rhs + src * wav[n]
        return p, a * rhs

    with torch.set_grad_enabled(residual is not None):
        x = c.detach().clone().requires_grad_(residual is not None)
        c2dt2 = g.pad(x, "replicate") ** 2 * dt**2
        p_prev = torch.zeros(shots, *g.padded_shape, device=dev, dtype=dtype)
        p = torch.zeros_like(p_prev)
        obs = [sample(p)]
        blk = block if residual is not None else None
        for n0 in range(0, n_steps, blk or n_steps):
            n1 = min(n_steps, n0 + (blk or n_steps))

            def seg(pp, pc, m, n0=n0, n1=n1):
                out = []
                for n in range(n0, n1):
                    pp, pc = step(pp, pc, m, n)
                    if n + 1 in record:
                        out.append(sample(pc))
                return (pp, pc, *out)

            res = seg(p_prev, p, c2dt2) if not blk else checkpoint(seg, p_prev, p, c2dt2,
                                                                   use_reentrant=False)
            p_prev, p = res[0], res[1]
            obs.extend(res[2:])
        traces = torch.stack(obs, -1)
        if residual is None:
            return traces, None
        loss = (traces * residual[:, : traces.shape[-1]]).sum()
        (grad,) = torch.autograd.grad(loss, x)
    return traces.detach(), grad


def autograd_leapfrog(op, c, shots, n_steps, residual, block=200):
    """dJ/dc of the reference's plain PyTorch path for the first ``shots`` shots
    (J = sum(residual * traces); ``residual`` holds at least those shots' rows)."""
    rows = sum(op.receiver_counts[:shots])
    return reference_leapfrog(op, c, range(shots), n_steps, residual[:rows], block)[1]


def autograd_comparator(op, c, residual, repeats=3, ckpt_steps=1000):
    """Time the plain PyTorch autograd path on the same GPU: with every step's tape kept
    (grad_mode="autograd") and with checkpoint blocks of 200 (grad_mode="checkpoint"), each at
    the source batch that fits HBM for the whole workload.

    Store-all: one batch of that size over all of the workload's steps, scaled only in shots.
    Checkpointed: every shot in one batch (it fits), over ``ckpt_steps`` steps, scaled in
    steps (its cost per step is the same in every block). Each is run ``repeats`` times after
    a warm-up; the min and the median are reported, the median is used.

    Batch sizing: the tape bytes per shot-step are measured on a short probe; store-all keeps
    n_steps of them per shot, checkpointing keeps one block of 200 plus two wavefields per
    snapshot. The budget is 85 % of the free device memory."""
    import statistics

    import torch

    n_shots = len(op.shot_sources)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    base = torch.cuda.memory_allocated()
    torch.cuda.reset_peak_memory_stats()
    probe = 40
    autograd_leapfrog(op, c, 1, probe, residual, block=None)  # also warms MIOpen, allocator
    torch.cuda.synchronize()
    tape = (torch.cuda.max_memory_allocated() - base) / probe
    torch.cuda.empty_cache()
    budget = 0.85 * torch.cuda.mem_get_info()[0]
    state = op.grid.padded_shape[0] * op.grid.padded_shape[1] * c.element_size()
    per_shot = {"autograd": tape * op.n_steps,
                "checkpoint": tape * 200 + 2 * state * -(-op.n_steps // 200)}
    out = {}
    for mode, block in (("autograd", None), ("checkpoint", 200)):
        batch = int(max(1, min(n_shots, budget // per_shot[mode])))
        steps = op.n_steps if mode == "autograd" else min(op.n_steps, ckpt_steps)
        autograd_leapfrog(op, c, batch, 20, residual, block=block)  # warm up this batch
        torch.cuda.synchronize()
        times = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            autograd_leapfrog(op, c, batch, steps, residual, block=block)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
            print(f"[bench] autograd comparator {mode}: {batch} shots x {steps} steps "
                  f"{times[-1]:.2f} s", file=sys.stderr, flush=True)
            torch.cuda.empty_cache()
        scale = (op.n_steps / steps) * (n_shots / batch)
        med = statistics.median(times)
        out[mode] = {
            "ms_per_gradient_scaled": round(med * scale * 1e3, 1),
            "ms_per_gradient_scaled_min": round(min(times) * scale * 1e3, 1),
            "source_batch": batch,
            "repeats": repeats,
            "sample": f"source_batch {batch} (tape {tape / 2**20:.1f} MiB per shot-step, "
                      f"{per_shot[mode] / 2**30:.1f} GiB per shot, budget "
                      f"{budget / 2**30:.0f} GiB): {batch} shots x {steps} steps measured "
                      f"{repeats} times, median {med:.2f} s, min {min(times):.2f} s, scaled "
                      f"x{scale:g} to {n_shots} shots x {op.n_steps} steps",
        }
        torch.cuda.empty_cache()
    return out


def gradient_error_vs_fp64(op, wl, c32):
    """Accuracy of the timed plan itself: the production operator ``op`` (64 shots, the timed
    plan, its graphs and kept blocks) differentiates a residual that is nonzero on two shots
    (shot 0 and the shot with the most receivers); its float32 traces and gradient are compared
    with the float64 HIP path on those two shots (which equals the FP64 oracle to ~1e-14 at this
    size and the reference's float64 autograd to ~1e-15 in the parity tests)."""
    import torch

    from nefi_amd.physics.batched import ShotBatchWaveOperator

    counts = [int(k) for k in wl["counts"]]
    picked = [0, int(np.argmax(counts))]
    starts = np.concatenate([[0], np.cumsum(counts)])
    rows = np.concatenate([np.arange(starts[s], starts[s + 1]) for s in picked])
    gen = torch.Generator("cuda").manual_seed(7)
    res64 = torch.zeros(int(starts[-1]), op.n_t, device="cuda", dtype=torch.float64)
    rsel = torch.as_tensor(rows, device="cuda")
    res64[rsel] = torch.randn(len(rows), op.n_t, device="cuda", dtype=torch.float64,
                              generator=gen)
    x = c32.detach().clone().requires_grad_(True)
    y = op({"c": x})
    (g32,) = torch.autograd.grad((y * res64.float()).sum(), x)
    y32 = y.detach()[rsel].double()
    sub = ShotBatchWaveOperator(op.domain, [wl["shots"][s] for s in picked],
                                grad_mode="checkpoint", checkpoint_every=200, **wl["kw"])
    x64 = c32.detach().double().requires_grad_(True)
    y64 = sub({"c": x64})
    (g64,) = torch.autograd.grad((y64 * res64[rsel]).sum(), x64)
    eg = float((g32.double() - g64).norm() / g64.norm())
    et = float((y32 - y64.detach()).norm() / y64.detach().norm())
    del sub
    torch.cuda.empty_cache()
    return {"rel_l2": eg, "traces_rel_l2": et,
            "sample": f"the timed {len(op.shot_sources)}-shot plan with the residual on shots "
                      f"{picked}, full grid and {op.n_steps} steps, against the float64 HIP "
                      f"path on those shots"}


def misfit_bench(rows, n_t, dt_obs, f_max, dev, reps=5):
    """wave_fwi's band-limited misfit (BandLimitedMSE) and its adjoint source on traces of the
    workload's shape, timed with the library's HIP events on the filter launches."""
    from types import SimpleNamespace

    import torch

    from nefi_amd import _native as nat
    from nefi_amd.instances.wave_fwi import BandLimitedMSE

    g = torch.Generator(dev).manual_seed(3)
    pred = torch.randn(rows, n_t, device=dev, generator=g).requires_grad_(True)
    obs = torch.randn(rows, n_t, device=dev, generator=g)
    ctx = SimpleNamespace(pred=pred, obs=SimpleNamespace(data=obs, mask=None))
    loss_fn = BandLimitedMSE(f_max, dt_obs)
    for _ in range(2):
        torch.autograd.grad(loss_fn(ctx), pred)
    torch.cuda.synchronize()
    nat.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        torch.autograd.grad(loss_fn(ctx), pred)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / reps
    ms, launches = nat.profile_read(nat.K_FILTER)
    nat.profile_enable(False)
    # fused residual pass: read pred, obs, write Z; obs pass: read obs; adjoint: read Z, write
    bytes_per_call = 24 * rows * n_t
    per_call_s = ms / 1e3 / reps
    return {
        "workload": f"BandLimitedMSE(f_max={f_max:g} Hz) forward + adjoint source, {rows} traces "
                    f"x {n_t} samples, float32",
        "ms_per_call": round(per_call_s * 1e3, 3),
        "wall_ms_per_call": round(wall * 1e3, 3),
        "filter_launches_per_call": launches // reps,
        "hbm_bytes_per_call": bytes_per_call,
        "achieved_gbs": round(bytes_per_call / per_call_s / 1e9, 1),
        "frac_of_hbm": round(bytes_per_call / per_call_s / 1e9 / HBM_PEAK_GBS, 4),
    }


def pat3d_config(n=192, dx=1e-4, seed=2, n_det=4096, n_abs=20):
    """Inputs of the 3-D photoacoustic configuration (host, float64): 192^3 cells of 0.1 mm,
    sound speed 1500 m/s with +-5 % smooth seeded heterogeneity (Gaussian-filtered N(0,1),
    sigma 6 cells), initial pressure of ``n_abs`` seeded ellipsoidal absorbers (amplitudes
    U[0.5, 1], semi-axes U[0.4, 1.2] mm, random orientations, tanh edges 0.15 mm wide, summed
    and clipped to 1), and ``n_det`` detectors on a hemispherical cap: radius 0.45 L around the
    centre of the top face (depth = last axis), reaching into the volume, sampled by a
    Fibonacci lattice (equal-area bands in depth, golden-angle azimuths). The absorbers lie
    inside the cap. Returns ``(L, c, p0, detectors)``."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    L = n * dx
    pert = gaussian_filter(rng.standard_normal((n, n, n)), 6.0, mode="wrap")
    c = 1500.0 * (1.0 + 0.05 * pert / np.abs(pert).max())
    R = 0.45 * L
    centre = np.array([0.5 * L, 0.5 * L, 0.0])
    k = np.arange(n_det)
    cos_t = (k + 0.5) / n_det
    sin_t = np.sqrt(1.0 - cos_t**2)
    phi = np.pi * (3.0 - math.sqrt(5.0)) * k
    det = centre + R * np.stack([sin_t * np.cos(phi), sin_t * np.sin(phi), cos_t], 1)
    ax = (np.arange(n) + 0.5) * dx
    p0 = np.zeros((n, n, n))
    width = 1.5 * dx
    for _ in range(n_abs):
        while True:  # a centre in the cap, away from its rim
            q = centre + rng.uniform(-0.7 * R, 0.7 * R, 3) * np.array([1.0, 1.0, 0.0])
            q[2] = rng.uniform(0.15 * R, 0.75 * R)
            if np.linalg.norm(q - centre) < 0.7 * R:
                break
        semi = rng.uniform(4.0, 12.0, 3) * dx
        rot, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        amp = rng.uniform(0.5, 1.0)
        reach = semi.max() + 4 * width
        sl = [slice(max(0, int((q[a] - reach) / dx)), min(n, int((q[a] + reach) / dx) + 2))
              for a in range(3)]
        X, Y, Z = np.meshgrid(ax[sl[0]] - q[0], ax[sl[1]] - q[1], ax[sl[2]] - q[2],
                              indexing="ij")
        body = np.einsum("ij,...j->...i", rot.T, np.stack([X, Y, Z], -1)) / semi
        dist = (np.sqrt((body**2).sum(-1)) - 1.0) * semi.min()  # about the signed distance
        p0[tuple(sl)] += amp * 0.5 * (1.0 - np.tanh(dist / width))
    return L, c, np.minimum(p0, 1.0), det.tolist()


def pat3d_bench(args, dev) -> dict:
    """The 3-D photoacoustic operator (192^3 + 10-cell sponge = 212^3, order 2, 1500 steps; the
    inputs of ``pat3d_config``: 20 ellipsoidal absorbers, 4096 detectors on a hemispherical cap)
    forward + adjoint in float32: throughput, and the roofline of its dominant kernel from HIP
    events."""
    import torch

    from nefi_amd import _native as nat
    from nefi_amd.domain import Domain
    from nefi_amd.physics.wave import WaveInitialConditionOperator

    n, dx = 192, 1e-4
    L, c, p0_np, det = pat3d_config(n, dx)
    c_max = float(c.max())
    op = WaveInitialConditionOperator(Domain((n, n, n), ((0.0, L),) * 3), det, n_t=1501,
                                      dt_obs=0.3 * dx / c_max, c=torch.tensor(c), c_max=c_max,
                                      order=2, absorbing="sponge", absorb_width=10 * dx)
    p0 = torch.tensor(p0_np, dtype=torch.float32, device=dev, requires_grad=True)
    w = torch.randn(len(det), 1501, device=dev, generator=torch.Generator(dev).manual_seed(3))

    def step():
        p0.grad = None
        op({"p0": p0}).backward(w)

    for _ in range(args.warmup):
        step()
    # production path (each operator call one hipGraph launch after its first use)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = (time.perf_counter() - t0) / args.steps
    # kernel timings: the same calls run eagerly with HIP events around every launch
    nat.profile_enable(True)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    timing = {k: nat.profile_read(v) for k, v in (("step", nat.K_STEP), ("adjoint", nat.K_ADJOINT))}
    nat.profile_enable(False)
    padded = op.grid.padded_shape
    cells = padded[0] * padded[1] * padded[2]
    fast = next(iter(op._cache.values())).plan.option(nat.OPT_FAST_3D)
    dom = max(timing, key=lambda k: timing[k][0])
    ms, launches = timing[dom]
    avg_s = ms / 1e3 / launches
    bpu = 16.0  # step: p^n, p^{n-1}, m read, p^{n+1} written; reverse: lambda pair, m a, out
    achieved = bpu * cells / avg_s / 1e9
    return {
        "metric": "3-D photoacoustic forward+adjoint Gcell-updates/s",
        "value": round(cells * op.n_steps / elapsed / 1e9, 3),
        "unit": "Gcell-updates/s",
        "ms_per_step": round(elapsed * 1e3, 2),
        "steps": args.steps, "warmup": args.warmup, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"3-D PAT, 192^3 + 10-cell sponge = {padded[0]}^3, order 2, "
                               f"20 ellipsoidal absorbers, 4096 detectors on a hemispherical cap, "
                               f"{op.n_steps} steps, forward + adjoint"},
        "roofline": {"bound": "hbm",
                     "kernel": (f"k3d_{dom}<order 2, sponge> (2.5-D tiles)" if fast
                                else f"k_{dom}<float, order 2, sponge> (generic)"),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_cell_update": bpu,
                     "avg_launch_us": round(avg_s * 1e6, 2)},
        "kernels_per_gradient": {k: {"total_ms": round(v[0] / args.steps, 2),
                                     "launches": v[1] // args.steps} for k, v in timing.items()},
    }


def datagen_bench(args, dev) -> dict:
    """Float64 data generation on the 2x-finer grid, as nefi's wave_fwi data generator runs
    its operator (twice the resolution, float64, no gradient): the benchmark survey's geometry
    and model on 2048^2 cells of 5 m (+ the same 200 m sponge = 40 cells), 64 shots, forward
    only. A bounded run of ``--nt`` samples (default 101: 200 steps at two substeps per
    sample) gives the rate; the full 4001-sample generation is projected from it."""
    import torch

    from nefi_amd import _native as nat
    from nefi_amd.domain import Domain
    from nefi_amd.physics.batched import ShotBatchWaveOperator

    wl = survey(1024, args.shots, 4001, absorbing=args.absorbing)
    n_t = args.nt if args.nt != 4001 else 101
    kw = dict(wl["kw"], n_t=n_t)
    L = wl["L"]
    n = 2 * wl["n"]
    dom = Domain((n, n), ((0.0, L), (0.0, L)))
    op = ShotBatchWaveOperator(dom, wl["shots"], grad_mode="checkpoint", **kw)
    c = torch.tensor(np.kron(wl["c"], np.ones((2, 2))), dtype=torch.float64, device=dev)
    with torch.no_grad():
        for _ in range(args.warmup):
            op({"c": c})
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            y = op({"c": c})
        torch.cuda.synchronize()
        elapsed = (time.perf_counter() - t0) / args.steps
        nat.profile_enable(True)
        op({"c": c})
        torch.cuda.synchronize()
        ms, launches = nat.profile_read(nat.K_STEP)
        nat.profile_enable(False)
    assert bool(torch.isfinite(y).all())
    tiled = next(iter(op._cache.values())).plan.option(nat.OPT_FAST_F64)
    padded = op.grid.padded_shape
    updates = padded[0] * padded[1] * args.shots * op.n_steps
    avg_s = ms / 1e3 / max(1, launches)
    bpu = 32.0  # float64: read p^n, p^{n-1}, m; write p^{n+1}
    achieved = bpu * padded[0] * padded[1] * args.shots / avg_s / 1e9
    full_steps = op.n_steps * (4000 // (n_t - 1))
    return {
        "metric": "float64 data generation Gcell-updates/s (2x-fine 64-shot survey, forward)",
        "value": round(updates / elapsed / 1e9, 3),
        "unit": "Gcell-updates/s",
        "ms_per_step": round(elapsed * 1e3, 2),
        "steps": args.steps, "warmup": args.warmup, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2x-fine data generation: {n}^2 cells of 5 m (+40-cell sponge = "
                               f"{padded[0]}^2), {args.shots} shots, {op.n_steps} steps "
                               f"({n_t} samples, {op.substeps} substeps each), float64, "
                               f"forward only"},
        "full_generation_s": round(elapsed * full_steps / op.n_steps, 1),
        "full_generation_steps": full_steps,
        "roofline": {"bound": "hbm",
                     "kernel": ("k2d64_step<order 4, sponge, 16 rows> (LDS tiles)" if tiled
                                else "k_step<double, order 4, sponge> (generic)"),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_cell_update": bpu,
                     "avg_launch_us": round(avg_s * 1e6, 2)},
    }


def datagen3d_bench(args, dev) -> dict:
    """Float64 data generation for the 3-D photoacoustic workload on the 2x-finer grid
    (photoacoustic3d's data generator: 384^3 cells of 50 um + the same 1 mm sponge = 424^3,
    order 2, 4096 detectors, float64, forward only). A bounded run of ``--nt`` samples (default
    51: 100 steps at two substeps per sample) gives the rate; the full 1501-sample generation is
    projected from it."""
    import torch
    from scipy.ndimage import gaussian_filter, zoom

    from nefi_amd import _native as nat
    from nefi_amd.domain import Domain
    from nefi_amd.physics.wave import WaveInitialConditionOperator

    n, dx = 192, 1e-4
    L = n * dx
    rng = np.random.default_rng(2)
    pert = gaussian_filter(rng.standard_normal((n, n, n)), 6.0, mode="wrap")
    c = 1500.0 * (1.0 + 0.05 * pert / np.abs(pert).max())
    det = [[(i + 0.5) * L / 64, (j + 0.5) * L / 64, 0.5 * dx] for i in range(64)
           for j in range(64)]
    c_max = float(c.max())
    n_t = args.nt if args.nt != 4001 else 51
    fine = 2 * n
    op = WaveInitialConditionOperator(Domain((fine,) * 3, ((0.0, L),) * 3), det, n_t=n_t,
                                      dt_obs=0.3 * dx / c_max, c=torch.tensor(zoom(c, 2, order=1)),
                                      c_max=c_max, order=2, absorbing="sponge",
                                      absorb_width=10 * dx)
    p0 = torch.tensor(gaussian_filter(rng.standard_normal((fine,) * 3), 6.0, mode="wrap"),
                      dtype=torch.float64, device=dev)
    with torch.no_grad():
        for _ in range(args.warmup):
            op({"p0": p0})
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            y = op({"p0": p0})
        torch.cuda.synchronize()
        elapsed = (time.perf_counter() - t0) / args.steps
        nat.profile_enable(True)
        op({"p0": p0})
        torch.cuda.synchronize()
        ms, launches = nat.profile_read(nat.K_STEP)
        nat.profile_enable(False)
    assert bool(torch.isfinite(y).all())
    tiled = next(iter(op._cache.values())).plan.option(nat.OPT_FAST_F64)
    padded = op.grid.padded_shape
    cells = padded[0] * padded[1] * padded[2]
    avg_s = ms / 1e3 / max(1, launches)
    bpu = 32.0  # float64: read p^n, p^{n-1}, m; write p^{n+1}
    achieved = bpu * cells / avg_s / 1e9
    full_steps = op.substeps * 1500
    return {
        "metric": "3-D float64 data generation Gcell-updates/s (2x-fine photoacoustic, forward)",
        "value": round(cells * op.n_steps / elapsed / 1e9, 3),
        "unit": "Gcell-updates/s",
        "ms_per_step": round(elapsed * 1e3, 2),
        "steps": args.steps, "warmup": args.warmup, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2x-fine 3-D PAT data generation: {fine}^3 cells of 50 um "
                               f"(+20-cell sponge = {padded[0]}^3), 4096 detectors, order 2, "
                               f"{op.n_steps} steps ({n_t} samples, {op.substeps} substeps "
                               f"each), float64, forward only"},
        "full_generation_s": round(elapsed * full_steps / op.n_steps, 1),
        "full_generation_steps": full_steps,
        "roofline": {"bound": "hbm",
                     "kernel": ("k3d64_step<order 2, sponge, 4 rows, 8 planes> (2.5-D tiles)"
                                if tiled else "k_step<double, order 2, sponge> (generic)"),
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_cell_update": bpu,
                     "avg_launch_us": round(avg_s * 1e6, 2)},
    }


def measured_traffic(kernel_key: str, shots_per_group: int, xcd_map: int):
    """HBM bytes per launch of one kernel instance (with ``shots_per_group`` shots per workgroup
    and the tile order ``xcd_map``) from the committed rocprofv3 counter runs
    (profiles/pmc_traffic_*.md, FETCH_SIZE and WRITE_SIZE scaled by the calibration kernel)."""
    for path in sorted((ROOT / "profiles").glob("pmc_traffic_*.md")):
        text = path.read_text()
        start = text.rfind("```json")
        if start < 0:
            continue
        blob = json.loads(text[start + 7 : text.index("```", start + 7)])
        k = blob["kernels"].get(f"{kernel_key}|spw{shots_per_group}|xcd{xcd_map}")
        if k:
            return k["fetch_bytes"] + k["write_bytes"], path.name
    return None, None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--shots", type=int, default=64)
    ap.add_argument("--nt", type=int, default=4001)
    ap.add_argument("--ckpt", type=int, default=200)
    ap.add_argument("--workload", default="fwi",
                    choices=("fwi", "pat3d", "datagen", "datagen3d"),
                    help="fwi: the 64-shot 2-D gradient (the benchmark); pat3d: the 3-D "
                         "photoacoustic forward + adjoint; datagen: float64 forward on the "
                         "2x-finer grid (nefi's data generator); datagen3d: the same for the "
                         "3-D photoacoustic workload")
    ap.add_argument("--absorbing", default="sponge", choices=("sponge", "pml", "cerjan"),
                    help="absorbing layer of the survey (the benchmark workload uses the sponge)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autograd", action="store_true",
                    help="skip the plain-autograd comparator on the GPU")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the float32-vs-float64 gradient check")
    ap.add_argument("--no-misfit", action="store_true",
                    help="skip timing the band-limited misfit on the workload's traces")
    ap.add_argument("--calibrate", action="store_true",
                    help="first run a plain 512 MiB read + 512 MiB write elementwise kernel "
                         "(reference point for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build runs the propagator on a single MI355X "
              "(shots would shard across GPUs with one gradient all-reduce; not measured here)",
              file=sys.stderr)
        return 2

    import torch

    from nefi_amd import _native as nat
    from nefi_amd.domain import Domain
    from nefi_amd.physics.batched import ShotBatchWaveOperator

    dev = torch.device("cuda", 0)
    if args.calibrate:
        a = torch.ones(1 << 27, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "pat3d":
        print(json.dumps(pat3d_bench(args, dev)), flush=True)
        return 0
    if args.workload == "datagen":
        print(json.dumps(datagen_bench(args, dev)), flush=True)
        return 0
    if args.workload == "datagen3d":
        print(json.dumps(datagen3d_bench(args, dev)), flush=True)
        return 0
    wl = survey(args.n, args.shots, args.nt, absorbing=args.absorbing)
    n = wl["n"]
    dom = Domain((n, n), ((0.0, wl["L"]), (0.0, wl["L"])))
    op = ShotBatchWaveOperator(dom, wl["shots"], grad_mode="checkpoint",
                               checkpoint_every=args.ckpt, **wl["kw"])
    c = torch.tensor(wl["c"], dtype=torch.float32, device=dev, requires_grad=True)
    rows = int(sum(wl["counts"]))
    residual = torch.randn(rows, op.n_t, device=dev, generator=torch.Generator(dev).manual_seed(1))
    padded = op.grid.padded_shape
    updates = padded[0] * padded[1] * op.n_steps * args.shots

    def step():
        c.grad = None
        y = op({"c": c})
        y.backward(residual)

    print(f"[bench] {args.shots} shots, {padded[0]}x{padded[1]} padded, {op.n_steps} steps, "
          f"K={args.ckpt}, receivers {rows}", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        step()
        torch.cuda.synchronize()
        print(f"[bench] warmup {i} done", file=sys.stderr, flush=True)
    kinds = {"step": nat.K_STEP, "steps_blocked": nat.K_STEPS, "adjoint": nat.K_ADJOINT,
             "adjoint_blocked": nat.K_ADJOINT_STEPS, "sample": nat.K_SAMPLE,
             "inject": nat.K_INJECT}
    # timed region 1 (kernel timings): every launch bracketed by HIP events, calls eager, the
    # reverse sweep on one stream so that each launch has the GPU to itself
    overlap_env = os.environ.get("NEFI_AMD_OVERLAP")
    os.environ["NEFI_AMD_OVERLAP"] = "0"
    step()  # warm the one-stream schedule
    torch.cuda.synchronize()
    nat.profile_enable(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        print(f"[bench] kernel-timing step {i} queued", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed_eager = time.perf_counter() - t0
    if next(iter(op._cache.values())).block != args.ckpt:
        raise SystemExit("the kernel-timing run changed the snapshot interval")
    timing = {k: nat.profile_read(v) for k, v in kinds.items()}
    units = {k: nat.profile_units(v) for k, v in kinds.items()}
    nat.profile_enable(False)
    if overlap_env is None:
        del os.environ["NEFI_AMD_OVERLAP"]
    else:
        os.environ["NEFI_AMD_OVERLAP"] = overlap_env
    # timed region 2 (the job): the production path, i.e. captured hipGraphs replayed (when the
    # plan enables them) and the reverse sweep overlapping the next block's recompute
    plan = next(iter(op._cache.values())).plan
    step()  # captures any call whose graph is not cached yet
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        print(f"[bench] step {i} queued", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    print(f"[bench] production path: {elapsed * 1e3 / args.steps:.1f} ms per gradient, "
          f"{plan.option(nat.OPT_GRAPH_COUNT)} graphs", file=sys.stderr, flush=True)
    prop = next(iter(op._cache.values()))
    if prop.block != args.ckpt:  # the workload snapshots every --ckpt steps; nothing else is valid
        raise SystemExit(f"snapshot interval {prop.block} used instead of {args.ckpt}")
    # blocks the forward sweep keeps for the reverse sweep (one keep buffer each)
    kept_blocks = sum(1 for name in prop.ws.bufs if name.startswith("keep"))
    del prop  # the comparator later frees the operator's workspace; hold no reference to it
    ms_per_step = elapsed * 1e3 / args.steps
    value = updates / (elapsed / args.steps) / 1e9

    tiles = {name: dict(zip(("rows", "shots_per_group"), plan.tiles(kind)))
             for kind, name in ((0, "step"), (1, "adjoint"))}
    tiles["xcd_map"] = plan.option(nat.OPT_XCD_MAP)
    tiles["rows_per_iter"] = plan.option(nat.OPT_ROWS_PER_ITER)
    tiles["lds"] = [plan.option(nat.OPT_LDS_TILES_STEP), plan.option(nat.OPT_LDS_TILES_ADJOINT)]
    tiles["hip_graph"] = plan.option(nat.OPT_GRAPH)
    tiles["time_block"] = plan.option(nat.OPT_TIME_BLOCK)
    tiles["time_block_adjoint"] = plan.option(nat.OPT_TIME_BLOCK_ADJOINT)
    tiles["keep_laplacian"] = plan.option(nat.OPT_KEEP_LAP)
    tiles["source"] = (plan.tune_choice or {}).get("source", "library defaults")
    dom_kind = max(("step", "steps_blocked", "adjoint", "adjoint_blocked"),
                   key=lambda k: timing[k][0])
    reverse = dom_kind.startswith("adjoint")
    rows_t = tiles["adjoint" if reverse else "step"]["rows"]
    lds_t = tiles["lds"][1 if reverse else 0]
    fam = "k3" if lds_t else "k2"
    blocked = dom_kind in ("steps_blocked", "adjoint_blocked")
    depth = tiles["time_block_adjoint" if reverse else "time_block"]
    kname = {"step": f"{fam}_step" if rows_t else "k_step",
             "steps_blocked": "k4_steps",
             "adjoint": f"{fam}_adjoint" if rows_t else "k_adjoint",
             "adjoint_blocked": "k5_adjoint"}[dom_kind]
    pml = op.absorbing == "pml"
    if pml and rows_t:  # PML plans always run the LDS-tile face kernels, 8 or 16 rows
        kname = {"step": "k3_step_pml", "adjoint": "k3_adjoint_pml"}[dom_kind]
        rows_t = 8 if rows_t <= 8 else 16
    kname += f"<order 4, {op.absorbing}{', rows %d' % rows_t if rows_t else ''}"
    kname += f", depth {depth}>" if blocked else ">"
    tot_ms, launches = timing[dom_kind]
    # (under a counter-collecting profiler the library's events record nothing)
    avg_s = tot_ms / 1e3 / launches if launches and tot_ms > 0 else float("nan")
    steps_per_launch = units[dom_kind] / launches if launches else 1.0
    cells_per_launch = padded[0] * padded[1] * args.shots * steps_per_launch
    if pml:
        bpu = (BYTES_PER_UPDATE_PML_KEPT_LAP if tiles["keep_laplacian"]
               else BYTES_PER_UPDATE_PML)[dom_kind]
    else:
        bpu = BYTES_PER_UPDATE[dom_kind]
    achieved = bpu * cells_per_launch / avg_s / 1e9 if avg_s == avg_s else 0.0
    r_code = 2 if op.order == 4 else 1
    a_code = {"none": 0, "sponge": 1, "cerjan": 2, "pml": 3}[op.absorbing]
    inst = (f"nw::{kname.split('<')[0]}<{r_code}, {rows_t}" if pml and rows_t else
            f"nw::{kname.split('<')[0]}<{r_code}, {a_code}, {rows_t}")
    if fam == "k2" and dom_kind == "step" and not pml:
        inst += f", {tiles['rows_per_iter']}"
    if blocked:
        inst += f", {depth}"
    spw_t = tiles["adjoint" if reverse else "step"]["shots_per_group"]
    traffic, traffic_src = measured_traffic(inst + ">", spw_t, tiles["xcd_map"])
    roofline = {
        "bound": "hbm",
        "kernel": kname,
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_source": traffic_src,
        "bytes_per_cell_update": bpu,
        "steps_per_launch": round(steps_per_launch, 3),
        "avg_launch_us": round(avg_s * 1e6, 2),
    }
    if traffic and avg_s == avg_s:
        roofline["measured_gbs"] = round(traffic / avg_s / 1e9, 1)
        roofline["measured_frac"] = round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4)
    gbytes = GRADIENT_BYTES[op.absorbing]
    gradient_roof = {"bytes_per_cell_update": gbytes,
                     "effective_gbs": round(value * gbytes, 1),
                     "effective_frac": round(value * gbytes / HBM_PEAK_GBS, 4)}
    kernel_ms = {k: {"total_ms": round(v[0] / args.steps, 2), "launches": v[1] // args.steps,
                     "steps": units[k] // args.steps}
                 for k, v in timing.items()}
    misfit = None
    if not args.no_misfit:
        misfit = misfit_bench(rows, op.n_t, op.dt_obs, 0.5 * wl["kw"]["f0"], dev)
        print(f"[bench] misfit: {misfit}", file=sys.stderr, flush=True)
    grad_err = None
    if not args.no_accuracy:
        grad_err = gradient_error_vs_fp64(op, wl, c)
        print(f"[bench] accuracy of the timed plan: {grad_err}", file=sys.stderr, flush=True)
    comp = None
    if not args.no_autograd:
        op._cache.clear()  # release the HIP path's workspace: the comparator gets the device
        torch.cuda.empty_cache()
        comp = autograd_comparator(op, c, residual)
        for entry in comp.values():
            entry["speedup"] = round(entry["ms_per_gradient_scaled"] / ms_per_step, 1)
            entry["speedup_vs_min"] = round(entry["ms_per_gradient_scaled_min"] / ms_per_step, 1)
        print(f"[bench] autograd comparator: {comp}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, op)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)
    line = {
        "metric": "fwd+adjoint Gcell-updates/s (64-shot 2-D FWI gradient)",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "path": {"hip_graph": bool(plan.option(nat.OPT_GRAPH)),
                 "overlapped_reverse_sweep": os.environ.get("NEFI_AMD_OVERLAP", "1") != "0",
                 "snapshot_every": args.ckpt, "kept_blocks": kept_blocks},
        "kernel_timing_run": {
            "what": "eager calls, every launch bracketed by HIP events, one-stream reverse "
                    "sweep (the roofline and kernels_per_gradient come from this run)",
            "value": round(updates / (elapsed_eager / args.steps) / 1e9, 3),
            "ms_per_step": round(elapsed_eager * 1e3 / args.steps, 2)},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"2-D FWI gradient, {n}^2 grid (+20-cell {op.absorbing} = {padded[0]}^2), "
                        f"{args.shots} shots, src depth U[2,40] cells, receivers U[64,512]/shot, "
                        f"{op.n_steps} steps, snapshot every {args.ckpt}",
            "grid": list(padded),
            "shots": args.shots,
            "steps_per_gradient": op.n_steps,
            "receivers": rows,
            "cell_updates_per_step": updates,
        },
        "roofline": roofline,
        "gradient": gradient_roof,
        "kernels_per_gradient": kernel_ms,
        "tiles": tiles,
        "cpu_baseline": cpu,
        "autograd_gpu": comp,
        "gradient_error_vs_fp64": grad_err,
        "misfit": misfit,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
