"""Neural-field full-waveform inversion: a coordinate MLP maps grid coordinates to sound speed,
the HIP propagator turns it into traces, and AdamW fits the traces of a seeded layered model.

This is the build's own driver for the 512² / 32-shot / 3000-step workload; nefi's field, head
and solver are restated here from their formulas, not imported:

* annealed Fourier features (nefi/fields/encoding.py:48-202): bands 2^k π, k < K, on
  coordinates in [-1, 1], gated by w_k = ½(1 − cos(π clamp(βK − k, 0, 1))) with β the
  annealing progress, raw coordinates prepended. "64 frequencies" maps to K = 8 octaves per
  axis (16 axis frequencies, 32 sin/cos features): bands above 2^8 π alias on 512 samples.
* the MLP (nefi/fields/neural.py:42-144): tanh hidden layers of width 256 with the encoding
  concatenated again at the middle layer, Xavier weights, zero biases, output weights × 0.1;
* the Bounded head (nefi/fields/heads.py:102-123): c = lo + (hi − lo) σ(h), lo = 1500,
  hi = 4500, output bias at the logit of the starting speed;
* the solver defaults (nefi/solve/curriculum.py:24-71, solver.py:344-473): AdamW, lr 1e-3 with
  cosine decay to 1 %, weight decay 1e-4, gradient norm clipped at 1, annealing over the run.

The misfit is the mean squared trace residual divided by the mean square of the observed
traces (a scale-free form of nefi's MSE: raw traces here are ~1e-8, which would put the
gradients far below Adam's epsilon).

``python -m nefi_amd.drivers.neural_fwi --iters 500`` runs the full workload on the GPU.
"""

from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch
from torch import nn

from ..domain import Domain
from ..physics.wave import WaveOperator

LO, HI = 1500.0, 4500.0


class FourierFeatures(nn.Module):
    """Annealed octave Fourier features of coordinates in [-1, 1]."""

    def __init__(self, in_dim: int, n_octaves: int = 8):
        super().__init__()
        self.in_dim, self.n_octaves = in_dim, n_octaves
        self.register_buffer("freqs", math.pi * 2.0 ** torch.arange(n_octaves, dtype=torch.float32))
        self.out_dim = in_dim + 2 * in_dim * n_octaves

    def gates(self, progress: float) -> torch.Tensor:
        k = torch.arange(self.n_octaves, dtype=torch.float32, device=self.freqs.device)
        beta = progress * self.n_octaves
        return 0.5 * (1.0 - torch.cos(math.pi * torch.clamp(beta - k, 0.0, 1.0)))

    def forward(self, x: torch.Tensor, progress: float) -> torch.Tensor:
        ang = x.unsqueeze(-2) * self.freqs.view(-1, 1)                 # (N, K, D)
        sc = torch.cat([torch.sin(ang), torch.cos(ang)], dim=-1)      # (N, K, 2D)
        sc = sc * self.gates(progress).view(-1, 1)
        return torch.cat([x, sc.flatten(-2)], dim=-1)


class NeuralSpeed(nn.Module):
    """Coordinate MLP with a mid-network skip and a bounded sound-speed head."""

    def __init__(self, ndim: int = 2, hidden: int = 256, depth: int = 4, n_octaves: int = 8,
                 init_speed: float = 3000.0):
        super().__init__()
        self.enc = FourierFeatures(ndim, n_octaves)
        e = self.enc.out_dim
        self.skip_at = depth // 2
        dims = [e] + [hidden] * depth
        self.layers = nn.ModuleList(
            nn.Linear(hidden + e if i == self.skip_at else dims[i], hidden) for i in range(depth))
        self.out = nn.Linear(hidden, 1)
        for layer in self.layers:
            nn.init.xavier_uniform_(layer.weight)
            nn.init.zeros_(layer.bias)
        nn.init.xavier_uniform_(self.out.weight)
        u = (init_speed - LO) / (HI - LO)
        with torch.no_grad():
            self.out.weight.mul_(0.1)
            self.out.bias.fill_(math.log(u / (1.0 - u)))

    def forward(self, coords: torch.Tensor, progress: float) -> torch.Tensor:
        e = self.enc(coords, progress)
        h = e
        for i, layer in enumerate(self.layers):
            if i == self.skip_at:
                h = torch.cat([h, e], dim=-1)
            h = torch.tanh(layer(h))
        return LO + (HI - LO) * torch.sigmoid(self.out(h)[..., 0])


def layered_model(n: int = 512, dx: float = 10.0, seed: int = 0) -> np.ndarray:
    """Six layers with undulating interfaces (amplitude 20 cells) and two Gaussian inclusions,
    axis 0 horizontal, axis 1 depth, clipped to [1500, 4500] m/s."""
    L = n * dx
    rng = np.random.default_rng(seed)
    xs = (np.arange(n) + 0.5) * dx
    zs = (np.arange(n) + 0.5) * dx
    speeds = [1600.0, 2000.0, 2500.0, 3000.0, 3600.0, 4200.0]
    c = np.full((n, n), speeds[0])
    for k in range(1, 6):
        depth = k * L / 6 + 20 * dx * np.sin(2 * np.pi * xs / L * (k + 1) + rng.uniform(0, 6))
        c[zs[None, :] > depth[:, None]] = speeds[k]
    for cx, cz, a in ((0.3 * L, 0.4 * L, 400.0), (0.7 * L, 0.65 * L, -300.0)):
        c += a * np.exp(-((xs[:, None] - cx) ** 2 + (zs[None, :] - cz) ** 2) / (2 * (n / 20.48 * dx) ** 2))
    return np.clip(c, LO, HI)


def survey(n: int = 512, shots: int = 32, receivers: int = 256, n_t: int = 3001,
           dx: float = 10.0) -> WaveOperator:
    """Sources evenly spaced half a cell below the top, receivers along the top row; 20-cell
    sponge, order 4, CFL 0.4 at 4500 m/s."""
    L = n * dx
    src = [[(i + 0.5) * L / shots, 0.5 * dx] for i in range(shots)]
    rec = [[(i + 0.5) * L / receivers, 0.5 * dx] for i in range(receivers)]
    return WaveOperator(Domain((n, n), ((0.0, L), (0.0, L))), src, rec, n_t=n_t,
                        dt_obs=0.4 * dx / HI, f0=15.0, c_max=HI, order=4, absorbing="sponge",
                        absorb_width=20 * dx, grad_mode="checkpoint")


def run(iters: int = 500, n: int = 512, shots: int = 32, receivers: int = 256,
        n_t: int = 3001, seed: int = 0, device: str = "cuda", lr: float = 1e-3,
        log_every: int = 10, out=sys.stderr) -> dict:
    """Fit the traces of layered_model(n) from a uniform 3000 m/s start; returns the misfit
    history, the model error and timings."""
    torch.manual_seed(seed)
    op = survey(n, shots, receivers, n_t)
    c_true = torch.tensor(layered_model(n, seed=seed), dtype=torch.float32, device=device)
    with torch.no_grad():
        obs = op({"c": c_true})
    scale = float((obs ** 2).mean())
    ax = (torch.arange(n, dtype=torch.float32, device=device) + 0.5) / n * 2.0 - 1.0
    coords = torch.stack(torch.meshgrid(ax, ax, indexing="ij"), -1).reshape(-1, 2)
    field = NeuralSpeed().to(device)
    with torch.no_grad():
        c_start = field(coords, 0.0).reshape(n, n)
        start_err = float(torch.linalg.norm(c_start - c_true) / torch.linalg.norm(c_true))
    opt = torch.optim.AdamW(field.parameters(), lr=lr, weight_decay=1e-4)
    history, times = [], []
    for step in range(iters):
        t0 = time.perf_counter()
        progress = min(1.0, step / max(1.0, iters))
        for group in opt.param_groups:
            group["lr"] = lr * (0.01 + 0.99 * 0.5 * (1.0 + math.cos(math.pi * step / max(1, iters))))
        c = field(coords, progress).reshape(n, n)
        loss = ((op({"c": c}) - obs) ** 2).mean() / scale
        opt.zero_grad(set_to_none=True)
        loss.backward()
        torch.nn.utils.clip_grad_norm_(field.parameters(), 1.0)
        opt.step()
        history.append(float(loss.detach()))
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        if log_every and (step % log_every == 0 or step == iters - 1):
            print(f"[neural_fwi] iter {step}: misfit {history[-1]:.4e} ({times[-1]:.3f} s)",
                  file=out, flush=True)
    with torch.no_grad():
        c = field(coords, 1.0).reshape(n, n)
        model_err = float(torch.linalg.norm(c - c_true) / torch.linalg.norm(c_true))
    steady = times[1:] if len(times) > 1 else times
    s_iter = float(np.mean(steady))
    cells = op.grid.padded_shape[0] * op.grid.padded_shape[1]
    return {"misfit": history, "model_rel_error": model_err,
            "model_rel_error_start": start_err, "s_per_iter": s_iter,
            "gcell_updates_per_s": cells * op.n_steps * shots / s_iter / 1e9,
            "padded": list(op.grid.padded_shape), "steps": op.n_steps, "shots": shots}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--iters", type=int, default=500)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--shots", type=int, default=32)
    ap.add_argument("--receivers", type=int, default=256)
    ap.add_argument("--nt", type=int, default=3001)
    args = ap.parse_args(argv)
    res = run(args.iters, args.n, args.shots, args.receivers, args.nt)
    res["misfit_first_last"] = [res["misfit"][0], res["misfit"][-1]]
    res["misfit_every_50"] = res["misfit"][::50]
    del res["misfit"]
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
