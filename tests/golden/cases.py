"""Seeded test cases shared by the golden-vector writer and the parity tests.

Each case is plain data (domain, geometry, operator keywords) plus a seeded input generator, so
the reference (``make_golden.py``), the FP64 oracle and the HIP operators all see identical
inputs. Sizes are small enough for the reference's CPU autograd to finish in seconds, except
``cfg1`` which is the 128² / 1000-sample forward configuration itself.
"""

from __future__ import annotations

import math

import numpy as np


def _gauss_anomalies(shape, extent, rng, n, amp, sigma_cells):
    """``n`` Gaussian blobs of amplitude ±amp, centres U[0.2, 0.8]·L."""
    axes = [lo + (np.arange(k) + 0.5) * (hi - lo) / k for (lo, hi), k in zip(extent, shape)]
    mesh = np.stack(np.meshgrid(*axes, indexing="ij"), -1)
    h = (extent[0][1] - extent[0][0]) / shape[0]
    out = np.zeros(shape)
    for _ in range(n):
        ctr = np.array([lo + rng.uniform(0.2, 0.8) * (hi - lo) for lo, hi in extent])
        sign = rng.choice([-1.0, 1.0])
        r2 = ((mesh - ctr) ** 2).sum(-1)
        out += sign * amp * np.exp(-r2 / (2 * (sigma_cells * h) ** 2))
    return out


def _smooth_noise(shape, rng, sigma_cells, amp):
    from scipy.ndimage import gaussian_filter

    x = gaussian_filter(rng.standard_normal(shape), sigma_cells, mode="wrap")
    return amp * x / np.abs(x).max()


def cfg1():
    """128², dx = 10 m, sponge 20 cells, one 15 Hz Ricker at (640, 5), 32 receivers at depth 5,
    dt_obs = 0.4·dx/c_max (one leapfrog step per sample), 1000 samples (999 steps)."""
    n, dx = 128, 10.0
    L = n * dx
    rng = np.random.default_rng(0)
    c = 1500.0 + _gauss_anomalies((n, n), ((0, L), (0, L)), rng, 3, 300.0, 8.0)
    cmax = float(c.max())
    rec = [[(i + 0.5) * 40.0, 5.0] for i in range(32)]
    return dict(
        name="cfg1", kind="fwi", shape=(n, n), extent=((0.0, L), (0.0, L)),
        sources=[[640.0, 5.0]], receivers=rec,
        kw=dict(n_t=1000, dt_obs=0.4 * dx / cmax, f0=15.0, c_max=cmax, order=4,
                absorbing="sponge", absorb_width=200.0),
        c=c, grad=False,
    )


def grad_sponge():
    """Reduced config-2 geometry: 48², 3 shots, 16 receivers, linear-in-depth model plus a
    smooth perturbation; L2 misfit against data of the perturbed model."""
    n, dx = 48, 10.0
    L = n * dx
    rng = np.random.default_rng(1)
    z = (np.arange(n) + 0.5) / n
    c0 = np.broadcast_to(2000.0 + 1500.0 * z[None, :], (n, n)).copy()
    c_true = c0 + _smooth_noise((n, n), rng, 4.0, 100.0)
    cmax = 3700.0
    src = [[(i + 0.5) * L / 3, 25.0] for i in range(3)]
    rec = [[(i + 0.5) * L / 16, 15.0] for i in range(16)]
    return dict(
        name="grad_sponge", kind="fwi", shape=(n, n), extent=((0.0, L), (0.0, L)),
        sources=src, receivers=rec,
        kw=dict(n_t=151, dt_obs=0.0015, f0=25.0, c_max=cmax, order=4, absorbing="sponge",
                absorb_width=100.0, grad_mode="checkpoint", checkpoint_every=37),
        c=c0, c_true=c_true, grad="misfit",
    )


def _small(name, absorbing, order, width, dim=2, seed=2, n_t=60, dt_obs=0.02, grad_mode="autograd"):
    rng = np.random.default_rng(seed)
    if dim == 1:
        shape, extent = (40,), ((0.0, 1.0),)
        src, rec = [[0.3], [0.55]], [[0.2], [0.75], [0.9]]
    else:
        shape, extent = (20, 24), ((0.0, 1.0), (0.0, 1.2))
        src = [[0.2, 0.3], [0.25, 0.8]]
        rec = [[0.8, y] for y in (0.2, 0.45, 0.7, 1.0)] + [[0.05, 0.6]]
    c = np.clip(2.0 + 0.2 * rng.standard_normal(shape), 1.6, 2.4)
    w = rng.standard_normal((len(src), len(rec), n_t))
    return dict(
        name=name, kind="fwi", shape=shape, extent=extent, sources=src, receivers=rec,
        kw=dict(n_t=n_t, dt_obs=dt_obs, f0=5.0, c_max=2.5, order=order, absorbing=absorbing,
                absorb_width=width, grad_mode=grad_mode),
        c=c, weights=w, grad="weights",
    )


def pml2d():
    return _small("pml2d", "pml", 4, 0.2, seed=2)


def cerjan2d():
    return _small("cerjan2d", "cerjan", 4, 0.25, seed=3)


def none2d_o2():
    return _small("none2d_o2", "none", 2, None, seed=4)


def sponge2d_o2():
    return _small("sponge2d_o2", "sponge", 2, 0.2, seed=5, grad_mode="checkpoint")


def pml1d_o2():
    return _small("pml1d_o2", "pml", 2, 0.15, dim=1, seed=6, n_t=80, dt_obs=0.01)


def pat2d():
    """Initial-value operator on a ring of 24 receivers (PML), the shape of the reference's
    linearity / time-reversal test at a smaller size."""
    n = 24
    th = 2 * math.pi * np.arange(24) / 24
    rec = np.stack([0.9 * np.cos(th), 0.9 * np.sin(th)], 1)
    rng = np.random.default_rng(7)
    ax = -1.0 + (2 * np.arange(n) + 1.0) / n
    X, Y = np.meshgrid(ax, ax, indexing="ij")
    p0 = np.exp(-((X - 0.3) ** 2 + (Y + 0.2) ** 2) / (2 * 0.1**2)) + 0.1 * rng.random((n, n))
    w = rng.standard_normal((24, 90))
    return dict(
        name="pat2d", kind="ic", shape=(n, n), extent=((-1.0, 1.0), (-1.0, 1.0)),
        receivers=rec.tolist(),
        kw=dict(n_t=90, dt_obs=0.02, c=1.0, absorb_width=0.25),
        p0=p0, weights=w, grad="weights",
    )


def pat3d():
    """3-D initial-value operator, sponge, heterogeneous known sound speed, 16 detectors on
    the top face (the shape of the photoacoustic3d instance at a tiny size)."""
    shape = (10, 10, 8)
    extent = ((0.0, 1.0), (0.0, 1.0), (0.0, 0.8))
    rng = np.random.default_rng(8)
    c = 1.5 * (1.0 + 0.05 * _smooth_noise(shape, rng, 2.0, 1.0))
    p0 = rng.random(shape)
    g = (np.arange(4) + 0.5) / 4
    rec = [[x, y, 0.0] for x in g for y in g]
    w = rng.standard_normal((16, 40))
    return dict(
        name="pat3d", kind="ic", shape=shape, extent=extent, receivers=rec,
        kw=dict(n_t=40, dt_obs=0.02, c=c, absorbing="sponge", absorb_width=0.2, order=2),
        p0=p0, weights=w, grad="weights",
    )


def cerjan_pat2d():
    """Initial-value operator with the Cerjan taper (exercises the tapered-previous-state rule
    from the very first step)."""
    case = pat2d()
    case.update(name="cerjan_pat2d")
    case["kw"] = dict(n_t=90, dt_obs=0.02, c=1.0, absorb_width=0.3, absorbing="cerjan")
    return case


ALL = [cfg1, grad_sponge, pml2d, cerjan2d, none2d_o2, sponge2d_o2, pml1d_o2, pat2d, pat3d,
       cerjan_pat2d]
