"""Pins of the FP64 oracle (CPU): reference golden vectors, closed forms and invariants.

The oracle is the checker of the HIP kernels, so every part of it is pinned here by something
other than its own formulas: the reference's own float64 outputs (golden vectors written by
``tests/golden/make_golden.py``), the discrete energy identity, the adjoint dot-product test,
central finite differences, brute-force transposes on tiny grids and textbook special cases.
"""

from __future__ import annotations

import math
import os
from pathlib import Path

import numpy as np
import pytest

from _support import NAMES, known_c, load, oracle_problem, rel
from oracle.leapfrog import Problem


@pytest.mark.parametrize("name", NAMES)
def test_oracle_traces_match_reference(name):
    meta, z = load(name)
    P = oracle_problem(meta, z)
    assert P.substeps == int(z["f64_substeps"]) and P.n_steps == int(z["f64_n_steps"])
    assert math.isclose(P.dt, float(z["f64_dt"]), rel_tol=1e-15)
    if meta["kind"] == "fwi":
        tr = P.fwi_traces(z["in_c"])
    else:
        c, _ = known_c(meta, z)
        tr = P.ic_traces(c, z["in_p0"])[0]
    assert rel(tr, z["f64_traces"]) < 1e-12


@pytest.mark.parametrize("name", [n for n in NAMES if n != "cfg1"])
def test_oracle_gradient_matches_reference_autograd(name):
    meta, z = load(name)
    P = oracle_problem(meta, z)
    if meta["kind"] == "fwi":
        if meta["grad"] == "weights":
            g = P.fwi_gradient(z["in_c"], z["in_weights"])
        else:
            g = P.fwi_gradient(z["in_c"], P.fwi_traces(z["in_c"]) - z["f64_d_obs"])
    else:
        c, _ = known_c(meta, z)
        g = P.ic_gradient(c, z["in_weights"])[0]
    assert rel(g, z["f64_grad"]) < 1e-12


def _tiny(absorbing, order=4, d=2, **kw):
    shape = (9, 11) if d == 2 else (14,) if d == 1 else (6, 7, 5)
    extent = tuple((0.0, 0.1 * n) for n in shape)
    rec = [[0.1 * n * f for n in shape] for f in (0.3, 0.55, 0.8)]
    src = [[0.1 * n * f for n in shape] for f in (0.4, 0.62)]
    return Problem(shape, extent, rec, sources=src, n_t=12, dt_obs=0.02, f0=6.0, c_max=2.2,
                   order=order, absorbing=absorbing, absorb_width=0.2 if absorbing != "none" else None,
                   **kw)


@pytest.mark.parametrize("absorbing", ["none", "sponge", "cerjan", "pml"])
@pytest.mark.parametrize("d", [1, 2])
def test_oracle_gradient_equals_central_differences(absorbing, d):
    P = _tiny(absorbing, d=d)
    rng = np.random.default_rng(0)
    c = 1.8 + 0.3 * rng.random(P.shape)
    w = rng.standard_normal((2, 3, P.n_t))
    g = P.fwi_gradient(c, w)
    for _ in range(3):
        e = rng.standard_normal(P.shape)
        eps = 1e-6
        f = lambda x: float((P.fwi_traces(x) * w).sum())  # noqa: E731
        fd = (f(c + eps * e) - f(c - eps * e)) / (2 * eps)
        assert abs(fd - float((g * e).sum())) < 1e-7 * max(1.0, abs(fd))


def test_oracle_3d_sponge_gradient_equals_central_differences():
    P = _tiny("sponge", order=2, d=3)
    rng = np.random.default_rng(1)
    c = 1.8 + 0.3 * rng.random(P.shape)
    w = rng.standard_normal((2, 3, P.n_t))
    g = P.fwi_gradient(c, w)
    e = rng.standard_normal(P.shape)
    f = lambda x: float((P.fwi_traces(x) * w).sum())  # noqa: E731
    fd = (f(c + 1e-6 * e) - f(c - 1e-6 * e)) / 2e-6
    assert abs(fd - float((g * e).sum())) < 1e-7 * max(1.0, abs(fd))


@pytest.mark.parametrize("absorbing", ["none", "sponge", "cerjan", "pml"])
def test_oracle_initial_value_adjoint_dot_product(absorbing):
    """<A a, y> = <a, A^T y> for the linear initial-value operator (reference test
    tests/test_photoacoustic3d.py:49-59 states the same identity)."""
    shape, extent = (10, 12), ((0.0, 1.0), (0.0, 1.2))
    rec = [[0.2, 0.3], [0.8, 0.9], [0.5, 0.1]]
    P = Problem(shape, extent, rec, n_t=30, dt_obs=0.02, c_max=1.5, absorbing=absorbing,
                absorb_width=0.2 if absorbing != "none" else None)
    rng = np.random.default_rng(2)
    c = np.full(shape, 1.2) + 0.2 * rng.random(shape)
    a = rng.standard_normal(shape)
    y = rng.standard_normal((3, 30))
    Aa = P.ic_traces(c, a)[0]
    ATy = P.ic_gradient(c, y)[0]
    assert (Aa * y).sum() == pytest.approx((a * ATy).sum(), rel=1e-11)
    b = rng.standard_normal(shape)
    lin = P.ic_traces(c, 2 * a - 3 * b)[0] - (2 * Aa - 3 * P.ic_traces(c, b)[0])
    assert np.abs(lin).max() < 1e-12 * np.abs(Aa).max()


@pytest.mark.parametrize("order,d", [(2, 1), (4, 1), (2, 2), (4, 2), (4, 3)])
def test_oracle_laplacian_is_the_symmetric_cross_stencil(order, d):
    """Brute force on a tiny grid: the matrix of the stencil is symmetric, its rows are the
    coefficients of wave.py:121 divided by h², and a quadratic is differentiated exactly in
    the interior."""
    shape = {1: (9,), 2: (7, 8), 3: (6, 5, 7)}[d]
    P = Problem(shape, tuple((0.0, 0.3 * n) for n in shape), [[0.15] * d], n_t=2, dt_obs=0.01,
                c_max=1.0, absorbing="none", order=order)
    N = P.cells
    M = np.stack([P.laplacian(np.eye(N)[k]) for k in range(N)], 1)
    assert np.allclose(M, M.T, atol=0, rtol=1e-14)
    h = P.h
    x = np.stack(np.meshgrid(*[(np.arange(n) + 0.5) * hh for n, hh in zip(shape, h)],
                             indexing="ij"), -1)
    q = (x**2).sum(-1)
    lq = P.laplacian(q.reshape(-1)).reshape(shape)
    r = 2 if order == 4 else 1
    inner = tuple(slice(r, n - r) for n in shape)
    assert np.allclose(lq[inner], 2.0 * d, rtol=1e-9)
    # stability bound lambda_max <= sum rho / h^2 (timestep.py:177)
    rho = {2: 4.0, 4: 16.0 / 3.0}[order]
    assert np.abs(np.linalg.eigvalsh(M)).max() <= sum(rho / hh**2 for hh in h) * (1 + 1e-12)


def test_oracle_energy_is_conserved_without_absorption():
    """Discrete leapfrog energy (wave.py:171-189) is constant once the source has stopped."""
    shape, extent = (30, 30), ((0.0, 1.0), (0.0, 1.0))
    P = Problem(shape, extent, [[0.2, 0.2]], sources=[[0.5, 0.5]], n_t=100, dt_obs=0.01, f0=8.0,
                c_max=1.5, absorbing="none")
    c = np.full(shape, 1.0)
    c[15:, :] = 1.5
    m = P.model(c)
    energies = []
    for nt in (60, 75, 90, 100):
        Q = Problem(shape, extent, [[0.2, 0.2]], sources=[[0.5, 0.5]], n_t=nt, dt_obs=0.01,
                    f0=8.0, c_max=1.5, absorbing="none")
        _, st = Q.fwi_traces(c, final_state=True)
        prev, cur = st[0, 0], st[1, 0]
        kin = (cur - prev) ** 2 / m
        pot = -cur * Q.laplacian(prev)
        energies.append(0.5 * np.prod(Q.h) * (kin + pot).sum())
    e = np.array(energies)
    assert (e.max() - e.min()) / e.mean() < 1e-10


def test_oracle_ricker_and_cfl_closed_forms():
    from oracle.leapfrog import ricker

    t = np.linspace(0, 1, 2001)
    w = ricker(t, 5.0)
    assert abs(w.max() - 1.0) < 1e-6 and abs(t[w.argmax()] - 0.24) < 1e-3
    assert abs(w[0]) < 1e-4 and abs(w.sum() * (t[1] - t[0])) < 1e-4
    P = Problem((10, 10), ((0, 1), (0, 1)), [[0.5, 0.5]], n_t=3, dt_obs=1.0, c_max=1.0, order=2,
                courant=1.0, absorbing="none")
    assert math.isclose(P.dt, 1.0 / math.ceil(1.0 / (0.1 / math.sqrt(2)) - 1e-12))


def test_oracle_order2_1d_courant_one_is_an_exact_shift():
    """Order 2, 1-D, c·dt = h: leapfrog propagates a pulse by exactly one cell per step."""
    n = 60
    P = Problem((n,), ((0.0, float(n)),), [[30.5]], n_t=2, dt_obs=1.0, c_max=1.0, order=2,
                courant=1.0, absorbing="none")
    assert P.substeps == 1 and P.dt == 1.0
    c = np.ones(n)
    x = np.arange(n) + 0.5
    u0 = np.exp(-((x - 20.5) ** 2) / 4.0)
    u1 = np.exp(-((x - 21.5) ** 2) / 4.0)  # right-moving pulse one step later
    # run 10 steps from (u0, u1) through the initial-value machinery of the C loop
    from oracle import leapfrog as L

    m = P.model(c)
    prev, cur = u0.copy(), u1.copy()
    for _ in range(10):
        nxt = 2 * cur - prev + m * P.laplacian(cur)
        prev, cur = cur, nxt
    want = np.exp(-((x - 31.5) ** 2) / 4.0)
    inner = slice(5, n - 5)
    assert np.abs(cur[inner] - want[inner]).max() < 1e-12
    assert L.Problem is Problem


@pytest.mark.parametrize("absorbing", ["none", "sponge", "cerjan", "pml"])
@pytest.mark.parametrize("d", [1, 2, 3])
def test_oracle_checkpointed_gradient_is_bit_identical_to_store_all(absorbing, d):
    """Snapshots every K steps with per-block recomputation (K dividing the step count or not,
    K = 1, K = N) give exactly the bits of the store-all reverse sweep."""
    if d == 3 and absorbing == "pml":
        pytest.skip("the reference refuses a 3-D PML")
    P = _tiny(absorbing, order=2 if d == 3 else 4, d=d)
    rng = np.random.default_rng(5)
    c = 1.8 + 0.3 * rng.random(P.shape)
    w = rng.standard_normal((2, 3, P.n_t))
    g_all = P.fwi_gradient(c, w)
    for every in (1, 4, 5, P.n_steps - 1, P.n_steps, 10 * P.n_steps):
        assert np.array_equal(P.fwi_gradient(c, w, every=every), g_all), every


@pytest.mark.parametrize("name", ["grad_sponge", "pml2d", "cerjan2d"])
def test_oracle_checkpointed_gradient_matches_reference_autograd(name):
    meta, z = load(name)
    P = oracle_problem(meta, z)
    every = max(1, P.n_steps // 3 - 1)
    if meta["grad"] == "weights":
        g = P.fwi_gradient(z["in_c"], z["in_weights"], every=every)
    else:
        g = P.fwi_gradient(z["in_c"], P.fwi_traces(z["in_c"]) - z["f64_d_obs"], every=every)
    assert rel(g, z["f64_grad"]) < 1e-12


def test_oracle_results_do_not_depend_on_the_thread_count():
    """The OpenMP loops write disjoint cells and sum nothing across them: one thread and many
    threads give the same bits (PML, the layer with the most parallel loops)."""
    import subprocess
    import sys

    code = (
        "import sys, numpy as np; sys.path[:0] = [sys.argv[1], sys.argv[1] + '/tests'];"
        "import test_oracle as t;"
        "P = t._tiny('pml', d=2); rng = np.random.default_rng(2);"
        "c = 1.8 + 0.3 * rng.random(P.shape); w = rng.standard_normal((2, 3, P.n_t));"
        "g = P.fwi_gradient(c, w, every=4); y = P.fwi_traces(c);"
        "sys.stdout.write(np.concatenate([g.ravel(), y.ravel()]).tobytes().hex())"
    )
    root = str(Path(__file__).resolve().parent.parent)
    outs = []
    for threads in ("1", "7"):
        env = dict(os.environ, OMP_NUM_THREADS=threads)
        r = subprocess.run([sys.executable, "-c", code, root], env=env, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        outs.append(r.stdout)
    assert outs[0] == outs[1] and len(outs[0]) > 0


def _tr_case(name):
    import json

    from _support import GOLDEN

    z = np.load(GOLDEN / "time_reversal.npz", allow_pickle=False)
    meta = json.loads(str(z["meta"]))[name]
    traces = z[f"{name}_in_traces"]
    c = z[f"{name}_in_c"] if meta["c_map"] else np.full(tuple(meta["shape"]), 1.0)
    P = Problem(meta["shape"], meta["extent"], meta["receivers"], n_t=traces.shape[-1],
                c_max=float(np.max(c)), **meta["kw"])
    return z, P, traces, c


@pytest.mark.parametrize("name", ["pat2d", "pat3d"])
@pytest.mark.parametrize("mode", ["dirichlet", "adjoint"])
def test_oracle_time_reversal_matches_reference(name, mode):
    """The receiver clamp propagated from rest (Dirichlet mode) and the adjoint image, against
    the reference's own float64 time_reversal images (PML ring in 2-D, sponge in 3-D)."""
    z, P, traces, c = _tr_case(name)
    img = P.time_reversal(traces, c, mode)
    assert img.shape == tuple(z[f"{name}_{mode}_f64"].shape)
    assert rel(img, z[f"{name}_{mode}_f64"]) < 1e-12


@pytest.mark.parametrize("d", [1, 2, 3])
@pytest.mark.parametrize("order", [2, 4])
def test_oracle_laplacian_matches_reference(d, order):
    import json

    from _support import GOLDEN

    z = np.load(GOLDEN / "laplacian.npz", allow_pickle=False)
    spacing = json.loads(str(z["meta"]))[f"d{d}"]["spacing"]
    x = z[f"d{d}_in"]
    shape = x.shape[1:]
    P = Problem(shape, [(0.0, n * h) for n, h in zip(shape, spacing)],
                [[0.5 * n * h for n, h in zip(shape, spacing)]], n_t=2, dt_obs=1e-3,
                c_max=1.0, order=order, absorbing="none")
    assert P.padded == tuple(shape)
    out = P.laplacian(x.reshape(x.shape[0], -1)).reshape(x.shape)
    assert rel(out, z[f"d{d}_o{order}_f64"]) < 1e-13
