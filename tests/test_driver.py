"""The config-5 neural-field driver: its field pieces against their closed forms (CPU) and a
short inversion that lowers the trace misfit (GPU)."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from nefi_amd.drivers import neural_fwi as nf


def test_fourier_gates_follow_the_cosine_ramp():
    enc = nf.FourierFeatures(2, 8)
    for progress in (0.0, 0.3, 0.5, 0.77, 1.0):
        beta = progress * 8
        want = [0.5 * (1 - math.cos(math.pi * min(1.0, max(0.0, beta - k)))) for k in range(8)]
        assert np.allclose(enc.gates(progress).numpy(), want, atol=1e-7)
    x = torch.tensor([[0.25, -0.5]])
    f = enc(x, 1.0)
    assert f.shape == (1, 2 + 2 * 2 * 8)
    k = 3  # band 3: sin and cos of 2^3 pi x, both axes, after the raw coordinates
    ang = math.pi * 2**k * np.array([0.25, -0.5])
    assert np.allclose(f[0, 2 + 4 * k: 2 + 4 * k + 4].numpy(),
                       np.concatenate([np.sin(ang), np.cos(ang)]), atol=1e-5)


def test_bounded_head_starts_at_the_initial_speed_and_stays_in_bounds():
    torch.manual_seed(0)
    field = nf.NeuralSpeed(init_speed=2500.0)
    with torch.no_grad():
        c = field(torch.rand(1000, 2) * 2 - 1, 0.5)
    assert abs(float(c.mean()) - 2500.0) < 60.0
    with torch.no_grad():
        field.out.bias.fill_(50.0)
        assert float(field(torch.zeros(4, 2), 1.0).max()) <= nf.HI
        field.out.bias.fill_(-50.0)
        assert float(field(torch.zeros(4, 2), 1.0).min()) >= nf.LO


def test_layered_model_bounds_and_layers():
    c = nf.layered_model(128)
    assert c.shape == (128, 128) and c.min() >= nf.LO and c.max() <= nf.HI
    assert c[:, 2].mean() < c[:, -3].mean()  # speed grows with depth


@pytest.mark.gpu
def test_short_inversion_lowers_the_misfit():
    res = nf.run(iters=12, n=96, shots=4, receivers=48, n_t=301, log_every=4)
    h = res["misfit"]
    print(f"misfit {h[0]:.4e} -> {h[-1]:.4e}, model error {res['model_rel_error']:.3f}, "
          f"{res['s_per_iter']:.3f} s per iteration", flush=True)
    assert all(np.isfinite(h)) and h[-1] < 0.8 * h[0]
