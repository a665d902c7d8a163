"""The HIP trace filters and fused misfit against the FP64 oracle.

Every path goes through ``nw_filter_apply``: ``lowpass`` and ``gaussian_lowpass`` (float32 within
1e-5 rel-L2 of the float64 oracle, float64 within 1e-12), their backward (self-adjoint: the
autograd dot-product test), ``BandLimitedMSE`` value and gradient (adjoint source) with and
without masks, filters and normalisation, ``SensorBandwidth`` around the 2-D initial-value wave
operator, and the misfit of a survey the size of the 64-shot 1024² benchmark (18 432 traces of
4001 samples; rows sampled against the oracle).
"""

from __future__ import annotations

import time

import numpy as np
import pytest
import torch

from _support import rel
from nefi_amd import _native as nat
from nefi_amd.instances.photoacoustic3d import SensorBandwidth, gaussian_lowpass
from nefi_amd.instances.wave_fwi import BandLimitedMSE, lowpass
from oracle import filters as of

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


class _Obs:
    def __init__(self, data, mask=None):
        self.data, self.mask = data, mask


class _Ctx:
    def __init__(self, pred, obs):
        self.pred, self.obs = pred, obs


@pytest.mark.parametrize("shape", [(3, 37), (5, 1000), (2, 3, 2001), (7, 4001), (1, 1), (4, 64)])
def test_lowpass_matches_oracle(shape):
    print(f"[case lowpass {shape}]", flush=True)
    rng = np.random.default_rng(sum(shape))
    x = rng.standard_normal(shape)
    dt, f_max, taper = 0.002, 25.0, 0.3
    want = of.lowpass(x, dt, f_max, taper)
    for dtype, tol in ((torch.float32, 1e-5), (torch.float64, 1e-12)):
        y = lowpass(torch.tensor(x, dtype=dtype, device=DEV), dt, f_max, taper)
        assert y.shape == x.shape and y.dtype == dtype
        e = rel(y.double().cpu().numpy(), want)
        print(f"lowpass {shape} {dtype}: rel-L2 {e:.2e}", flush=True)
        assert e < tol


@pytest.mark.parametrize("n", [50, 1501, 3001])
def test_gaussian_lowpass_matches_oracle(n):
    print(f"[case gaussian_lowpass n={n}]", flush=True)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((6, n))
    want = of.gaussian_lowpass(x, 0.01, 4.0)
    for dtype, tol in ((torch.float32, 1e-5), (torch.float64, 1e-12)):
        y = gaussian_lowpass(torch.tensor(x, dtype=dtype, device=DEV), 0.01, 4.0)
        e = rel(y.double().cpu().numpy(), want)
        print(f"gaussian n={n} {dtype}: rel-L2 {e:.2e}", flush=True)
        assert e < tol
    xt = torch.tensor(x, device=DEV)
    assert gaussian_lowpass(xt, 0.01, None) is xt


def test_filter_backward_is_the_adjoint_and_cpu_inputs_come_back_on_cpu():
    rng = np.random.default_rng(4)
    x = torch.tensor(rng.standard_normal((9, 777)), device=DEV, requires_grad=True)
    w = torch.tensor(rng.standard_normal((9, 777)), device=DEV)
    y = lowpass(x, 0.004, 30.0)
    (g,) = torch.autograd.grad((y * w).sum(), x)
    lhs, rhs = float((y.detach() * w).sum()), float((x.detach() * g).sum())
    print(f"<Fx, w> {lhs:.12e} <x, F w> {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    xc = torch.tensor(rng.standard_normal((2, 300)))  # CPU float64 (nefi's data generators)
    yc = lowpass(xc, 0.004, 30.0)
    assert yc.device.type == "cpu" and yc.dtype == torch.float64
    assert rel(yc.numpy(), of.lowpass(xc.numpy(), 0.004, 30.0)) < 1e-12


@pytest.mark.parametrize("f_max,relative,mask_kind", [(15.0, True, "broadcast"),
                                                      (15.0, False, None),
                                                      (None, True, "full"),
                                                      (None, False, None),
                                                      (40.0, True, "bool")])
def test_band_limited_mse_value_and_adjoint_source(f_max, relative, mask_kind):
    print(f"[case misfit f_max={f_max} relative={relative} mask={mask_kind}]", flush=True)
    rng = np.random.default_rng(5)
    S, R, n = 3, 17, 1201
    pred, obs = rng.standard_normal((2, S, R, n))
    mask = None
    if mask_kind == "broadcast":
        mask = (rng.random((S, R, 1)) > 0.3).astype(float)
    elif mask_kind == "full":
        mask = rng.random((S, R, n))
    elif mask_kind == "bool":
        mask = rng.random((S, R, 1)) > 0.5
    want, gwant = of.band_limited_mse(pred, obs, None if mask is None else mask.astype(float),
                                      0.002, f_max, 0.3, relative)
    for dtype, tol in ((torch.float32, 1e-5), (torch.float64, 1e-11)):
        p = torch.tensor(pred, dtype=dtype, device=DEV, requires_grad=True)
        o = torch.tensor(obs, dtype=dtype, device=DEV)
        m = None if mask is None else torch.tensor(mask, device=DEV)
        if m is not None and m.dtype != torch.bool:
            m = m.to(dtype)
        loss = BandLimitedMSE(f_max, 0.002, relative=relative)(_Ctx(p, _Obs(o, m)))
        (g,) = torch.autograd.grad(loss, p)
        el = abs(float(loss) - want) / abs(want)
        eg = rel(g.double().cpu().numpy(), gwant)
        print(f"{dtype}: loss {float(loss):.9e} (oracle {want:.9e}, rel {el:.1e}), "
              f"gradient rel-L2 {eg:.2e}", flush=True)
        assert loss.dtype == dtype and el < tol and eg < tol


@pytest.mark.parametrize("n", [1, 33, 100, 200, 400, 600, 1500, 3000, 4096, 5000, 8192])
def test_every_transform_length(n):
    """Every FFT length of the plans (L = 64 .. 16384 for float32, .. 8192 for float64: each has
    its own pass plan), through the fused misfit (residual, full mask, sums) with an odd row
    count (the last row of the pair is absent) and through the plain filter."""
    print(f"[case transform length n={n}]", flush=True)
    rng = np.random.default_rng(n + 11)
    rows = 5
    pred, obs = rng.standard_normal((2, rows, n))
    mask = rng.random((rows, n))
    want, gwant = of.band_limited_mse(pred, obs, mask, 0.002, 60.0, 0.3, True)
    ywant = of.lowpass(pred, 0.002, 60.0, 0.3)
    for dtype, tol in ((torch.float32, 1e-5), (torch.float64, 1e-11)):
        if dtype == torch.float64 and n > 4096:
            continue
        p = torch.tensor(pred, dtype=dtype, device=DEV, requires_grad=True)
        o = torch.tensor(obs, dtype=dtype, device=DEV)
        m = torch.tensor(mask, dtype=dtype, device=DEV)
        loss = BandLimitedMSE(60.0, 0.002)(_Ctx(p, _Obs(o, m)))
        (g,) = torch.autograd.grad(loss, p)
        y = lowpass(p.detach(), 0.002, 60.0)
        el = abs(float(loss) - want) / abs(want)
        eg = rel(g.double().cpu().numpy(), gwant)
        ey = rel(y.double().cpu().numpy(), ywant)
        print(f"n={n} {dtype}: loss rel {el:.1e}, gradient rel-L2 {eg:.2e}, lowpass {ey:.2e}",
              flush=True)
        assert el < tol and eg < tol and ey < tol


def test_misfit_sums_are_deterministic():
    rng = np.random.default_rng(6)
    p = torch.tensor(rng.standard_normal((300, 2001)), device=DEV, dtype=torch.float32)
    o = torch.tensor(rng.standard_normal((300, 2001)), device=DEV, dtype=torch.float32)
    loss = BandLimitedMSE(20.0, 0.002)
    vals = [loss(_Ctx(p, _Obs(o))).item() for _ in range(3)]
    assert vals[0] == vals[1] == vals[2]


def test_sensor_bandwidth_around_the_initial_value_operator():
    from nefi_amd.domain import Domain
    from nefi_amd.physics.wave import WaveInitialConditionOperator

    n, h = 48, 1e-4
    L = n * h
    det = [[(i + 0.5) * L / 12, 0.5 * h] for i in range(12)]
    inner = WaveInitialConditionOperator(Domain((n, n), ((0.0, L), (0.0, L))), det, n_t=301,
                                         dt_obs=0.3 * h / 1500.0, c=1500.0, c_max=1500.0,
                                         order=4, absorbing="sponge", absorb_width=8 * h)
    f_c = 0.05 / inner.dt_obs
    op = SensorBandwidth(inner, inner.dt_obs, f_c)
    assert op.fidelity_tag.startswith(inner.fidelity_tag) and op.primary == "p0"
    rng = np.random.default_rng(7)
    p0 = torch.tensor(rng.standard_normal((n, n)), device=DEV, requires_grad=True)
    w = torch.tensor(rng.standard_normal((12, 301)), device=DEV)
    y = op({"p0": p0})
    y_in = inner({"p0": p0.detach()})
    assert rel(y.detach().cpu().numpy(),
               of.gaussian_lowpass(y_in.cpu().numpy(), inner.dt_obs, f_c)) < 1e-12
    (g,) = torch.autograd.grad((y * w).sum(), p0)
    lhs, rhs = float((y.detach() * w).sum()), float((p0.detach() * g).sum())
    print(f"SensorBandwidth dot product: {lhs:.12e} vs {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


def test_misfit_at_the_benchmark_survey_size():
    """18 432 traces (64 shots x 288 receivers) x 4001 samples, float32: rows sampled against
    the oracle; the fused forward and the adjoint-source pass timed with HIP events."""
    rng = torch.Generator(DEV).manual_seed(8)
    rows, n = 64 * 288, 4001
    pred = torch.randn(rows, n, device=DEV, generator=rng)
    obs = torch.randn(rows, n, device=DEV, generator=rng)
    mask = (torch.rand(rows, 1, device=DEV, generator=rng) > 0.1).float()
    pred.requires_grad_(True)
    loss_fn = BandLimitedMSE(12.0, 0.001)
    for _ in range(2):  # warm (plan creation)
        (g,) = torch.autograd.grad(loss_fn(_Ctx(pred, _Obs(obs, mask))), pred)
    torch.cuda.synchronize()
    nat.profile_enable(True)
    t0 = time.perf_counter()
    loss = loss_fn(_Ctx(pred, _Obs(obs, mask)))
    (g,) = torch.autograd.grad(loss, pred)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms, launches = nat.profile_read(nat.K_FILTER)
    nat.profile_enable(False)
    print(f"misfit + adjoint source, {rows} x {n}: {wall * 1e3:.2f} ms wall, filter kernels "
          f"{ms:.2f} ms over {launches} launches", flush=True)
    # the oracle on 40 sampled rows (the misfit couples rows only through the global sums)
    idx = torch.randperm(rows, generator=torch.Generator().manual_seed(9))[:40]
    P, O, M = (t.detach()[idx].double().cpu().numpy() for t in (pred, obs, mask))
    full_sums = [float(v) for v in _sums(pred.detach(), obs, mask, 12.0, 0.001)]
    # value: the sums of all rows in float64 on the GPU vs the kernel's
    want_loss = full_sums[0] / full_sums[1]
    print(f"loss {float(loss):.9e} vs float64 sums {want_loss:.9e}", flush=True)
    assert abs(float(loss) - want_loss) < 1e-5 * abs(want_loss)
    # gradient rows: oracle gradient of the sampled rows, rescaled to the full normalisation
    r = of.lowpass(P - O, 0.001, 12.0)
    g_rows = of.lowpass(2.0 * M * r, 0.001, 12.0)
    msum, den = full_sums[2], full_sums[3]
    g_want = g_rows / (msum * den)
    e = rel(g[idx].double().cpu().numpy(), g_want)
    print(f"adjoint source rows vs oracle: rel-L2 {e:.2e}", flush=True)
    assert e < 1e-5


def _sums(pred, obs, mask, f_max, dt):
    """num, 1, sum(mask), den of the benchmark misfit in float64 (torch FFT on the GPU, used
    only as a reference value)."""
    n = pred.shape[-1]
    from nefi_amd.physics.filters import cosine_response, next_fast_len

    n_fft = next_fast_len(2 * n)
    w = torch.tensor(cosine_response(n_fft, dt, f_max, 0.3), device=DEV)

    def lp(x):
        return torch.fft.irfft(torch.fft.rfft(x.double(), n=n_fft, dim=-1) * w, n=n_fft,
                               dim=-1)[..., :n]

    m = mask.double().expand_as(pred)
    msum = m.sum()
    num = (lp(pred - obs) ** 2 * m).sum() / msum
    den = (lp(obs) ** 2 * m).sum() / msum
    return num / den, 1.0, msum, den
