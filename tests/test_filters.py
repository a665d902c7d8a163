"""The trace-filter oracle pinned without the GPU, and the host side of the filter plans.

The oracle (``oracle/filters.py``) restates nefi's FFT low-passes and band-limited misfit; it is
checked here against the reference's own filter tests, a brute-force direct convolution with
taps from a cosine-sum inverse DFT, the cosine window's closed form at its roll-off midpoint,
self-adjointness, and finite differences of the misfit. The product's host-side pieces (FFT
length rule, responses, taps) are checked against the oracle and scipy; the C ABI's argument
checks run without a GPU.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
import scipy.fft

from nefi_amd import _native as nat
from nefi_amd.physics import filters as pf
from oracle import filters as of


def test_reference_lowpass_test_passes_on_the_oracle():
    # tests/test_instances_wave_family.py:49-54
    dt = 0.004
    t = np.arange(500) * dt
    low, high = np.sin(2 * math.pi * 2.0 * t), np.sin(2 * math.pi * 20.0 * t)
    y = of.lowpass(low + high, dt, f_max=6.0)
    core = slice(100, 400)
    assert np.abs(y[core] - low[core]).max() < 0.05


def test_reference_gaussian_lowpass_test_passes_on_the_oracle():
    # tests/test_photoacoustic3d.py:40-46
    t = np.arange(200) * 0.05
    slow, fast = np.sin(2 * np.pi * 0.2 * t), np.sin(2 * np.pi * 4.0 * t)
    out = of.gaussian_lowpass(np.stack([slow, fast]), 0.05, 1.0)
    assert np.abs((out[0] - slow)[20:-20]).max() < 0.05
    assert np.abs(out[1][20:-20]).max() < 1e-3
    assert of.gaussian_lowpass(slow, 0.05, None) is slow


@pytest.mark.parametrize("n", [7, 37, 100, 251])
def test_oracle_filters_equal_direct_convolution_with_cosine_sum_taps(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal((3, n))
    dt, f_max, taper = 0.01, 12.0, 0.3
    n_fft = scipy.fft.next_fast_len(2 * n, real=True)
    w = of._window(n_fft, dt, f_max, taper)
    y_direct = of.toeplitz_filter(x, of.taps_direct(w, n_fft, n))
    assert np.abs(of.lowpass(x, dt, f_max, taper) - y_direct).max() < 1e-12 * np.abs(x).max()
    f = np.fft.rfftfreq(2 * n, d=dt)
    g = np.exp(-0.5 * (f / 9.0) ** 2)
    y_direct = of.toeplitz_filter(x, of.taps_direct(g, 2 * n, n))
    assert np.abs(of.gaussian_lowpass(x, dt, 9.0) - y_direct).max() < 1e-12 * np.abs(x).max()


def test_odd_fft_lengths_occur_and_are_covered():
    # N = next_fast_len(2n) is odd for some n (e.g. 2n = 14 -> 15): the cosine-sum taps and
    # the product's taps take the odd-N branch there
    assert scipy.fft.next_fast_len(14, real=True) == 15
    w = of._window(15, 0.01, 12.0, 0.3)
    assert np.allclose(of.taps_direct(w, 15, 7), np.fft.irfft(w, 15)[:7], atol=1e-15)


def test_window_midpoint_halves_a_long_sinusoid():
    # the roll-off is 0.5 (1 + cos(pi (f - f1) / (f_max - f1))): 1/2 at its midpoint, 1 below
    # f1, 0 above f_max
    dt, f_max, taper, n = 0.002, 20.0, 0.4, 8000
    f1 = (1 - taper) * f_max
    t = np.arange(n) * dt
    core = slice(2000, 6000)
    for f, want in ((0.5 * (f1 + f_max), 0.5), (0.5 * f1, 1.0), (1.5 * f_max, 0.0)):
        y = of.lowpass(np.sin(2 * np.pi * f * t), dt, f_max, taper)
        amp = np.sqrt(2.0 * np.mean(y[core] ** 2))
        assert abs(amp - want) < 0.01, (f, amp)


def test_oracle_filters_are_self_adjoint():
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal((2, 4, 300))
    for F in (lambda v: of.lowpass(v, 0.004, 30.0, 0.3), lambda v: of.gaussian_lowpass(v, 0.004, 25.0)):
        a, b = (F(x) * y).sum(), (x * F(y)).sum()
        assert abs(a - b) < 1e-12 * abs(a)


@pytest.mark.parametrize("f_max,relative,masked", [(15.0, True, True), (15.0, False, False),
                                                   (None, True, True), (None, False, True)])
def test_oracle_misfit_gradient_matches_central_differences(f_max, relative, masked):
    rng = np.random.default_rng(3)
    pred, obs = rng.standard_normal((2, 3, 5, 120))
    mask = (rng.random((3, 5, 1)) > 0.3).astype(float) if masked else None
    loss, grad = of.band_limited_mse(pred, obs, mask, 0.004, f_max, 0.3, relative)
    for k in range(4):
        e = rng.standard_normal(pred.shape)
        e /= np.linalg.norm(e)
        eps = 1e-4
        lp = of.band_limited_mse(pred + eps * e, obs, mask, 0.004, f_max, 0.3, relative)[0]
        lm = of.band_limited_mse(pred - eps * e, obs, mask, 0.004, f_max, 0.3, relative)[0]
        fd = (lp - lm) / (2 * eps)
        ad = float((grad * e).sum())
        assert abs(fd - ad) <= 1e-7 * max(abs(fd), 1e-3 * abs(loss)), (k, fd, ad)


def test_fft_length_rule_matches_scipy():
    for n in list(range(1, 3000)) + [4001, 6002, 8002, 16384, 20001]:
        assert pf.next_fast_len(n) == scipy.fft.next_fast_len(n, real=True), n


def test_product_taps_equal_the_oracles():
    for n, dt, f_max, taper in ((37, 0.01, 12.0, 0.3), (1000, 0.002, 40.0, 0.25)):
        n_fft = pf.next_fast_len(2 * n)
        w = pf.cosine_response(n_fft, dt, f_max, taper)
        assert np.array_equal(w, of._window(n_fft, dt, f_max, taper))
        assert np.allclose(pf.taps_of(w, n_fft, n), of.taps_direct(w, n_fft, n), atol=1e-14)
        g = pf.gaussian_response(2 * n, dt, 9.0)
        assert np.allclose(pf.taps_of(g, 2 * n, n), of.taps_direct(g, 2 * n, n), atol=1e-14)
    with pytest.raises(ValueError):
        pf.taps_of(np.ones(5), 8, 5)  # 8 < 2*5 - 1: the filter would wrap


def test_filter_plan_rejects_bad_arguments_without_a_gpu():
    lib = nat.lib()
    h = ctypes.c_void_p()
    taps = (ctypes.c_double * 4)(1.0, 0.5, 0.25, 0.0)
    assert lib.nw_filter_create(nat.NW_F32, 0, taps, ctypes.byref(h)) == 3
    assert b"trace length" in lib.nw_last_error()
    assert lib.nw_filter_create(nat.NW_F64, 4097, taps, ctypes.byref(h)) == 3
    assert lib.nw_filter_create(7, 4, taps, ctypes.byref(h)) == 1
    assert lib.nw_filter_create(nat.NW_F32, 4, None, ctypes.byref(h)) == 1
    assert lib.nw_filter_apply(None, None, None, None, None, 1, 1, 1.0, None, None) == 1
    assert lib.nw_filter_length(None) == 0
