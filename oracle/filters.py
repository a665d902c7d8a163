"""FP64 oracle of nefi's trace filters and band-limited misfit (test infrastructure).

Only the tests, ``__graft_entry__.smoke`` and ``bench.py``'s CPU baseline use this module, and
only as the checker; the product path is ``nefi_amd.physics.filters`` over the HIP library.

Restated (numpy float64, FFTs from ``numpy.fft``):

* ``lowpass`` — ``nefi/instances/wave_fwi/losses.py:29-50``: cosine-tapered window on
  ``rfftfreq(N, dt)``, ``N = next_fast_len(2n)`` (``nefi/utils/tensor.py:21-23``, scipy's real
  rule), ``irfft(rfft(x, N) * w, N)[..., :n]``.
* ``gaussian_lowpass`` — ``nefi/instances/photoacoustic3d/sensor.py:24-38``: ``N = 2n``,
  ``H = exp(-(f/f_c)^2 / 2)``; ``f_c`` None or <= 0 returns the input.
* ``band_limited_mse`` — ``losses.py:65-80`` with ``masked_mean`` (``nefi/measurement.py:86-91``):
  the loss and its gradient in ``pred`` (the filter is self-adjoint for N >= 2n - 1, so the
  gradient is ``F(2 m F(r)) / (sum m * den)``).

Pinned by (tests/test_filters.py): the reference's own filter tests
(``tests/test_instances_wave_family.py:49-54``, ``tests/test_photoacoustic3d.py:40-46``); a
brute-force O(n^2) direct convolution with taps from a textbook cosine-sum inverse DFT
(``toeplitz_filter`` + ``taps_direct``: catches a wrong N, frequency grid or crop); the window's
closed form at the roll-off midpoint (amplitude 1/2 of a long sinusoid); self-adjointness; the
misfit's gradient against central finite differences.
"""

from __future__ import annotations

import math

import numpy as np
import scipy.fft


def _window(n_fft: int, dt: float, f_max: float, taper: float) -> np.ndarray:
    """losses.py:29-37."""
    f = np.fft.rfftfreq(n_fft, d=dt)
    f1 = (1.0 - taper) * f_max
    w = np.ones_like(f)
    roll = (f > f1) & (f < f_max)
    w[roll] = 0.5 * (1.0 + np.cos(math.pi * (f[roll] - f1) / max(f_max - f1, 1e-30)))
    w[f >= f_max] = 0.0
    return w


def lowpass(x: np.ndarray, dt: float, f_max: float, taper: float = 0.3) -> np.ndarray:
    """losses.py:40-50 in float64."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[-1]
    n_fft = scipy.fft.next_fast_len(2 * n, real=True)
    w = _window(n_fft, dt, f_max, taper)
    return np.fft.irfft(np.fft.rfft(x, n=n_fft, axis=-1) * w, n=n_fft, axis=-1)[..., :n]


def gaussian_lowpass(x: np.ndarray, dt: float, f_cutoff: float | None) -> np.ndarray:
    """sensor.py:24-38 in float64."""
    if f_cutoff is None or f_cutoff <= 0:
        return x
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[-1]
    size = 2 * n
    f = np.fft.rfftfreq(size, d=float(dt))
    h = np.exp(-0.5 * (f / float(f_cutoff)) ** 2)
    return np.fft.irfft(np.fft.rfft(x, n=size, axis=-1) * h, n=size, axis=-1)[..., :n]


def _masked_mean(x: np.ndarray, mask: np.ndarray | None) -> float:
    """measurement.py:86-91."""
    if mask is None:
        return float(x.mean())
    m = np.broadcast_to(mask, x.shape)
    return float((x * m).sum() / max(float(m.sum()), 1e-12))


def band_limited_mse(pred, obs, mask, dt: float, f_max: float | None, taper: float = 0.3,
                     relative: bool = True) -> tuple[float, np.ndarray]:
    """BandLimitedMSE.forward (losses.py:65-80) and its gradient in ``pred`` (float64)."""
    pred = np.asarray(pred, dtype=np.float64)
    obs = np.broadcast_to(np.asarray(obs, dtype=np.float64), pred.shape)
    m = None if mask is None else np.broadcast_to(np.asarray(mask, np.float64), pred.shape)
    filt = (lambda v: v) if f_max is None else (lambda v: lowpass(v, dt, f_max, taper))
    r = filt(pred - obs)
    num = _masked_mean(r**2, m)
    den = max(_masked_mean(filt(obs) ** 2, m), 1e-30) if relative else 1.0
    msum = float(pred.size) if m is None else max(float(m.sum()), 1e-12)
    z = r if m is None else m * r
    grad = filt(2.0 * z) / (msum * den)
    return num / den, grad


# ---- independent checks of the FFT restatement ------------------------------------------------
def taps_direct(response: np.ndarray, n_fft: int, n: int) -> np.ndarray:
    """h[k] = irfft(response, n_fft)[k] for k < n by the textbook cosine sum of a real, even
    spectrum (no FFT): (H_0 + 2 sum_{j=1}^{N/2-1} H_j cos(2 pi j k / N) + H_{N/2} (-1)^k) / N
    for even N (the last term absent for odd N)."""
    H = np.asarray(response, dtype=np.float64)
    k = np.arange(n)[:, None]
    half = n_fft // 2
    j = np.arange(1, half if n_fft % 2 == 0 else half + 1)[None, :]
    s = H[0] + 2.0 * (H[1 : j.shape[1] + 1][None, :] * np.cos(2 * np.pi * j * k / n_fft)).sum(1)
    if n_fft % 2 == 0:
        s = s + H[half] * np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    return s / n_fft


def toeplitz_filter(x: np.ndarray, h: np.ndarray) -> np.ndarray:
    """y[t] = sum_s h[|t - s|] x[s] along the last axis, by direct summation."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[-1]
    idx = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :])
    return x @ h[idx].T
