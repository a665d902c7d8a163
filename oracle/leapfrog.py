"""FP64 oracle of nefi's acoustic leapfrog — TEST INFRASTRUCTURE, not part of the product.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` may
import this module. It restates, in numpy and float64, the host-side geometry of
``nefi/physics/wave.py`` and drives the C time loop of ``leapfrog64.c``:

* padded grid, ``n_abs = ceil(W/h - 1e-9)`` (wave.py:204-216), default ``W = 10·max(h)``
  unless ``absorbing == "none"`` (wave.py:385-387);
* CFL step and substeps (wave.py:394-397, timestep.py:177-207);
* layer depth ξ at centres/faces (wave.py:248-261); sponge σ, PML ζ, Cerjan taper and the
  full-grid coefficient arrays a, b, zz, α, β (wave.py:431-504);
* multilinear point stencils with the edge clamp (wave.py:272-300), dense source maps
  (wave.py:312-318), Ricker wavelet (wave.py:101-111, 665-668);
* ``m = pad_replicate(c)^2 dt^2`` (wave.py:718) and the initial-value start
  ``p^{-1} = p0p + m/2 Δ p0p`` (wave.py:849-850);
* gradients: ``dJ/dc = 2 dt^2 c · fold(dJ/dm)`` with the replicate-pad adjoint, and
  ``dJ/dp0 = crop(ḡ_cur + ḡ_prev + Δ(m/2 ḡ_prev))`` for the initial-value operator.

Parity of every function here is pinned by the golden vectors under ``tests/golden`` (written
by the reference itself) and by closed forms and invariants (``tests/test_oracle.py``).
"""

from __future__ import annotations

import ctypes
import math
import subprocess
from ctypes import POINTER, c_double, c_int, c_long, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle64.so"

_CO = {2: (-2.0, 1.0), 4: (-2.5, 4.0 / 3.0, -1.0 / 12.0)}
_RHO = {2: 4.0, 4: 16.0 / 3.0}


class _Prob(ctypes.Structure):
    _fields_ = [
        ("ndim", c_int), ("order", c_int), ("absorbing", c_int),
        ("dims", c_long * 3), ("h", c_double * 3),
        ("n_steps", c_int), ("n_t", c_int), ("substeps", c_int), ("n_shots", c_int),
        ("a", c_void_p), ("b", c_void_p), ("zz", c_void_p), ("taper", c_void_p),
        ("alpha", c_void_p * 3), ("beta", c_void_p * 3),
        ("src", c_void_p), ("wavelet", c_void_p),
        ("rec_taps", c_int), ("n_rec", c_int), ("rec_idx", c_void_p), ("rec_w", c_void_p),
    ]


_lib = None


def build() -> Path:
    subprocess.run(["sh", str(HERE / "build.sh")], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        h = ctypes.CDLL(str(LIB))
        vp = c_void_p
        h.or_forward.argtypes = [POINTER(_Prob), vp, vp, vp, vp, vp]
        h.or_gradient.argtypes = [POINTER(_Prob), vp, vp, vp, vp, vp, vp, vp]
        h.or_gradient_ckpt.argtypes = [POINTER(_Prob), vp, vp, vp, vp, vp, vp, vp, c_int, vp]
        h.or_laplacian.argtypes = [POINTER(_Prob), vp, vp]
        h.or_forward_clamped.argtypes = [POINTER(_Prob), vp, vp, vp, vp, c_int, vp, vp]
        h.or_forward_clamped.restype = c_int
        h.or_forward.restype = h.or_gradient.restype = h.or_gradient_ckpt.restype = c_int
        _lib = h
    return _lib


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def ricker(t, f0, t0=None):
    t0 = 1.2 / f0 if t0 is None else t0
    a2 = (math.pi * f0 * (np.asarray(t, dtype=np.float64) - t0)) ** 2
    return (1.0 - 2.0 * a2) * np.exp(-a2)


class Problem:
    """One leapfrog configuration in float64 (the shapes of ``WaveOperator`` /
    ``WaveInitialConditionOperator``). ``sources=None`` means an initial-value problem."""

    def __init__(self, shape, extent, receivers, *, n_t, dt_obs, c_max, sources=None, f0=None,
                 t0=None, amplitude=1.0, order=4, absorbing="pml", absorb_width=None,
                 absorb_R=1e-3, cerjan_gamma=0.3, courant=0.9):
        self.shape = tuple(int(n) for n in shape)
        self.d = len(self.shape)
        self.extent = tuple((float(lo), float(hi)) for lo, hi in extent)
        self.h = tuple((hi - lo) / n for (lo, hi), n in zip(self.extent, self.shape))
        self.order, self.absorbing = int(order), absorbing
        self.c_max, self.R, self.gamma = float(c_max), float(absorb_R), float(cerjan_gamma)
        if absorb_width is None:
            absorb_width = 10.0 * max(self.h) if absorbing != "none" else 0.0
        W = float(absorb_width) if absorbing != "none" else 0.0
        self.n_abs = tuple(int(math.ceil(W / h - 1e-9)) if W > 0 else 0 for h in self.h)
        self.padded = tuple(n + 2 * a for n, a in zip(self.shape, self.n_abs))
        self.lo_pad = tuple(lo - a * h for (lo, _), a, h in zip(self.extent, self.n_abs, self.h))
        lam = sum(_RHO[self.order] / h**2 for h in self.h)
        dt_max = courant * 2.0 / (self.c_max * math.sqrt(lam))
        self.substeps = max(1, int(math.ceil(dt_obs / dt_max - 1e-12)))
        self.dt = dt_obs / self.substeps
        self.n_t = int(n_t)
        self.n_steps = (self.n_t - 1) * self.substeps
        self.receivers = np.asarray(receivers, dtype=np.float64).reshape(-1, self.d)
        self.rec_idx, self.rec_w = self.point_weights(self.receivers)
        self.sources = None if sources is None else np.asarray(sources, np.float64).reshape(-1, self.d)
        if self.sources is not None:
            t = np.arange(self.n_steps, dtype=np.float64) * self.dt
            self.wavelet = float(amplitude) * ricker(t, float(f0), t0)
        self._coefficients()

    # ---- geometry ------------------------------------------------------------------------
    @property
    def cells(self) -> int:
        return int(np.prod(self.padded))

    def xi(self, ax, faces=False):
        a, n = self.n_abs[ax], self.shape[ax]
        pos = np.arange(n + 2 * a + (1 if faces else 0), dtype=np.float64)
        if not faces:
            pos = pos + 0.5
        if a == 0:
            return np.zeros_like(pos)
        return np.clip(np.maximum(a - pos, pos - (a + n)), 0.0, None) / a

    def _bc(self, v, ax, shape=None):
        view = [1] * self.d
        view[ax] = -1
        return np.broadcast_to(v.reshape(view), shape or self.padded)

    def _rate_max(self, ax, factor):
        return factor * self.c_max * math.log(1.0 / self.R) / (self.n_abs[ax] * self.h[ax])

    def _coefficients(self):
        self.a = self.b = self.zz = self.taper = None
        self.alpha = [None] * self.d
        self.beta = [None] * self.d
        self.pml = self.absorbing == "pml" and any(self.n_abs)
        dt = self.dt
        if self.absorbing == "sponge" and any(self.n_abs):
            sig = np.zeros(self.padded)
            for ax in range(self.d):
                if self.n_abs[ax]:
                    sig = sig + self._rate_max(ax, 3.0) * self._bc(self.xi(ax), ax) ** 2
            half = 0.5 * sig * dt
            self.a, self.b = 1.0 / (1.0 + half), 1.0 - half
        elif self.absorbing == "cerjan" and any(self.n_abs):
            x = np.max(np.stack([self._bc(self.xi(ax), ax) for ax in range(self.d)]), axis=0)
            self.taper = np.exp(-((self.gamma * x) ** 2))
        elif self.pml:
            def zeta(ax, faces=False):
                if self.n_abs[ax] == 0:
                    return self.xi(ax, faces) * 0.0
                return self._rate_max(ax, 1.5) * self.xi(ax, faces) ** 2

            zc = [self._bc(zeta(ax), ax) for ax in range(self.d)]
            zsum = sum(zc)
            self.a = 1.0 / (1.0 + 0.5 * dt * zsum)
            self.b = 1.0 - 0.5 * dt * zsum
            if self.d == 2:
                self.zz = dt**2 * zc[0] * zc[1]
            for ax in range(self.d):
                fshape = list(self.padded)
                fshape[ax] += 1
                zf = self._bc(zeta(ax, True), ax, tuple(fshape))
                other = sum((self._bc(zeta(b), b) for b in range(self.d) if b != ax),
                            np.zeros(self.padded))
                # broadcast the other axes' centre rates onto the faces of this axis
                other_f = np.broadcast_to(
                    np.concatenate([other, other.take([-1], axis=ax)], axis=ax), fshape)
                den = 1.0 + 0.5 * dt * zf
                self.alpha[ax] = np.ascontiguousarray((1.0 - 0.5 * dt * zf) / den)
                self.beta[ax] = np.ascontiguousarray(dt * (other_f - zf) / (den * self.h[ax]))
        for name in ("a", "b", "zz", "taper"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, np.ascontiguousarray(v, dtype=np.float64))

    def point_weights(self, pts):
        pts = np.asarray(pts, dtype=np.float64).reshape(-1, self.d)
        n = pts.shape[0]
        flat = np.zeros((n, 1), dtype=np.int64)
        wts = np.ones((n, 1))
        for ax in range(self.d):
            h, lo, m = self.h[ax], self.lo_pad[ax], self.padded[ax]
            u = (pts[:, ax] - lo) / h - 0.5
            if (u < -1e-9).any() or (u > m - 1 + 1e-9).any():
                raise ValueError(f"point outside the simulation grid along axis {ax}")
            u = np.clip(u, 0.0, m - 1)
            i0 = np.minimum(np.floor(u), max(m - 2, 0)).astype(np.int64)
            t = u - i0
            pi = np.stack([i0, np.minimum(i0 + 1, m - 1)], 1)
            pw = np.stack([1.0 - t, t], 1)
            stride = int(np.prod(self.padded[ax + 1 :]))
            flat = (flat[:, :, None] + stride * pi[:, None, :]).reshape(n, -1)
            wts = (wts[:, :, None] * pw[:, None, :]).reshape(n, -1)
        return np.ascontiguousarray(flat), np.ascontiguousarray(wts)

    def nearest_index(self, pts):
        pts = np.asarray(pts, dtype=np.float64).reshape(-1, self.d)
        flat = np.zeros(pts.shape[0], dtype=np.int64)
        for ax in range(self.d):
            i = np.clip(np.floor((pts[:, ax] - self.lo_pad[ax]) / self.h[ax]), 0,
                        self.padded[ax] - 1).astype(np.int64)
            flat = flat * self.padded[ax] + i
        return flat

    def source_maps(self):
        idx, w = self.point_weights(self.sources)
        out = np.zeros((idx.shape[0], self.cells))
        vol = float(np.prod(self.h))
        for s in range(idx.shape[0]):
            np.add.at(out[s], idx[s], w[s] / vol)
        return out * self.dt**2

    def pad_replicate(self, c):
        idx = [np.clip(np.arange(-a, n + a), 0, n - 1) for a, n in zip(self.n_abs, self.shape)]
        return np.asarray(c, np.float64)[np.ix_(*idx)]

    def pad_zeros(self, x):
        lead = x.ndim - self.d
        return np.pad(x, [(0, 0)] * lead + [(a, a) for a in self.n_abs])

    def crop(self, x):
        lead = x.ndim - self.d
        sl = (slice(None),) * lead + tuple(slice(a, a + n) for a, n in zip(self.n_abs, self.shape))
        return x[sl]

    def fold(self, g):
        """Adjoint of ``pad_replicate``: sum every padded cell onto the field cell it copies."""
        out = np.zeros(self.shape)
        idx = [np.clip(np.arange(-a, n + a), 0, n - 1) for a, n in zip(self.n_abs, self.shape)]
        np.add.at(out, np.ix_(*idx), np.asarray(g).reshape(self.padded))
        return out

    # ---- C interface -------------------------------------------------------------------------
    def _struct(self, n_shots, src=None):
        P = _Prob()
        P.ndim, P.order = self.d, self.order
        P.absorbing = {"none": 0, "sponge": 1, "cerjan": 2, "pml": 3}[self.absorbing]
        if self.absorbing == "pml" and not self.pml:
            P.absorbing = 0
        lead = 3 - self.d
        dims = (1,) * lead + self.padded
        hs = (1.0,) * lead + self.h
        for k in range(3):
            P.dims[k], P.h[k] = dims[k], hs[k]
        P.n_steps, P.n_t, P.substeps, P.n_shots = self.n_steps, self.n_t, self.substeps, n_shots
        P.a, P.b, P.zz, P.taper = _p(self.a), _p(self.b), _p(self.zz), _p(self.taper)
        for ax in range(self.d):
            P.alpha[lead + ax] = _p(self.alpha[ax])
            P.beta[lead + ax] = _p(self.beta[ax])
        if src is not None:
            P.src, P.wavelet = _p(src), _p(self.wavelet)
        P.rec_taps, P.n_rec = self.rec_idx.shape[1], self.rec_idx.shape[0]
        P.rec_idx, P.rec_w = _p(self.rec_idx), _p(self.rec_w)
        return P

    def laplacian(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.empty_like(x)
        P = self._struct(1)
        flat_in = x.reshape(-1, self.cells)
        flat_out = out.reshape(-1, self.cells)
        for k in range(flat_in.shape[0]):
            lib().or_laplacian(ctypes.byref(P), flat_in[k].ctypes.data, flat_out[k].ctypes.data)
        return out

    def model(self, c):
        return np.ascontiguousarray(self.pad_replicate(c) ** 2 * self.dt**2).reshape(-1)

    # ---- FWI -------------------------------------------------------------------------------
    def fwi_traces(self, c, final_state=False):
        S = self.sources.shape[0]
        src = np.ascontiguousarray(self.source_maps())
        m = self.model(c)
        z = np.zeros((S, self.cells))
        tr = np.zeros((S, self.rec_idx.shape[0], self.n_t))
        st = np.zeros((2, S, self.cells)) if final_state else None
        P = self._struct(S, src)
        rc = lib().or_forward(ctypes.byref(P), m.ctypes.data, z.ctypes.data, z.ctypes.data,
                              tr.ctypes.data, _p(st))
        assert rc == 0
        return (tr, st) if final_state else tr

    def fwi_gradient(self, c, residual, every=None, with_traces=False):
        """Gradient of ``sum(residual * traces(c))`` with respect to ``c``.

        ``every``: keep the forward state every ``every`` steps and recompute each block before
        its reverse steps (host memory ~ (every + 1) states instead of n_steps + 1); None keeps
        every state. Both give identical bits. ``with_traces``: return ``(gradient, traces)``,
        the traces observed on the recomputed states (the bits of ``fwi_traces``)."""
        S = self.sources.shape[0]
        src = np.ascontiguousarray(self.source_maps())
        c = np.asarray(c, np.float64)
        m = self.model(c)
        z = np.zeros((S, self.cells))
        res = np.ascontiguousarray(residual, dtype=np.float64)
        gm = np.zeros(self.cells)
        P = self._struct(S, src)
        tr = np.zeros((S, self.rec_idx.shape[0], self.n_t)) if with_traces else None
        rc = lib().or_gradient_ckpt(ctypes.byref(P), m.ctypes.data, z.ctypes.data, z.ctypes.data,
                                    res.ctypes.data, gm.ctypes.data, None, None,
                                    int(every or self.n_steps), _p(tr))
        assert rc == 0
        g = 2.0 * self.dt**2 * c * self.fold(gm)
        return (g, tr) if with_traces else g

    # ---- initial-value problem ----------------------------------------------------------------
    def ic_state(self, c, p0):
        m = self.model(c)
        p0p = self.pad_zeros(np.asarray(p0, np.float64)).reshape(-1, self.cells)
        lap = self.laplacian(p0p)
        return m, p0p + 0.5 * m * lap, p0p

    def ic_traces(self, c, p0):
        m, pm1, p0p = self.ic_state(c, p0)
        B = p0p.shape[0]
        tr = np.zeros((B, self.rec_idx.shape[0], self.n_t))
        P = self._struct(B)
        rc = lib().or_forward(ctypes.byref(P), m.ctypes.data, np.ascontiguousarray(pm1).ctypes.data,
                              np.ascontiguousarray(p0p).ctypes.data, tr.ctypes.data, None)
        assert rc == 0
        return tr

    def ic_gradient(self, c, residual):
        """Gradient of ``sum(residual * traces(p0))`` with respect to ``p0`` (linear operator)."""
        m = self.model(c)
        res = np.ascontiguousarray(residual, dtype=np.float64).reshape(-1, self.rec_idx.shape[0],
                                                                        self.n_t)
        B = res.shape[0]
        z = np.zeros((B, self.cells))
        gm = np.zeros(self.cells)
        gprev = np.zeros((B, self.cells))
        gcur = np.zeros((B, self.cells))
        P = self._struct(B)
        rc = lib().or_gradient(ctypes.byref(P), m.ctypes.data, z.ctypes.data, z.ctypes.data,
                               res.ctypes.data, gm.ctypes.data, gprev.ctypes.data, gcur.ctypes.data)
        assert rc == 0
        g = gcur + gprev + self.laplacian(0.5 * m * gprev)
        return self.crop(g.reshape(B, *self.padded))

    # ---- time reversal (wave.py:878-968) ----------------------------------------------------
    def reversed_traces(self, traces):
        """Traces at reversed time ``T - n dt`` for n = 0..n_steps, linearly interpolated
        between samples (wave.py:941-947): ``(n_steps + 1, n_receivers)``."""
        tr = np.asarray(traces, np.float64)
        n_t, m = tr.shape[-1], self.substeps
        out = np.empty((self.n_steps + 1, tr.shape[0]))
        for n in range(self.n_steps + 1):
            u = (n_t - 1) - n / m
            j0 = min(max(int(math.floor(u)), 0), n_t - 1)
            j1 = min(j0 + 1, n_t - 1)
            t = u - j0
            out[n] = (1.0 - t) * tr[:, j0] + t * tr[:, j1]
        return out

    def time_reversal(self, traces, c, mode="dirichlet"):
        """``time_reversal`` of the initial-value operator on this problem's grid: the receiver
        clamp propagated from rest (``dirichlet``: p^{n} at the receivers' nearest cells set to
        the reversed traces after every step, wave.py:949-967) or the adjoint image
        ``A^T traces`` (``adjoint``, wave.py:943-947)."""
        if mode == "adjoint":
            return self.ic_gradient(c, np.asarray(traces, np.float64)[None])[0]
        m = self.model(c)
        vals = np.ascontiguousarray(self.reversed_traces(traces))
        cells = np.ascontiguousarray(self.nearest_index(self.receivers), dtype=np.int64)
        p0 = np.zeros(self.cells)
        p0[cells] = vals[0]
        z = np.zeros(self.cells)
        st = np.zeros((2, self.cells))
        P = self._struct(1)
        rc = lib().or_forward_clamped(ctypes.byref(P), m.ctypes.data, z.ctypes.data,
                                      p0.ctypes.data, cells.ctypes.data, len(cells),
                                      vals.ctypes.data, st.ctypes.data)
        assert rc == 0
        return self.crop(st[1].reshape(self.padded))

