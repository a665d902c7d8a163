"""Device plans and the autograd Functions of the wave operators.

A :class:`HostPlan` holds what ``_LeapfrogBase._static`` (nefi/physics/wave.py:459-504) and
``WaveOperator._extra_static`` (wave.py:670-674) derive on the host, in float64, as 1-D
profiles and tap lists. A :class:`DevicePlan` uploads it once per (device, dtype) and owns the
``nw_plan`` of ``libnefi_wave.so``. :class:`Propagator` runs forward and reverse sweeps through
the library; the two ``autograd.Function`` classes put the hand-written adjoint behind
``torch.autograd`` in place of autograd's taped per-step graph (precedent for the convention:
``nefi/operators/pde/adjoint.py:280-370``: forward under no_grad, ``once_differentiable``
backward).
"""

from __future__ import annotations

import contextlib
import ctypes
import math
import os
import weakref
import zlib
from dataclasses import dataclass, field

import torch
from torch.autograd.function import once_differentiable

from .. import _native as nat
from ..errors import ConfigError


@dataclass
class HostPlan:
    """Host-side float64 description of one simulation batch (see ``struct nw_desc``)."""

    ndim: int
    order: int
    absorbing: int
    dims: tuple[int, int, int]
    n_abs: tuple[int, int, int]
    field_dims: tuple[int, int, int]
    h: tuple[float, float, float]
    dt: float
    n_steps: int
    n_t: int
    substeps: int
    n_shots: int
    damp: list = field(default_factory=lambda: [None, None, None])
    pml_alpha: list = field(default_factory=lambda: [None, None, None])
    pml_kappa: list = field(default_factory=lambda: [None, None, None])
    pml_zf: list = field(default_factory=lambda: [None, None, None])
    pml_zc: list = field(default_factory=lambda: [None, None, None])
    src_idx: torch.Tensor | None = None  # (S, taps) int64
    src_val: torch.Tensor | None = None  # (S, taps) f64
    wavelet: torch.Tensor | None = None  # (n_steps,) f64
    rec_first: torch.Tensor | None = None
    rec_count: torch.Tensor | None = None
    trace_first: torch.Tensor | None = None
    rec_idx: torch.Tensor | None = None  # (rows, taps) int64
    rec_w: torch.Tensor | None = None  # (rows, taps) f64
    n_trace_rows: int = 0
    pitch: int = 0  # row pitch of wavefield buffers (0 = dims[2])

    @property
    def row_pitch(self) -> int:
        return self.pitch or self.dims[2]

    @property
    def cells(self) -> int:
        """Elements of one shot's (pitched) wavefield."""
        return self.dims[0] * self.dims[1] * self.row_pitch

    def pitched(self, idx: torch.Tensor) -> torch.Tensor:
        """Dense flat cell indices of the padded grid -> indices in pitched buffers."""
        idx = torch.as_tensor(idx, dtype=torch.int64)
        d2 = self.dims[2]
        out = (idx // d2) * self.row_pitch + idx % d2
        return torch.where(idx < 0, idx, out)

    def wavefield_view(self, flat: torch.Tensor, batch: int) -> torch.Tensor:
        """``(batch, *padded)`` view of a pitched flat buffer (padding columns dropped)."""
        lead = self.dims[3 - self.ndim : 2]
        v = flat.view(batch, *lead, self.row_pitch)
        return v[..., : self.dims[2]]

    @property
    def used_axes(self) -> list[int]:
        return list(range(3 - self.ndim, 3))


def injection_tables(rec_idx: torch.Tensor, rec_w: torch.Tensor, groups):
    """Transpose of receiver stencils grouped by distinct cell, per receiver group.

    ``groups`` lists ``(first_row, count)`` of the receiver rows each shot samples; shots with
    the same group share the tables. Returns ``(first, count, cell, off, rec, w)`` as in
    ``nw_desc.inj_*``, entries ordered by (cell, receiver, tap).
    """
    cache: dict = {}
    cells, offs, recs, ws = [], [0], [], []
    first, count = [], []
    for g in groups:
        if g not in cache:
            r0, nr = g
            start = len(cells)
            if nr > 0:
                idx = rec_idx[r0 : r0 + nr].reshape(-1)
                w = rec_w[r0 : r0 + nr].reshape(-1)
                taps = rec_idx.shape[1]
                rr = torch.arange(nr, dtype=torch.int64).repeat_interleave(taps)
                order = torch.argsort(idx * (nr + 1) + rr, stable=True)
                idx, w, rr = idx[order], w[order], rr[order]
                uniq, counts = torch.unique_consecutive(idx, return_counts=True)
                for u, n in zip(uniq.tolist(), counts.tolist()):
                    cells.append(u)
                    offs.append(offs[-1] + n)
                recs.extend(rr.tolist())
                ws.extend(w.tolist())
            cache[g] = (start, len(cells) - start)
        f, n = cache[g]
        first.append(f)
        count.append(n)
    i32 = torch.int32
    return (
        torch.tensor(first, dtype=i32),
        torch.tensor(count, dtype=i32),
        torch.tensor(cells, dtype=i32),
        torch.tensor(offs, dtype=i32),
        torch.tensor(recs, dtype=i32),
        torch.tensor(ws, dtype=torch.float64),
    )


class DevicePlan:
    """One ``nw_plan`` plus the device copies of its tables (kept alive with the plan)."""

    def __init__(self, host: HostPlan, device: torch.device, dtype: torch.dtype) -> None:
        if dtype not in (torch.float32, torch.float64):
            raise ConfigError(f"wave kernels compute in float32 or float64, got {dtype}")
        lib = nat.lib()
        self.host, self.device, self.dtype = host, device, dtype
        self._hold: list[torch.Tensor] = []
        self.tune_choice: dict | None = None  # set by the autotuner (tiles and their source)

        def dev(t, dt=None):
            if t is None:
                return None
            x = t.to(device=device, dtype=dt or dtype).contiguous()
            self._hold.append(x)
            return x.data_ptr()

        d = nat.NwDesc()
        d.ndim, d.order, d.absorbing = host.ndim, host.order, host.absorbing
        d.dtype = nat.NW_F64 if dtype == torch.float64 else nat.NW_F32
        d.pitch = host.row_pitch
        for a in range(3):
            d.dims[a], d.n_abs[a], d.field_dims[a] = host.dims[a], host.n_abs[a], host.field_dims[a]
            d.h[a] = host.h[a]
            d.damp[a] = dev(host.damp[a])
            d.pml_alpha[a] = dev(host.pml_alpha[a])
            d.pml_kappa[a] = dev(host.pml_kappa[a])
            d.pml_zf[a] = dev(host.pml_zf[a])
            d.pml_zc[a] = dev(host.pml_zc[a])
        d.dt = host.dt
        d.n_steps, d.n_t, d.substeps, d.n_shots = host.n_steps, host.n_t, host.substeps, host.n_shots
        if host.src_idx is not None and host.src_idx.numel():
            d.src_taps = host.src_idx.shape[1]
            d.src_idx = dev(host.pitched(host.src_idx), torch.int32)
            d.src_val = dev(host.src_val)
            d.wavelet = dev(host.wavelet)
        d.rec_taps = 1 << host.ndim
        d.n_trace_rows = host.n_trace_rows
        if host.rec_idx is not None and host.rec_idx.numel():
            d.rec_max = int(host.rec_count.max())
            d.rec_first = dev(host.rec_first, torch.int32)
            d.rec_count = dev(host.rec_count, torch.int32)
            d.trace_first = dev(host.trace_first, torch.int32)
            ridx = host.pitched(host.rec_idx)
            d.rec_idx = dev(ridx, torch.int32)
            d.rec_w = dev(host.rec_w)
            groups = list(zip(host.rec_first.tolist(), host.rec_count.tolist()))
            f, n, cell, off, rec, w = injection_tables(ridx, host.rec_w, groups)
            d.inj_max = int(n.max()) if n.numel() else 0
            d.inj_first, d.inj_count = dev(f, torch.int32), dev(n, torch.int32)
            d.inj_cell, d.inj_off = dev(cell, torch.int32), dev(off, torch.int32)
            d.inj_rec, d.inj_w = dev(rec, torch.int32), dev(w)
        self.desc = d
        handle = ctypes.c_void_p()
        nat.check(lib.nw_plan_create(ctypes.byref(d), ctypes.byref(handle)), "nw_plan_create")
        self.handle = handle
        self.esize = 8 if dtype == torch.float64 else 4
        self.ckpt_slot = lib.nw_ckpt_slot_bytes(handle) // self.esize
        S, cells = host.n_shots, host.cells
        self.shot_numel = S * cells
        # PML face buffers as the library lays them out: elements per buffer and the row pitch
        # of each used axis's face grid (axis 2 face rows are pitched)
        self.face_numel, self.face_pitch = [], []
        for a in host.used_axes if host.absorbing == nat.NW_PML else []:
            elems, pitch = ctypes.c_int64(), ctypes.c_int64()
            nat.check(lib.nw_face_layout(handle, a, ctypes.byref(elems), ctypes.byref(pitch)),
                      "nw_face_layout")
            self.face_numel.append(elems.value)
            self.face_pitch.append(pitch.value)

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        h = getattr(self, "handle", None)
        if h is not None and nat._LIB is not None:
            nat._LIB.nw_plan_destroy(h)
            self.handle = None

    @property
    def keep_slot(self) -> int:
        """Elements of one keep slot for the plan's current tiles and options (a slot of kept
        Laplacians carries no PML face fields)."""
        return int(nat.lib().nw_keep_slot_bytes(self.handle)) // self.esize

    # ---- tiles of the fast 2-D kernels -----------------------------------------------------
    def tiles(self, kind: int) -> tuple[int, int]:
        rows, spw = ctypes.c_int32(), ctypes.c_int32()
        nat.check(nat.lib().nw_plan_get_tiles(self.handle, kind, ctypes.byref(rows),
                                              ctypes.byref(spw)), "nw_plan_get_tiles")
        return rows.value, spw.value

    def set_tiles(self, kind: int, rows: int, shots_per_group: int) -> None:
        nat.check(nat.lib().nw_plan_set_tiles(self.handle, kind, rows, shots_per_group),
                  "nw_plan_set_tiles")

    def set_option(self, option: int, value: int) -> None:
        nat.check(nat.lib().nw_plan_set_option(self.handle, option, value), "nw_plan_set_option")

    def option(self, option: int) -> int:
        v = ctypes.c_int32()
        nat.check(nat.lib().nw_plan_get_option(self.handle, option, ctypes.byref(v)),
                  "nw_plan_get_option")
        return v.value

    @contextlib.contextmanager
    def eager(self):
        """Run calls without graph capture (calls whose buffers or step ranges never repeat)."""
        was = self.option(nat.OPT_GRAPH)
        self.set_option(nat.OPT_GRAPH, 0)
        try:
            yield self
        finally:
            self.set_option(nat.OPT_GRAPH, was)

    def grad_slices(self) -> int:
        return int(nat.lib().nw_grad_slices(self.handle))

    # ---- helpers ---------------------------------------------------------------------------
    @property
    def pml(self) -> bool:
        return self.host.absorbing == nat.NW_PML

    def stream(self) -> int:
        return nat.stream_of(self.device)

    def zeros(self, n: int, dtype=None) -> torch.Tensor:
        return torch.zeros(n, device=self.device, dtype=dtype or self.dtype)

    def empty(self, n: int, dtype=None) -> torch.Tensor:
        return torch.empty(n, device=self.device, dtype=dtype or self.dtype)

    def new_state(self) -> tuple[list[torch.Tensor], list[torch.Tensor]]:
        state = [self.zeros(self.shot_numel), self.zeros(self.shot_numel)]
        psi = [self.zeros(n) for n in self.face_numel] + [self.zeros(n) for n in self.face_numel]
        return state, psi

    @staticmethod
    def _reorder(tensors: list[torch.Tensor], ptrs) -> list[torch.Tensor]:
        by_ptr = {t.data_ptr(): t for t in tensors}
        return [by_ptr[p] for p in ptrs]

    # ---- library calls ---------------------------------------------------------------------
    def model(self, c: torch.Tensor, with_ma: bool) -> tuple[torch.Tensor, torch.Tensor | None]:
        cells = self.host.cells
        m = self.empty(cells)
        ma = self.empty(cells) if with_ma else None
        nat.check(
            nat.lib().nw_model(self.handle, c.data_ptr(), m.data_ptr(), nat.ptr(ma), self.stream()),
            "nw_model",
        )
        return m, ma

    def forward(self, m, state, psi, n0, n1, traces=None, ckpt=None, ckpt_every=0, keep=None,
                first_step=0):
        """Advance ``state``/``psi`` (lists, updated in place) over steps ``[n0, n1)``."""
        sp = nat.ptr_array([state[0].data_ptr(), state[1].data_ptr()])
        pp = nat.ptr_array([t.data_ptr() for t in psi]) if self.pml else None
        rc = nat.lib().nw_forward(
            self.handle, m.data_ptr(), sp, pp, n0, n1, nat.ptr(traces), nat.ptr(ckpt),
            ckpt_every, nat.ptr(keep), n0, first_step, self.stream(),
        )
        nat.check(rc, "nw_forward")
        state[:] = self._reorder(state, [sp[0], sp[1]])
        if self.pml:
            k = len(self.face_numel)
            new = []
            for a in range(k):
                pair = [psi[a], psi[k + a]]
                new.append(self._reorder(pair, [pp[a], pp[k + a]]))
            psi[:] = [p[0] for p in new] + [p[1] for p in new]

    def adjoint(self, ma, lam, psib, n0, n1, keep=None, residual=None, inject_end=False,
                grad=None, first_step=0):
        lp = nat.ptr_array([lam[0].data_ptr(), lam[1].data_ptr()])
        pp = nat.ptr_array([t.data_ptr() for t in psib]) if self.pml else None
        rc = nat.lib().nw_adjoint(
            self.handle, ma.data_ptr(), ma.data_ptr(), lp, pp, n0, n1, nat.ptr(keep), n0,
            nat.ptr(residual), int(inject_end), nat.ptr(grad), first_step, self.stream(),
        )
        nat.check(rc, "nw_adjoint")
        lam[:] = self._reorder(lam, [lp[0], lp[1]])
        if self.pml:
            k = len(self.face_numel)
            new = []
            for a in range(k):
                new.append(self._reorder([psib[a], psib[k + a]], [pp[a], pp[k + a]]))
            psib[:] = [p[0] for p in new] + [p[1] for p in new]

    def ckpt_views(self, ckpt: torch.Tensor, slot: int):
        base = slot * self.ckpt_slot
        n = self.shot_numel
        prev = ckpt[base : base + n]
        cur = ckpt[base + n : base + 2 * n]
        faces, off = [], base + 2 * n
        for fn in self.face_numel:
            faces.append(ckpt[off : off + fn])
            off += fn
        return prev, cur, faces


def grad_block(plan: DevicePlan, n_steps: int, mode: str, every: int | None,
               held: int = 0) -> int:
    """Snapshot interval K of the reverse sweep.

    ``autograd`` keeps every step of the forward sweep (no recomputation); ``checkpoint`` and
    ``adjoint`` keep the state every K steps (``checkpoint_every`` or ``ceil(sqrt(n_steps))``)
    and recompute one block at a time. A keep buffer that would not fit in free device memory
    (plus ``held``, the bytes of the caller's keep buffer, which it reuses) falls back to
    ``ceil(sqrt(n_steps))``.
    """
    n = max(1, n_steps)
    k = n if mode == "autograd" else int(every or math.ceil(math.sqrt(n)))
    k = max(1, min(k, n))
    if plan.device.type == "cuda":
        free, _ = torch.cuda.mem_get_info(plan.device)
        need = (k + 1) * plan.keep_slot * plan.esize
        if need > 0.6 * (free + held):
            k = max(1, min(k, math.ceil(math.sqrt(n))))
    return k


class Workspace:
    """Named device buffers reused from call to call.

    Reuse keeps buffer addresses stable, so the hipGraph the library captured for a call is
    replayed by the next identical call instead of being captured again. A workspace is held
    (``busy``) from a forward sweep that saved state for the reverse sweep until that reverse
    sweep ran or the autograd graph was freed; a call finding it busy uses a fresh one.
    """

    def __init__(self, plan: DevicePlan) -> None:
        self.plan = plan
        self.bufs: dict[str, torch.Tensor] = {}
        self.busy = False
        self.generation = 0  # forward sweeps that saved reverse-sweep state here
        self._side: torch.cuda.Stream | None = None

    def side_stream(self) -> torch.cuda.Stream:
        """A second stream (default priority) for work that overlaps the caller's stream."""
        if self._side is None:
            self._side = torch.cuda.Stream(device=self.plan.device)
        return self._side

    # Each buffer starts a name-dependent multiple of 16 KiB (0..60 of them) past its
    # allocation. PyTorch hands out large blocks 2 MiB-aligned, so without the offsets the
    # wavefields one step reads and writes (lambda pair, kept slot, model) sit at equal
    # offsets from 2 MiB boundaries and tend to load the same HBM channels together. Measured on
    # the 64-shot 1064² gradient, alternating on one box, on two boxes: reverse steps 250-265 µs
    # with the offsets (units of 4, 16 or 64 KiB; the physical placement of each run still
    # decides between the two ends) against 260-270 µs without; 1880-1917 against 1909-1937 ms
    # per gradient.
    STAGGER_BYTES = 16384
    STAGGER_SLOTS = 61

    def get(self, name: str, numel: int, dtype=None, zero: bool = False,
            zero_new: bool = False) -> torch.Tensor:
        """Buffer ``name`` of at least ``numel`` elements; ``zero`` clears it on every call,
        ``zero_new`` only when it is (re)allocated (checkpoint and keep slots: PML face fields
        inside the layer-free box are never written and must read as zero)."""
        dt = dtype or self.plan.dtype
        esize = torch.empty(0, dtype=dt).element_size()
        pad = (zlib.crc32(name.encode()) % self.STAGGER_SLOTS) * self.STAGGER_BYTES // esize
        t = self.bufs.get(name)
        if t is None or t.numel() < numel + pad or t.dtype != dt:
            self.bufs.pop(name, None)
            alloc = torch.zeros if zero_new else torch.empty
            t = alloc(max(1, numel + pad), device=self.plan.device, dtype=dt)
            self.bufs[name] = t
        v = t[pad:pad + numel]
        if zero:
            v.zero_()
        return v

    def state(self, prefix: str = "s") -> tuple[list[torch.Tensor], list[torch.Tensor]]:
        P = self.plan
        st = [self.get(f"{prefix}0", P.shot_numel, zero=True),
              self.get(f"{prefix}1", P.shot_numel, zero=True)]
        faces = [self.get(f"{prefix}f{i}", n, zero=True) for i, n in enumerate(P.face_numel * 2)]
        return st, faces

    def hold(self) -> int:
        """Mark the workspace held by a new forward sweep; returns that sweep's generation."""
        self.busy = True
        self.generation += 1
        return self.generation

    def release(self, generation: int | None = None) -> None:
        """End the hold of forward sweep ``generation`` (``None``: whatever holds it). A release
        for an older sweep leaves a newer sweep's hold in place."""
        if generation is None or generation == self.generation:
            self.busy = False


class _Lease:
    """Releases a held workspace when the saved reverse-sweep state is dropped."""

    def __init__(self, ws: Workspace, generation: int) -> None:
        self._fin = weakref.finalize(self, ws.release, generation)


@dataclass
class SavedSweep:
    """What a forward sweep leaves for its reverse sweep: the model, its snapshots every K
    steps and the kept blocks, all inside workspace ``ws`` as forward sweep ``generation``
    wrote them. The first reverse sweep rotates the kept blocks through its own recomputes,
    so it marks the record ``spent``; a later reverse sweep through the same forward (a second
    ``autograd.grad`` with ``retain_graph=True``) runs the forward sweep again from ``c``."""

    m: torch.Tensor
    ma: torch.Tensor
    ckpt: torch.Tensor | None
    keep: torch.Tensor
    K: int
    nb: int
    ws: Workspace
    kept: int
    generation: int
    lease: _Lease | None
    spent: bool = False

    def usable(self) -> bool:
        return not self.spent and self.ws.generation == self.generation


class Propagator:
    """Forward / reverse sweeps of one :class:`DevicePlan`."""

    def __init__(self, plan: DevicePlan, grad_mode: str, checkpoint_every: int | None) -> None:
        self.plan = plan
        self.grad_mode = grad_mode
        self.checkpoint_every = checkpoint_every
        self.ws = Workspace(plan)
        self.block: int | None = None  # snapshot interval of the last gradient forward sweep

    def _workspace(self) -> Workspace:
        return Workspace(self.plan) if self.ws.busy else self.ws

    # ---- FWI: c -> traces ------------------------------------------------------------------
    def fwi_forward(self, c: torch.Tensor, need_grad: bool):
        P = self.plan
        h = P.host
        N = h.n_steps
        hold = need_grad and N > 0
        ws = self._workspace()
        gen = ws.hold() if hold else 0
        m = ws.get("m", h.cells)
        ma = ws.get("ma", h.cells) if need_grad else None
        nat.check(nat.lib().nw_model(P.handle, c.data_ptr(), m.data_ptr(), nat.ptr(ma),
                                     P.stream()), "nw_model")
        state, psi = ws.state()
        traces = ws.get("traces", h.n_trace_rows * h.n_t)
        if not hold:
            P.forward(m, state, psi, 0, N, traces=traces)
            return traces.clone(), None
        # the workspace's keep buffers are reused, so they count as available memory
        held = sum(t.numel() * t.element_size() for k, t in ws.bufs.items()
                   if k.startswith("keep"))
        K = grad_block(P, N, self.grad_mode, self.checkpoint_every, held)
        self.block = K
        nb = -(-N // K)
        last0 = (nb - 1) * K
        ckpt = ws.get("ckpt", (nb - 1) * P.ckpt_slot, zero_new=True) if nb > 1 else None
        # with room for the overlapped reverse sweep's keep buffers, this sweep keeps the last
        # J blocks (one per buffer), so the reverse sweep recomputes J - 1 blocks fewer
        J = self._kept_blocks(ws, K, nb)
        kept0 = (nb - J) * K
        if kept0 > 0:
            P.forward(m, state, psi, 0, kept0, traces=traces, ckpt=ckpt, ckpt_every=K)
        for b in range(nb - J, nb):
            P.forward(m, state, psi, b * K, min(N, (b + 1) * K), traces=traces,
                      keep=self._keep_buf(ws, K, nb - 1 - b))
        keep = self._keep_buf(ws, K, 0)
        saved = SavedSweep(m, ma, ckpt, keep, K, nb, ws, J, gen,
                           _Lease(ws, gen) if ws is self.ws else None)
        return traces.clone(), saved

    def fwi_backward(self, c: torch.Tensor, saved: SavedSweep | None,
                     residual: torch.Tensor) -> torch.Tensor:
        """dJ/dc for the residual dJ/dtraces. ``saved`` is the record of the forward sweep
        (None only for a run of zero steps, whose traces do not depend on ``c``). A record
        already spent by an earlier reverse sweep, or whose workspace a later forward sweep
        reused, is rebuilt by running the forward sweep from ``c`` again (as
        ``torch.utils.checkpoint`` recomputes on every backward), so every backward through one
        forward gives the same gradient."""
        P = self.plan
        h = P.host
        nf = h.field_dims[0] * h.field_dims[1] * h.field_dims[2]
        grad_c = P.zeros(nf)
        if saved is None:
            if h.n_steps > 0:
                raise RuntimeError("wave operator: the forward sweep saved no reverse-sweep "
                                   "state although the input required a gradient")
            return grad_c
        if not saved.usable():
            _, saved = self.fwi_forward(c, True)
        saved.spent = True
        m, ma, ckpt, keep, K, nb, ws, kept = (saved.m, saved.ma, saved.ckpt, saved.keep,
                                              saved.K, saved.nb, saved.ws, saved.kept)
        N = h.n_steps
        res = ws.get("residual", residual.numel())
        res.copy_(residual.reshape(-1))
        lam, psib = ws.state("lam")
        grad_m = ws.get("grad_m", P.grad_slices() * h.cells, torch.float64, zero=True)
        state, psi = ws.state("s")
        if kept > 1 or (nb > 2 and self._overlap_fits(ws, K)):
            self._backward_overlapped(m, ma, ckpt, K, nb, ws, res, lam, psib, grad_m,
                                      state, psi, kept)
            nb = 0  # done
        for b in range(nb - 1, -1, -1):
            n0, n1 = b * K, min(N, (b + 1) * K)
            if b < nb - 1:
                prev, cur, faces = P.ckpt_views(ckpt, b)
                state[0].copy_(prev)
                state[1].copy_(cur)
                for a, f in enumerate(faces):
                    psi[a].copy_(f)
                # keep slots hold p^n, or the Laplacian of p^n, which step n itself computes
                n_end = n1 if P.pml or P.option(nat.OPT_KEEP_LAP) else max(n0, n1 - 1)
                P.forward(m, state, psi, n0, n_end, keep=keep)
            P.adjoint(ma, lam, psib, n0, n1, keep=keep, residual=res,
                      inject_end=(b == nb - 1), grad=grad_m)
        nat.check(
            nat.lib().nw_model_grad(P.handle, c.data_ptr(), grad_m.data_ptr(), grad_c.data_ptr(),
                                    P.stream()),
            "nw_model_grad",
        )
        ws.release(saved.generation)
        return grad_c

    def _overlap_fits(self, ws: Workspace, K: int) -> bool:
        """Overlap needs a second keep buffer in free device memory (``NEFI_AMD_OVERLAP=0``
        turns it off). Temporally blocked reverse launches rotate through spare wavefields of
        their own, so they run beside the blocked forward recompute on the other stream."""
        P = self.plan
        if P.device.type != "cuda" or os.environ.get("NEFI_AMD_OVERLAP", "1") == "0":
            return False
        size = (K + 1) * P.keep_slot
        missing = sum(1 for name in ("keep", "keep2")
                      if not (name in ws.bufs and ws.bufs[name].numel() >= size))
        if missing == 0:
            return True
        free, _ = torch.cuda.mem_get_info(P.device)
        return missing * size * P.esize < 0.8 * free

    @staticmethod
    def _keep_buf(ws: Workspace, K: int, j: int) -> torch.Tensor:
        """Keep buffer j (block nb-1-j of the forward sweep, then every len(bufs)-th block
        before it in the reverse sweep)."""
        P = ws.plan
        return ws.get("keep" if j == 0 else f"keep{j + 1}", (K + 1) * P.keep_slot, zero_new=True)

    def _kept_blocks(self, ws: Workspace, K: int, nb: int) -> int:
        """Blocks the forward sweep keeps for the reverse sweep, one keep buffer each: 1, or
        with the overlapped reverse sweep 2 and more while their buffers fit in 85 % of free
        device memory (at most ``NEFI_AMD_KEEP_BLOCKS``, default 4). Each one more is one
        block fewer to recompute. When fewer than two buffers fit, 1 (the one-stream sweep)."""
        if nb <= 2 or not self._overlap_fits(ws, K):
            return 1
        P = self.plan
        size = (K + 1) * P.keep_slot
        cap = max(2, min(nb, int(os.environ.get("NEFI_AMD_KEEP_BLOCKS", "4"))))
        free, _ = torch.cuda.mem_get_info(P.device)
        budget = 0.85 * free
        j = 0
        while j < cap:
            name = "keep" if j == 0 else f"keep{j + 1}"
            if not (name in ws.bufs and ws.bufs[name].numel() >= size):
                if size * P.esize > budget:
                    break
                budget -= size * P.esize
            j += 1
        return j if j >= 2 else 1

    def _backward_overlapped(self, m, ma, ckpt, K, nb, ws, res, lam, psib, grad_m, state,
                             psi, kept: int) -> None:
        """The reverse sweep with earlier blocks recomputed on a second stream while the
        reverse steps of block b run: max(kept, 2) keep buffers rotate between blocks (block b
        reuses the buffer of block b + len(bufs)); a recompute waits until the reverse steps
        that read its buffer have finished, and the reverse steps of a block wait for its
        recompute. ``kept``: the forward sweep filled buffer j with block nb-1-j for j < kept."""
        P = self.plan
        N = P.host.n_steps
        nbuf = max(kept, 2)
        bufs = [self._keep_buf(ws, K, j) for j in range(nbuf)]
        main = torch.cuda.current_stream(P.device)
        side = ws.side_stream()

        def keep_of(b):
            return bufs[(nb - 1 - b) % nbuf]

        def recompute(b, after: torch.cuda.Event) -> torch.cuda.Event:
            n0, n1 = b * K, min(N, (b + 1) * K)
            with torch.cuda.stream(side):
                side.wait_event(after)
                prev, cur, faces = P.ckpt_views(ckpt, b)
                state[0].copy_(prev)
                state[1].copy_(cur)
                for a, f in enumerate(faces):
                    psi[a].copy_(f)
                n_end = n1 if P.pml or P.option(nat.OPT_KEEP_LAP) else max(n0, n1 - 1)
                P.forward(m, state, psi, n0, n_end, keep=keep_of(b))
                done = torch.cuda.Event()
                done.record(side)
            return done

        ready = {}
        if kept < nbuf:
            start = torch.cuda.Event()
            start.record(main)  # the forward sweep's snapshots and kept blocks are written
            for b in range(nb - 1 - kept, nb - 1 - nbuf, -1):  # blocks whose buffer is free
                ready[b] = recompute(b, start)
        for b in range(nb - 1, -1, -1):
            n0, n1 = b * K, min(N, (b + 1) * K)
            if b in ready:
                main.wait_event(ready.pop(b))
            P.adjoint(ma, lam, psib, n0, n1, keep=keep_of(b), residual=res,
                      inject_end=(b == nb - 1), grad=grad_m)
            if b >= nbuf:  # block b - nbuf reuses block b's keep buffer
                read = torch.cuda.Event()
                read.record(main)
                ready[b - nbuf] = recompute(b - nbuf, read)

    # ---- PAT: p0 -> traces -----------------------------------------------------------------
    def ic_forward(self, m: torch.Tensor, p0: torch.Tensor) -> torch.Tensor:
        P = self.plan
        h = P.host
        ws = self._workspace()
        state, psi = ws.state()
        nat.check(
            nat.lib().nw_ic_state(P.handle, m.data_ptr(), p0.data_ptr(), state[0].data_ptr(),
                                  state[1].data_ptr(), P.stream()),
            "nw_ic_state",
        )
        traces = ws.get("traces", h.n_trace_rows * h.n_t)
        P.forward(m, state, psi, 0, h.n_steps, traces=traces)
        return traces.clone()

    def ic_backward(self, m: torch.Tensor, ma: torch.Tensor, residual: torch.Tensor) -> torch.Tensor:
        P = self.plan
        h = P.host
        ws = self._workspace()
        res = ws.get("residual", residual.numel())
        res.copy_(residual.reshape(-1))
        lam, psib = ws.state("lam")
        P.adjoint(ma, lam, psib, 0, h.n_steps, residual=res, inject_end=True)
        pbar = ws.get("pbar", P.shot_numel)
        nat.check(nat.lib().nw_prev_adjoint(P.handle, lam[1].data_ptr(), pbar.data_ptr(),
                                            P.stream()), "nw_prev_adjoint")
        nf = h.field_dims[0] * h.field_dims[1] * h.field_dims[2]
        g = P.empty(h.n_shots * nf)
        nat.check(nat.lib().nw_ic_grad(P.handle, m.data_ptr(), lam[0].data_ptr(), pbar.data_ptr(),
                                       g.data_ptr(), P.stream()), "nw_ic_grad")
        return g


class FwiFunction(torch.autograd.Function):
    """``c`` (field grid, on the compute device, plan dtype) → flat traces; hand adjoint."""

    @staticmethod
    def forward(ctx, c: torch.Tensor, prop: Propagator):
        need = bool(ctx.needs_input_grad[0])
        with torch.no_grad():
            traces, saved = prop.fwi_forward(c.detach().contiguous(), need)
        ctx.prop = prop
        ctx.saved = saved
        ctx.save_for_backward(c)
        return traces

    @staticmethod
    @once_differentiable
    def backward(ctx, g: torch.Tensor):
        (c,) = ctx.saved_tensors
        prop = ctx.prop
        res = g.to(device=prop.plan.device, dtype=prop.plan.dtype).contiguous()
        saved = ctx.saved
        grad = prop.fwi_backward(c.detach().contiguous(), saved, res)
        if saved is not None:
            # a later backward through this forward reruns the forward sweep from c; keep only
            # the spent marker, not the workspace's buffers
            ctx.saved = SavedSweep(*(None,) * 6, saved.ws, 0, saved.generation, None, True)
        return grad.view_as(c), None


class IcFunction(torch.autograd.Function):
    """``p0`` (batch, field grid) → flat traces of the initial-value operator; hand adjoint."""

    @staticmethod
    def forward(ctx, p0: torch.Tensor, prop: Propagator, m: torch.Tensor, ma: torch.Tensor):
        with torch.no_grad():
            traces = prop.ic_forward(m, p0.detach().contiguous())
        ctx.prop, ctx.m, ctx.ma = prop, m, ma
        ctx.p0_shape = p0.shape
        return traces

    @staticmethod
    @once_differentiable
    def backward(ctx, g: torch.Tensor):
        prop = ctx.prop
        res = g.to(device=prop.plan.device, dtype=prop.plan.dtype).contiguous()
        grad = prop.ic_backward(ctx.m, ctx.ma, res)
        return grad.view(ctx.p0_shape), None, None, None
