"""Many shots of uneven source and receiver counts advanced together.

nefi's ``WaveOperator`` shares one receiver set across all sources (wave.py:589); a survey whose
shots see different receivers is expressed there as one operator per shot, or as the union of
receivers plus a mask. :class:`ShotBatchWaveOperator` keeps every shot's own geometry and runs
the whole batch through one plan: each time step is one launch over all shots, receivers are
sampled and residuals injected through per-shot (CSR) tables, and the gradient is the same
hand-written reverse sweep as ``WaveOperator``'s. Traces come back as the concatenation of the
shots' receiver rows, ``(sum_s R_s, n_t)``; :meth:`split` cuts them per shot.
"""

from __future__ import annotations

from collections.abc import Sequence

import torch

from ..domain import Domain
from ..errors import ShapeError
from ..operators.base import Fields
from ..registry import register
from ..utils import shape_tuple
from ._engine import FwiFunction, HostPlan
from .wave import WaveOperator, _as_points, _merge_taps

__all__ = ["ShotBatchWaveOperator"]


@register("operator", "wave_shot_batch")
class ShotBatchWaveOperator(WaveOperator):
    """FWI forward model over shots ``[(sources_s, receivers_s), ...]``.

    Every shot injects the same wavelet at each of its ``sources_s`` (``(k_s, d)`` points,
    ``k_s ≥ 1``) and records its own ``receivers_s`` (``(R_s, d)``). Other arguments as
    :class:`~nefi_amd.physics.wave.WaveOperator`; ``source_batch`` is not used (the batch is one
    plan).
    """

    def __init__(self, domain: Domain, shots: Sequence, **kwargs) -> None:
        kwargs.pop("source_batch", None)
        if not shots:
            raise ShapeError("ShotBatchWaveOperator needs at least one shot")
        d = domain.ndim
        self.shot_sources = [_as_points(s, d, "sources") for s, _ in shots]
        self.shot_receivers = [_as_points(r, d, "receivers") for _, r in shots]
        if any(s.shape[0] < 1 for s in self.shot_sources):
            raise ShapeError("every shot needs at least one source")
        all_rec = torch.cat(self.shot_receivers, 0)
        first_src = torch.stack([s[0] for s in self.shot_sources])
        super().__init__(domain, first_src, all_rec, **kwargs)
        self._shots_arg = [(s, r) for s, r in zip(self.shot_sources, self.shot_receivers)]
        self.receiver_counts = [int(r.shape[0]) for r in self.shot_receivers]

    def output_shape(self, shape):
        return (sum(self.receiver_counts), self.n_t)

    def split(self, traces: torch.Tensor) -> list[torch.Tensor]:
        """Per-shot views ``(R_s, n_t)`` of concatenated traces."""
        return list(torch.split(traces, self.receiver_counts, dim=0))

    def _shot_plan(self, s0: int, s1: int) -> HostPlan:
        hp = self._host_plan(s1 - s0)
        g = self.grid
        idxs, vals = [], []
        kmax = max(int(s.shape[0]) for s in self.shot_sources[s0:s1])
        taps = 1 << g.d
        for s in range(s0, s1):
            i, w = g.point_weights(self.shot_sources[s])
            i = i.reshape(1, -1)
            w = (w / g.cell_volume).reshape(1, -1)
            mi, mv = _merge_taps(i, w)
            pad = kmax * taps - mi.shape[1]
            if pad:
                mi = torch.cat([mi, torch.full((1, pad), -1, dtype=mi.dtype)], 1)
                mv = torch.cat([mv, torch.zeros(1, pad, dtype=mv.dtype)], 1)
            idxs.append(mi)
            vals.append(mv)
        hp.src_idx = torch.cat(idxs, 0)
        hp.src_val = torch.cat(vals, 0) * self.dt**2
        hp.wavelet = self.wavelet()
        ridx, rw = self._rec
        counts = torch.tensor(self.receiver_counts, dtype=torch.int64)
        starts = torch.cumsum(counts, 0) - counts
        hp.rec_idx, hp.rec_w = ridx, rw
        hp.rec_first = starts[s0:s1].clone()
        hp.rec_count = counts[s0:s1].clone()
        hp.trace_first = starts[s0:s1] - starts[s0]
        hp.n_trace_rows = int(counts[s0:s1].sum())
        return hp

    def at_resolution(self, shape):
        shape = shape_tuple(shape)
        if shape == tuple(self.domain.shape):
            return self
        if shape not in self._res:
            op = ShotBatchWaveOperator(
                self.domain.at(shape), self._shots_arg, f0=self.f0, t0=self.t0,
                amplitude=self.amplitude, field=self.primary, **self._kwargs(),
            )
            op.fidelity_tag = self.fidelity_tag
            self._res[shape] = op
        return self._res[shape]

    def simulate(self, c, observe=None, record=None, sources=None):
        if observe is not None or sources is not None:
            return super().simulate(c, observe, record, sources)
        y = self.forward({self.primary: c})
        return (), list(y.unbind(-1))

    def forward(self, fields: Fields) -> torch.Tensor:
        c = self.get_field(fields)
        prop, s0, s1, dev, work = self._prepare(c, None)
        x = c.to(device=dev, dtype=work)
        flat = FwiFunction.apply(x, prop)
        return flat.view(-1, self.n_t).to(device=c.device, dtype=c.dtype)
