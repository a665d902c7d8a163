"""The operator contract the solver calls (``nefi/operators/base.py:27-77``).

An operator is an ``nn.Module`` mapping a dict of named fields to a prediction tensor. The
attributes below are read by nefi's solver and diagnostics, so their names and defaults match the
reference exactly.
"""

from __future__ import annotations

from collections.abc import Mapping

import torch
from torch import nn

from ..errors import OperatorError

Fields = Mapping[str, torch.Tensor]


class Operator(nn.Module):
    """Differentiable forward model ``fields -> prediction``.

    Attributes:
        primary: name of the field the operator is (or is not) homogeneous in.
        homogeneity: ``p`` with ``F(s·x) = s^p F(x)``, or ``None``.
        batchable: accepts a leading batch axis with plain tensor ops only.
        traceable: may be traced by a whole-step compiler; time-loop operators are not.
    """

    primary: str = "x"
    homogeneity: float | None = None
    batchable: bool = False
    traceable: bool = True

    def forward(self, fields: Fields) -> torch.Tensor:  # pragma: no cover - abstract
        raise NotImplementedError

    def at_resolution(self, shape) -> Operator:
        return self

    def output_shape(self, shape) -> tuple[int, ...] | None:
        return None

    def required_fields(self) -> tuple[str, ...]:
        return (self.primary,)

    def get_field(self, fields: Fields, name: str | None = None) -> torch.Tensor:
        key = name or self.primary
        if key not in fields:
            raise OperatorError(
                f"{type(self).__name__} needs field {key!r}; available: {tuple(fields)}"
            )
        return fields[key]

    def apply(self, x: torch.Tensor, **others: torch.Tensor) -> torch.Tensor:
        return self({self.primary: x, **others})
