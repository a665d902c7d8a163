"""Helpers that turn a golden case into this package's operator and into the FP64 oracle."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = ["cfg1", "grad_sponge", "pml2d", "cerjan2d", "none2d_o2", "sponge2d_o2", "pml1d_o2",
         "pat2d", "pat3d", "cerjan_pat2d"]


def load(name: str):
    z = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    return meta, z


def rel(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def known_c(meta, z):
    """Known sound speed of an initial-value case: (field map, c_max)."""
    c = z["in_kw_c"] if "in_kw_c" in z else meta["kw"]["c"]
    field = np.broadcast_to(np.asarray(c, np.float64), tuple(meta["shape"])).copy()
    return field, float(np.max(c))


def oracle_problem(meta, z):
    from oracle.leapfrog import Problem

    kw = dict(meta["kw"])
    kw.pop("grad_mode", None)
    kw.pop("checkpoint_every", None)
    if meta["kind"] == "fwi":
        return Problem(meta["shape"], meta["extent"], meta["receivers"], sources=meta["sources"],
                       **kw)
    kw.pop("c", None)
    _, cmax = known_c(meta, z)
    return Problem(meta["shape"], meta["extent"], meta["receivers"], c_max=cmax, **kw)


def operator(meta, z, **override):
    import torch

    from nefi_amd.domain import Domain
    from nefi_amd.physics.wave import WaveInitialConditionOperator, WaveOperator

    dom = Domain(meta["shape"], meta["extent"])
    kw = dict(meta["kw"])
    kw.update(override)
    if meta["kind"] == "fwi":
        return WaveOperator(dom, meta["sources"], meta["receivers"], **kw)
    if "in_kw_c" in z:
        kw["c"] = torch.as_tensor(z["in_kw_c"], dtype=torch.float64)
    return WaveInitialConditionOperator(dom, meta["receivers"], **kw)


def field_input(meta, z):
    return z["in_c"] if meta["kind"] == "fwi" else z["in_p0"]


def objective(meta, z, y, d_obs=None):
    """The scalar the golden gradient differentiates, as a function of the traces ``y``."""
    import torch

    if meta["grad"] == "weights":
        w = torch.as_tensor(z["in_weights"], dtype=y.dtype, device=y.device)
        return (y * w).sum()
    return 0.5 * ((y - d_obs) ** 2).sum()
