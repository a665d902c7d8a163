"""Write the golden vectors of ``tests/golden/*.npz`` by running the reference itself.

Run in a container where the reference is mounted (``python tests/golden/make_golden.py``);
it imports ``nefi`` from ``/root/reference`` (or ``$NEFI_REFERENCE``), evaluates every case of
``cases.py`` with the reference ``WaveOperator`` / ``WaveInitialConditionOperator`` in float64
and float32 on the CPU, and stores inputs and outputs (traces and autograd gradients). Nothing
of the reference is stored but these numbers. The exact values pinned this way — wavelet
timing, the n = 0 sample, edge clamps of the stencils, replicate padding, PML face profiles —
are the ones closed forms and invariants leave open.
"""

from __future__ import annotations

import json
import os
import sys
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, os.environ.get("NEFI_REFERENCE", "/root/reference"))

import cases  # noqa: E402
from nefi.domain import Domain  # noqa: E402
from nefi.physics.wave import WaveInitialConditionOperator, WaveOperator  # noqa: E402


def _op(case, dtype):
    dom = Domain(case["shape"], case["extent"])
    kw = dict(case["kw"])
    if case["kind"] == "fwi":
        return WaveOperator(dom, case["sources"], case["receivers"], **kw)
    if isinstance(kw.get("c"), np.ndarray):
        kw["c"] = torch.as_tensor(kw["c"], dtype=torch.float64)
    return WaveInitialConditionOperator(dom, case["receivers"], **kw)


def run(case, dtype):
    op = _op(case, dtype)
    out = {}
    key = "c" if case["kind"] == "fwi" else "p0"
    x = torch.as_tensor(case[key], dtype=dtype).requires_grad_(case["grad"] is not False)
    y = op({op.primary: x})
    out["traces"] = y.detach().numpy()
    if case["grad"] == "weights":
        w = torch.as_tensor(case["weights"], dtype=dtype)
        (g,) = torch.autograd.grad((y * w).sum(), x)
        out["grad"] = g.numpy()
    elif case["grad"] == "misfit":
        with torch.no_grad():
            d_obs = op({op.primary: torch.as_tensor(case["c_true"], dtype=dtype)})
        loss = 0.5 * ((y - d_obs) ** 2).sum()
        (g,) = torch.autograd.grad(loss, x)
        out["grad"] = g.numpy()
        out["loss"] = np.array(float(loss))
        out["d_obs"] = d_obs.numpy()
    out["substeps"] = np.array(op.substeps)
    out["n_steps"] = np.array(op.n_steps)
    out["dt"] = np.array(op.dt)
    return out


def main(names=None):
    torch.set_num_threads(max(1, min(8, os.cpu_count() or 1)))
    for make in cases.ALL:
        case = make()
        if names and case["name"] not in names:
            continue
        blob = {}
        for dtype, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
            for k, v in run(case, dtype).items():
                blob[f"{tag}_{k}"] = v
        meta = {k: v for k, v in case.items() if not isinstance(v, np.ndarray)}
        blob["meta"] = np.array(json.dumps(meta, default=lambda o: o.tolist()))
        for k in ("c", "c_true", "p0", "weights"):
            if isinstance(case.get(k), np.ndarray):
                blob[f"in_{k}"] = case[k]
        if isinstance(case["kw"].get("c"), np.ndarray):
            blob["in_kw_c"] = case["kw"]["c"]
        np.savez_compressed(HERE / f"{case['name']}.npz", **blob)
        print(case["name"], {k: getattr(v, "shape", None) for k, v in blob.items()}, flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
