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
from nefi.physics.wave import (  # noqa: E402
    WaveInitialConditionOperator,
    WaveOperator,
    laplacian,
    time_reversal,
)


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


def time_reversal_golden():
    """``time_reversal`` images of the reference in both modes, float64 and float32, from the
    traces the reference itself computed for the pat2d and pat3d cases (their f64 traces)."""
    blob, meta = {}, {}
    for make in (cases.pat2d, cases.pat3d):
        case = make()
        name = case["name"]
        z = np.load(HERE / f"{name}.npz", allow_pickle=False)
        traces = z["f64_traces"][0] if z["f64_traces"].ndim == 3 else z["f64_traces"]
        kw = {k: v for k, v in case["kw"].items() if k not in ("n_t",)}
        c = kw.pop("c", 1.0)
        blob[f"{name}_in_traces"] = traces
        meta[name] = {"shape": case["shape"], "extent": case["extent"],
                      "receivers": case["receivers"], "kw": kw,
                      "c_map": isinstance(c, np.ndarray)}
        if isinstance(c, np.ndarray):
            blob[f"{name}_in_c"] = c
        for dtype, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
            cc = torch.as_tensor(c, dtype=dtype) if isinstance(c, np.ndarray) else float(c)
            for mode in ("dirichlet", "adjoint"):
                img = time_reversal(torch.as_tensor(traces, dtype=dtype),
                                    Domain(case["shape"], case["extent"]), case["receivers"],
                                    c=cc, mode=mode, **kw)
                blob[f"{name}_{mode}_{tag}"] = img.numpy()
    blob["meta"] = np.array(json.dumps(meta, default=lambda o: o.tolist()))
    np.savez_compressed(HERE / "time_reversal.npz", **blob)
    print("time_reversal", {k: getattr(v, "shape", None) for k, v in blob.items()}, flush=True)


def laplacian_golden():
    """The reference's ``laplacian`` at orders 2 and 4 on seeded batches of 1-D, 2-D and 3-D
    grids with unequal spacings, float64 and float32."""
    rng = np.random.default_rng(21)
    blob, meta = {}, {}
    for d, shape, spacing in ((1, (3, 17), (0.3,)), (2, (2, 9, 13), (0.5, 0.7)),
                              (3, (2, 6, 7, 5), (0.4, 0.6, 0.9))):
        p = rng.standard_normal(shape)
        blob[f"d{d}_in"] = p
        meta[f"d{d}"] = {"spacing": spacing}
        for order in (2, 4):
            for dtype, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
                out = laplacian(torch.as_tensor(p, dtype=dtype), spacing, order=order)
                blob[f"d{d}_o{order}_{tag}"] = out.numpy()
    blob["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(HERE / "laplacian.npz", **blob)
    print("laplacian", {k: getattr(v, "shape", None) for k, v in blob.items()}, flush=True)


def main(names=None):
    torch.set_num_threads(max(1, min(8, os.cpu_count() or 1)))
    if not names or "time_reversal" in names:
        time_reversal_golden()
    if not names or "laplacian" in names:
        laplacian_golden()
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
