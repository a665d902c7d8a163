"""Parity of the HIP operators against the reference's golden vectors and the FP64 oracle.

Every case of ``tests/golden`` runs through ``nefi_amd.physics.wave`` on the GPU:

* float64: traces and gradients must equal the reference's float64 results to round-off
  (rel-L2 ≤ 1e-10 / 1e-9): any difference in geometry, timing, stencil or adjoint shows here;
* float32: traces within 1e-5 rel-L2 of the reference's own float32 traces (the stated
  tolerance), gradients within 1e-4 of the float64 reference; the error of each float32 result
  against the float64 values is printed beside it.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from _support import NAMES, field_input, load, objective, operator, rel

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _run(name, dtype, **override):
    meta, z = load(name)
    op = operator(meta, z, **override)
    x = torch.tensor(field_input(meta, z), dtype=dtype, device=DEV)
    want_grad = meta["grad"] is not False
    x.requires_grad_(want_grad)
    y = op({op.primary: x})
    g = None
    if want_grad:
        tag = "f64" if dtype == torch.float64 else "f32"
        d_obs = None
        if meta["grad"] == "misfit":
            d_obs = torch.tensor(z[f"{tag}_d_obs"], dtype=dtype, device=DEV)
        (g,) = torch.autograd.grad(objective(meta, z, y, d_obs), x)
        g = g.cpu().double().numpy()
    return meta, z, y.detach().cpu().double().numpy(), g


@pytest.mark.parametrize("name", NAMES)
def test_float64_matches_reference(name):
    print(f"[case {name} f64]", flush=True)
    meta, z, y, g = _run(name, torch.float64)
    e = rel(y, z["f64_traces"])
    print(f"{name}: f64 traces rel-L2 vs reference {e:.3e}", flush=True)
    assert e < 1e-10
    if g is not None:
        eg = rel(g, z["f64_grad"])
        print(f"{name}: f64 gradient rel-L2 vs reference autograd {eg:.3e}", flush=True)
        assert eg < 1e-9


@pytest.mark.parametrize("name", NAMES)
def test_float32_matches_reference(name):
    print(f"[case {name} f32]", flush=True)
    meta, z, y, g = _run(name, torch.float32)
    e32 = rel(y, z["f32_traces"])
    e64 = rel(y, z["f64_traces"])
    eref = rel(z["f32_traces"], z["f64_traces"])
    print(f"{name}: f32 traces rel-L2 vs reference f32 {e32:.3e}; vs f64 {e64:.3e} "
          f"(reference f32 vs f64 {eref:.3e})", flush=True)
    assert e32 < 1e-5
    if g is not None:
        eg = rel(g, z["f64_grad"])
        egr = rel(z["f32_grad"], z["f64_grad"])
        print(f"{name}: f32 gradient rel-L2 vs f64 reference {eg:.3e} (reference f32 {egr:.3e})",
              flush=True)
        assert eg < 1e-4


@pytest.mark.parametrize("mode,every", [("autograd", None), ("checkpoint", 5),
                                        ("checkpoint", None), ("adjoint", 17)])
def test_grad_modes_agree(mode, every):
    _, z, _, g = _run("pml2d", torch.float64, grad_mode=mode, checkpoint_every=every)
    e = rel(g, z["f64_grad"])
    print(f"pml2d grad_mode={mode} every={every}: rel-L2 {e:.3e}", flush=True)
    assert e < 1e-9


def test_cpu_float64_input_is_staged_to_gpu():
    """nefi's data generators call the operator with CPU float64 fields (bench/base.py:79)."""
    meta, z = load("cerjan2d")
    op = operator(meta, z)
    y = op({"c": torch.tensor(z["in_c"], dtype=torch.float64)})
    assert y.device.type == "cpu" and y.dtype == torch.float64
    assert rel(y.numpy(), z["f64_traces"]) < 1e-10


def test_source_batch_equivalence_and_speed_guard():
    meta, z = load("pml2d")
    c = torch.tensor(z["in_c"], dtype=torch.float64, device=DEV)
    y = operator(meta, z)({"c": c})
    yb = operator(meta, z, source_batch=1)({"c": c})
    assert torch.allclose(y, yb, atol=1e-12, rtol=0)
    with pytest.raises(Exception, match="exceeds"):
        operator(meta, z)({"c": torch.full_like(c, 3.0)})
