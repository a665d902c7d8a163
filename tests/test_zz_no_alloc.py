"""The allocation-free trace filter entry point against the allocating one, through the C ABI.

``nw_filter_apply_noalloc`` runs the kernels of ``nw_filter_apply`` with the workgroups' float64
partial sums in the library's fixed per-GPU array instead of the plan's buffer: its filtered
traces and its sums must carry the same bits. Two calls in a row, the second on fewer rows
(the array keeps the first call's partial sums, which the second must not read), for the FFT
filter and the unfiltered masked sums, in float32 and float64.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from nefi_amd import _native as nat

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _plan(dtype: torch.dtype, n: int):
    t = np.arange(n, dtype=np.float64)
    taps = np.exp(-0.5 * (t / 7.0) ** 2) * np.cos(0.3 * t)  # a symmetric kernel's lags 0..n-1
    handle = ctypes.c_void_p()
    code = nat.NW_F64 if dtype == torch.float64 else nat.NW_F32
    arr = (ctypes.c_double * n)(*taps)
    nat.check(nat.lib().nw_filter_create(code, n, arr, ctypes.byref(handle)), "nw_filter_create")
    return handle


def _call(fn, handle, x, xs, mask, rows, filt, scale):
    y = torch.empty_like(x)
    sums = torch.full((2,), float("nan"), device=DEV, dtype=torch.float64)
    rc = fn(handle, x.data_ptr(), xs.data_ptr(), mask.data_ptr(), y.data_ptr(), rows, filt,
            scale, sums.data_ptr(), nat.stream_of(torch.device(DEV)))
    nat.check(rc, "filter apply")
    torch.cuda.synchronize()
    return y, sums


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("filt", [1, 0])
def test_noalloc_filter_apply_matches_allocating_apply(dtype, filt):
    lib = nat.lib()
    n = 1500
    handle = _plan(dtype, n)
    g = torch.Generator(DEV).manual_seed(17)
    try:
        for rows in (4097, 37):  # the second call smaller than the first
            print(f"[case noalloc filter {dtype} filter={filt} rows={rows}]", flush=True)
            x = torch.randn(rows, n, device=DEV, dtype=dtype, generator=g)
            xs = torch.randn(rows, n, device=DEV, dtype=dtype, generator=g)
            mask = (torch.rand(rows, n, device=DEV, generator=g) > 0.3).to(dtype)
            y_ref, s_ref = _call(lib.nw_filter_apply, handle, x, xs, mask, rows, filt, 0.75)
            y, s = _call(lib.nw_filter_apply_noalloc, handle, x, xs, mask, rows, filt, 0.75)
            print(f"sums {s.tolist()} vs {s_ref.tolist()}", flush=True)
            assert torch.isfinite(s).all() and float(s[1]) > 0
            assert torch.equal(y, y_ref) and torch.equal(s, s_ref)
    finally:
        lib.nw_filter_destroy(handle)


def test_noalloc_filter_apply_refuses_bad_arguments():
    lib = nat.lib()
    handle = _plan(torch.float32, 64)
    try:
        assert lib.nw_filter_apply_noalloc(handle, None, None, None, None, 3, 1, 1.0, None,
                                           None) == 1  # NW_ERR_ARG
        assert b"null argument" in lib.nw_last_error()
    finally:
        lib.nw_filter_destroy(handle)
