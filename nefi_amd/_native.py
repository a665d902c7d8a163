"""ctypes binding of ``libnefi_wave.so`` (declared in ``include/nefi_wave.h``,
``include/nefi_filter.h`` and ``include/nefi_stepper.h``).

The library is loaded after ``torch`` so that it binds to the HIP runtime torch already loaded
(both carry the soname ``libamdhip64.so.7``); torch's streams and device pointers are then valid
in it. Nothing here computes: every wave-path computation is a call into the library, and a
missing library or GPU raises :class:`~nefi_amd.errors.NativeError` instead of falling back.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_double, c_int32, c_int64, c_void_p
from pathlib import Path

import torch

from .errors import NativeError

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libnefi_wave.so"

NW_NONE, NW_SPONGE, NW_CERJAN, NW_PML = 0, 1, 2, 3
NW_F32, NW_F64 = 0, 1

ABI_VERSION = 4

# every symbol include/nefi_wave.h declares
EXPORTS = (
    "nw_abi_version",
    "nw_desc_bytes",
    "nw_last_error",
    "nw_plan_create",
    "nw_plan_destroy",
    "nw_model",
    "nw_forward",
    "nw_ckpt_slot_bytes",
    "nw_keep_slot_bytes",
    "nw_face_layout",
    "nw_adjoint",
    "nw_model_grad",
    "nw_ic_state",
    "nw_ic_grad",
    "nw_prev_adjoint",
    "nw_laplacian",
    "nw_clamp",
    "nw_profile_enable",
    "nw_profile_read",
    "nw_profile_units",
    "nw_plan_set_tiles",
    "nw_plan_get_tiles",
    "nw_grad_slices",
    "nw_plan_set_option",
    "nw_plan_get_option",
)
# every symbol include/nefi_filter.h declares
FILTER_EXPORTS = (
    "nw_filter_create",
    "nw_filter_destroy",
    "nw_filter_length",
    "nw_filter_apply",
    "nw_filter_apply_noalloc",
)
# every symbol include/nefi_stepper.h declares
STEPPER_EXPORTS = (
    "nw_ts_laplacian",
    "nw_ts_wave_step",
    "nw_ts_mur",
    "nw_ts_mur_adjoint",
    "nw_ts_wave_step_adjoint",
)
NW_TS_PERIODIC, NW_TS_NEUMANN, NW_TS_DIRICHLET, NW_TS_ABSORBING = 0, 1, 2, 3


class NwTsGrid(ctypes.Structure):
    """Mirror of ``struct nw_ts_grid`` (include/nefi_stepper.h)."""

    _fields_ = [
        ("ndim", c_int32),
        ("dtype", c_int32),
        ("boundary", c_int32),
        ("reserved", c_int32),
        ("batch", c_int64),
        ("shape", c_int64 * 3),
        ("h", c_double * 3),
    ]


class NwDesc(ctypes.Structure):
    """Mirror of ``struct nw_desc``; its size is checked against ``nw_desc_bytes()``."""

    _fields_ = [
        ("ndim", c_int32),
        ("dtype", c_int32),
        ("order", c_int32),
        ("absorbing", c_int32),
        ("dims", c_int64 * 3),
        ("n_abs", c_int64 * 3),
        ("field_dims", c_int64 * 3),
        ("pitch", c_int64),
        ("h", c_double * 3),
        ("dt", c_double),
        ("n_steps", c_int32),
        ("n_t", c_int32),
        ("substeps", c_int32),
        ("n_shots", c_int32),
        ("damp", c_void_p * 3),
        ("pml_alpha", c_void_p * 3),
        ("pml_kappa", c_void_p * 3),
        ("pml_zf", c_void_p * 3),
        ("pml_zc", c_void_p * 3),
        ("src_taps", c_int32),
        ("src_idx", c_void_p),
        ("src_val", c_void_p),
        ("wavelet", c_void_p),
        ("rec_taps", c_int32),
        ("rec_max", c_int32),
        ("n_trace_rows", c_int32),
        ("rec_first", c_void_p),
        ("rec_count", c_void_p),
        ("trace_first", c_void_p),
        ("rec_idx", c_void_p),
        ("rec_w", c_void_p),
        ("inj_max", c_int32),
        ("inj_first", c_void_p),
        ("inj_count", c_void_p),
        ("inj_cell", c_void_p),
        ("inj_off", c_void_p),
        ("inj_rec", c_void_p),
        ("inj_w", c_void_p),
    ]


_LIB: ctypes.CDLL | None = None


def _declare(lib: ctypes.CDLL) -> None:
    vp, i32, i64 = c_void_p, c_int32, c_int64
    vpp = POINTER(c_void_p)
    sig = {
        "nw_abi_version": (i32, []),
        "nw_desc_bytes": (i64, []),
        "nw_last_error": (ctypes.c_char_p, []),
        "nw_plan_create": (i32, [POINTER(NwDesc), POINTER(c_void_p)]),
        "nw_plan_destroy": (i32, [vp]),
        "nw_model": (i32, [vp, vp, vp, vp, vp]),
        "nw_forward": (i32, [vp, vp, vpp, vpp, i32, i32, vp, vp, i32, vp, i32, i32, vp]),
        "nw_ckpt_slot_bytes": (i64, [vp]),
        "nw_keep_slot_bytes": (i64, [vp]),
        "nw_face_layout": (i32, [vp, i32, POINTER(c_int64), POINTER(c_int64)]),
        "nw_adjoint": (i32, [vp, vp, vp, vpp, vpp, i32, i32, vp, i32, vp, i32, vp, i32, vp]),
        "nw_model_grad": (i32, [vp, vp, vp, vp, vp]),
        "nw_ic_state": (i32, [vp, vp, vp, vp, vp, vp]),
        "nw_ic_grad": (i32, [vp, vp, vp, vp, vp, vp]),
        "nw_prev_adjoint": (i32, [vp, vp, vp, vp]),
        "nw_laplacian": (i32, [vp, vp, vp, i64, vp]),
        "nw_clamp": (i32, [vp, vp, vp, vp, i32, vp]),
        "nw_profile_enable": (i32, [i32]),
        "nw_profile_read": (i32, [i32, POINTER(c_double), POINTER(c_int64)]),
        "nw_profile_units": (i32, [i32, POINTER(c_int64)]),
        "nw_plan_set_tiles": (i32, [vp, i32, i32, i32]),
        "nw_plan_get_tiles": (i32, [vp, i32, POINTER(c_int32), POINTER(c_int32)]),
        "nw_grad_slices": (i64, [vp]),
        "nw_plan_set_option": (i32, [vp, i32, i32]),
        "nw_plan_get_option": (i32, [vp, i32, POINTER(c_int32)]),
        "nw_filter_create": (i32, [i32, i64, POINTER(c_double), POINTER(c_void_p)]),
        "nw_filter_destroy": (i32, [vp]),
        "nw_filter_length": (i64, [vp]),
        "nw_filter_apply": (i32, [vp, vp, vp, vp, vp, i64, i32, c_double, vp, vp]),
        "nw_filter_apply_noalloc": (i32, [vp, vp, vp, vp, vp, i64, i32, c_double, vp, vp]),
        "nw_ts_laplacian": (i32, [POINTER(NwTsGrid), vp, vp, vp]),
        "nw_ts_wave_step": (i32, [POINTER(NwTsGrid), c_double, vp, vp, vp, vp, vp, vp, i32, vp]),
        "nw_ts_mur": (i32, [POINTER(NwTsGrid), i32, c_double, vp, vp, vp, vp]),
        "nw_ts_mur_adjoint": (i32, [POINTER(NwTsGrid), i32, c_double, vp, vp, vp, vp, vp, vp, vp]),
        "nw_ts_wave_step_adjoint": (i32, [POINTER(NwTsGrid), c_double, vp, vp, vp, vp, vp, vp, vp,
                                          vp, vp, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib() -> ctypes.CDLL:
    """The loaded library; raises :class:`NativeError` when it is not built or mismatched."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = Path(os.environ.get("NEFI_AMD_LIB", LIB_PATH))
    if not path.exists():
        raise NativeError(
            f"{path} is missing: build the HIP library first (python -c "
            "'import __graft_entry__ as g; g.build()')"
        )
    try:
        handle = ctypes.CDLL(str(path))
    except OSError as e:  # pragma: no cover - depends on the system
        raise NativeError(f"cannot load {path}: {e}") from e
    _declare(handle)
    if handle.nw_abi_version() != ABI_VERSION:
        raise NativeError(f"{path} has ABI {handle.nw_abi_version()}, expected {ABI_VERSION}")
    if handle.nw_desc_bytes() != ctypes.sizeof(NwDesc):
        raise NativeError(
            f"nw_desc layout mismatch: library {handle.nw_desc_bytes()} B, binding "
            f"{ctypes.sizeof(NwDesc)} B"
        )
    _LIB = handle
    return handle


K_STEP, K_ADJOINT, K_SAMPLE, K_INJECT, K_STEPS, K_ADJOINT_STEPS, K_FILTER = 0, 1, 2, 3, 4, 5, 6
OPT_XCD_MAP = 0
OPT_ROWS_PER_ITER = 1
OPT_LDS_TILES_STEP = 2
OPT_LDS_TILES_ADJOINT = 3
OPT_GRAPH = 4
OPT_GRAPH_COUNT = 5
OPT_TIME_BLOCK = 6
OPT_KEEP_LAP = 7
OPT_TIME_BLOCK_ADJOINT = 8
OPT_FAST_3D = 9
OPT_FAST_F64 = 10
OPT_WORK_QUEUE = 11
OPT_WORK_QUEUE_ITEMS = 12


def profile_enable(on: bool) -> None:
    check(lib().nw_profile_enable(int(on)), "nw_profile_enable")


def profile_read(kind: int) -> tuple[float, int]:
    """Summed duration (ms) and launch count of one kernel kind since profiling was enabled."""
    ms, n = c_double(), c_int64()
    check(lib().nw_profile_read(kind, ctypes.byref(ms), ctypes.byref(n)), "nw_profile_read")
    return ms.value, n.value


def profile_units(kind: int) -> int:
    """Time steps advanced (samples recorded, for the sample kind) by those launches."""
    u = c_int64()
    check(lib().nw_profile_units(kind, ctypes.byref(u)), "nw_profile_units")
    return u.value


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().nw_last_error()
        raise NativeError(f"{what} failed (code {rc}): {msg.decode() if msg else ''}")


def compute_device(t: torch.Tensor) -> torch.device:
    """Device the HIP kernels run on for an input living on ``t.device``.

    GPU tensors stay where they are. CPU tensors (nefi's data generators call the operators with
    CPU float64 fields) are staged to the current GPU; without a GPU this raises.
    """
    if t.is_cuda:
        return t.device
    if not torch.cuda.is_available():
        raise NativeError(
            "the nefi_amd wave operators run on a ROCm GPU (HIP kernels for gfx950) and none is "
            "visible; there is no CPU fallback"
        )
    return torch.device("cuda", torch.cuda.current_device())


def stream_of(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def ptr_array(ptrs) -> ctypes.Array:
    arr = (c_void_p * max(1, len(ptrs)))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr
