"""Tile autotuner of the fast 2-D kernels.

For a plan the fast kernels cover, time a few forward steps and a few reverse steps for every
candidate tile — rows of the strip one wavefront marches down × shots one workgroup advances
in turn, and for the forward step the temporal blocking depth (steps advanced per launch) —
and keep the fastest per kernel. Forward candidates are timed both as a plain sweep and as a
sweep that keeps every level (the recompute of the reverse sweep), per step advanced. Timing
uses the library's own per-launch HIP events (``nw_profile_*``), so it measures the kernels
alone. Results are cached in a JSON file keyed by the candidate set's version, device name,
padded shape, shot count, order and absorbing layer
(``$NEFI_AMD_TILE_CACHE``, default ``~/.cache/nefi_amd/tiles.json``); ``NEFI_AMD_AUTOTUNE=0``
keeps the library's defaults. Choices tuned on an MI355X for common workloads ship with the
package (``tiles_mi355x.json``, keyed by the architecture without its feature suffixes); a plan
whose key is in neither table is tuned on first use. Timing ties between candidates make a
fresh tuning differ from box to box, so the shipped table also keeps a workload's kernel
instances, and the counter profiles measured for them, fixed.

It tunes kernel launch shapes only and is unrelated to nefi's hyper-parameter package
``nefi/autotune``.
"""

from __future__ import annotations

import json
import os
from pathlib import Path

import torch

from .. import _native as nat
from ..errors import NativeError

ROWS = (4, 8, 16, 32, 64)
LDS_ROWS = (8, 16, 32)
PML_ROWS = (8, 16)  # PML plans step with LDS-tiled face kernels of these heights only
GROUPS = (1, 2, 4, 8, 16)
DEPTHS = (2, 3, 4)  # temporal blocking depths
BLOCK_STEPS = 12  # forward steps per timing: a multiple of every depth
# below this many cell-updates per step a launch is latency-bound and every tile ties
MIN_WORK = 1 << 18
# the forward choice for such plans: 16-row LDS tiles, one shot per workgroup, depth 4
# (a 168² single-shot forward: 4.7 µs per step against 7.9 at one launch per step)
SMALL_STEP = (16, 1, 1, 1, 4)
XCD_MARGIN = 0.98  # the plain tile order replaces the XCD map only when 2 % faster


def _cache_path() -> Path:
    p = os.environ.get("NEFI_AMD_TILE_CACHE")
    return Path(p) if p else Path.home() / ".cache" / "nefi_amd" / "tiles.json"


def _load() -> dict:
    try:
        return json.loads(_cache_path().read_text())
    except (OSError, ValueError):
        return {}


SHIPPED = Path(__file__).with_name("tiles_mi355x.json")


def _shipped() -> dict:
    try:
        return json.loads(SHIPPED.read_text())
    except (OSError, ValueError):
        return {}


def _store(table: dict) -> None:
    path = _cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(table, indent=1, sort_keys=True))
        tmp.replace(path)
    except OSError:
        pass


def plan_key(plan, base_arch: bool = False) -> str:
    """Cache key of a plan; ``base_arch`` drops the architecture's feature suffixes
    (``gfx950:sramecc+:xnack-`` -> ``gfx950``), as the shipped table's keys do."""
    h = plan.host
    if plan.device.type == "cuda":
        # the architecture name and CU count, which (unlike the marketing name) profilers
        # that wrap the process still report
        props = torch.cuda.get_device_properties(plan.device)
        arch = getattr(props, "gcnArchName", "") or props.name
        if base_arch:
            arch = arch.split(":")[0]
        name = f"{arch}/{props.multi_processor_count}"
    else:
        name = "cpu"
    return f"v3|{name}|{h.dims[1]}x{h.dims[2]}|S{h.n_shots}|o{h.order}|a{h.absorbing}"


def candidates(kind: int, n_shots: int, pml: bool = False):
    """(rows, shots per group, rows loaded per iteration, LDS tiles, blocking depth) to try."""
    groups = [g for g in GROUPS if g <= n_shots]
    if pml:  # no register-march or blocked variants carry the face fields
        return [(r, g, 1, 1, 1) for r in PML_ROWS for g in groups]
    out = [(r, g, u, 0, 1) for r in ROWS for g in groups
           for u in ((1, 2) if kind == 0 else (1,))]
    out += [(r, g, 1, 1, 1) for r in LDS_ROWS for g in groups]
    if kind == 0:
        out += [(r, g, 1, 1, t) for r in LDS_ROWS for g in groups for t in DEPTHS]
    # (blocked reverse launches need kept Laplacians, whose extra writes in the keep sweep
    # outweighed their gain when measured; they stay an explicit plan option)
    return out


def _apply(plan, kind: int, rows: int, spw: int, rpi: int, lds: int, depth: int = 1) -> None:
    plan.set_tiles(kind, rows, spw)
    if kind == 0:
        plan.set_option(nat.OPT_ROWS_PER_ITER, rpi)
        plan.set_option(nat.OPT_TIME_BLOCK, depth)
    else:
        plan.set_option(nat.OPT_TIME_BLOCK_ADJOINT, depth)
    plan.set_option(nat.OPT_LDS_TILES_STEP if kind == 0 else nat.OPT_LDS_TILES_ADJOINT, lds)


def _time_kind(plan, kind: int, cand, steps: int) -> float:
    """Time (ms) per step of one kernel kind at one candidate."""
    P = plan
    _apply(P, kind, *cand)
    cells = P.host.cells
    m = P.zeros(cells)
    state, psi = P.new_state()
    nat.profile_enable(True)  # (timed calls never go through graphs)
    if kind == 0:
        keep = P.zeros((steps + 1) * P.keep_slot)
        P.forward(m, state, psi, 0, steps)
        P.forward(m, state, psi, steps, 2 * steps, keep=keep)
        ms = nat.profile_read(nat.K_STEP)[0] + nat.profile_read(nat.K_STEPS)[0]
        n = nat.profile_units(nat.K_STEP) + nat.profile_units(nat.K_STEPS)
    else:
        keep = P.zeros((steps + 1) * P.keep_slot)
        lam = [P.zeros(P.shot_numel), P.zeros(P.shot_numel)]
        grad = P.zeros(P.grad_slices() * cells, torch.float64)
        res = P.zeros(max(1, P.host.n_trace_rows * P.host.n_t))  # residual injection included
        P.adjoint(m, lam, psi, 0, steps, keep=keep, grad=grad, residual=res)
        kinds = (nat.K_ADJOINT, nat.K_ADJOINT_STEPS)
        ms = sum(nat.profile_read(k)[0] for k in kinds) + nat.profile_read(nat.K_INJECT)[0]
        n = sum(nat.profile_units(k) for k in kinds)
    nat.profile_enable(False)
    return ms / max(1, n)


def tune(plan, steps: int = 3, force: bool = False) -> dict | None:
    """Pick and set the fastest tiles of ``plan`` (a ``DevicePlan``); returns the choice."""
    if os.environ.get("NEFI_AMD_AUTOTUNE", "1") == "0":
        return None
    if plan.tiles(0)[0] == 0:  # the fast kernels do not cover this plan
        return None
    h = plan.host
    if h.cells * h.n_shots < MIN_WORK and not force:
        # launch-bound: no timing, but fewer launches (4 steps per forward launch)
        small = (16, 1, 1, 1, 1) if plan.pml else SMALL_STEP
        _apply(plan, 0, *small)
        return {"step": dict(zip(("rows", "shots_per_group", "rows_per_iter", "lds", "depth"),
                                 small)), "static": True}
    key = plan_key(plan)
    table = _load()
    shipped = _shipped().get(plan_key(plan, base_arch=True))
    if key in table and not force:
        choice = dict(table[key], source="cache")
    elif shipped is not None and not force:
        choice = dict(shipped, source="shipped")
    else:
        steps = min(steps, max(1, h.n_steps))
        best_total, choice = None, None
        for xcd in (1, 0):
            plan.set_option(nat.OPT_XCD_MAP, xcd)
            pick = {"xcd_map": xcd, "table": {}}
            for kind, name in ((0, "step"), (1, "adjoint")):
                timings = {}
                n = min(BLOCK_STEPS, max(1, h.n_steps))
                for cand in candidates(kind, h.n_shots, plan.pml):
                    try:
                        _time_kind(plan, kind, cand, 1)  # warm the candidate once
                        timings[cand] = _time_kind(plan, kind, cand, n)
                    except NativeError:  # a variant this device refuses to launch
                        nat.profile_enable(False)
                (rows, spw, rpi, lds, depth), best = min(timings.items(), key=lambda kv: kv[1])
                pick[name] = {"rows": rows, "shots_per_group": spw, "rows_per_iter": rpi,
                              "lds": lds, "depth": depth, "ms": best}
                pick["table"][name] = {
                    f"{r}x{g}/{u}{'L' if l else ''}{f'T{t}' if t > 1 else ''}": round(ms, 4)
                    for (r, g, u, l, t), ms in timings.items()}
            # a gradient launches about two forward steps (sweep + recompute) per reverse step;
            # the XCD map (tried first) keeps a tile's halo rows in its XCD's L2, so the plain
            # order must be clearly faster to win (it moves more HBM bytes)
            total = 2.0 * pick["step"]["ms"] + pick["adjoint"]["ms"]
            if best_total is None or total < XCD_MARGIN * best_total:
                best_total, choice = total, pick
        table[key] = choice
        _store(table)
        choice = dict(choice, source="tuned")
    plan.set_option(nat.OPT_XCD_MAP, int(choice.get("xcd_map", 1)))
    for kind, name in ((0, "step"), (1, "adjoint")):
        c = choice[name]
        _apply(plan, kind, c["rows"], c["shots_per_group"], c.get("rows_per_iter", 1),
               c.get("lds", 0), c.get("depth", 1))
    return choice
