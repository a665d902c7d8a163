"""Host-side logic of the package (CPU, no GPU): the C ABI surface, geometry and tables that
the operators hand to the library, the reference's error behaviour, and the generic helpers of
nefi/physics/timestep.py."""

from __future__ import annotations

import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest
import torch

from _support import load, oracle_problem
from nefi_amd import _native as nat
from nefi_amd.domain import Domain
from nefi_amd.errors import ConfigError, NativeError, OperatorError, ShapeError
from nefi_amd.physics import timestep as ts
from nefi_amd.physics.wave import (
    ABSORBING,
    GRAD_MODES,
    WaveGrid,
    WaveInitialConditionOperator,
    WaveOperator,
    laplacian_kernel,
    ricker,
    stable_dt,
)
from nefi_amd.registry import build, get

ROOT = Path(__file__).resolve().parent.parent


# ---- C ABI ---------------------------------------------------------------------------------
def test_library_loads_and_exports_every_declared_symbol():
    def declared_in(name):
        text = (ROOT / "include" / name).read_text()
        return set(re.findall(r"^\s*(?:int32_t|int64_t|const char\*)\s+(nw_\w+)\(", text, re.M))

    assert declared_in("nefi_wave.h") == set(nat.EXPORTS)
    assert declared_in("nefi_filter.h") == set(nat.FILTER_EXPORTS)
    assert declared_in("nefi_stepper.h") == set(nat.STEPPER_EXPORTS)
    assert sorted(p.name for p in (ROOT / "include").glob("*.h")) == [
        "nefi_filter.h", "nefi_stepper.h", "nefi_wave.h"]
    declared = set(nat.EXPORTS) | set(nat.FILTER_EXPORTS) | set(nat.STEPPER_EXPORTS)
    lib = nat.lib()
    for name in declared:
        assert getattr(lib, name) is not None
    assert lib.nw_abi_version() == nat.ABI_VERSION
    assert lib.nw_desc_bytes() == ctypes.sizeof(nat.NwDesc)


def test_plan_create_rejects_bad_descriptors_without_a_gpu():
    lib = nat.lib()
    d = nat.NwDesc()
    h = ctypes.c_void_p()
    d.ndim = 4
    assert lib.nw_plan_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b"ndim" in lib.nw_last_error()
    d.ndim, d.order, d.n_shots, d.substeps, d.n_t = 2, 3, 1, 1, 1
    assert lib.nw_plan_create(ctypes.byref(d), ctypes.byref(h)) == 1
    assert b"order" in lib.nw_last_error()


def test_operators_raise_without_gpu_instead_of_falling_back():
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    meta, z = load("pml2d")
    op = WaveOperator(Domain(meta["shape"], meta["extent"]), meta["sources"], meta["receivers"],
                      **meta["kw"])
    with pytest.raises(NativeError, match="no CPU fallback"):
        op({"c": torch.tensor(z["in_c"])})


# ---- geometry against the independently written oracle geometry --------------------------
@pytest.mark.parametrize("name", ["cfg1", "pml2d", "cerjan2d", "pml1d_o2", "pat3d"])
def test_grid_and_stencils_match_oracle(name):
    meta, z = load(name)
    P = oracle_problem(meta, z)
    op_kw = {k: v for k, v in meta["kw"].items()}
    dom = Domain(meta["shape"], meta["extent"])
    if meta["kind"] == "fwi":
        op = WaveOperator(dom, meta["sources"], meta["receivers"], **op_kw)
        si, sw = op.grid.point_weights(op.sources)
        oi, ow = P.point_weights(np.asarray(meta["sources"]))
        assert np.array_equal(si.numpy(), oi) and np.allclose(sw.numpy(), ow, rtol=0, atol=0)
    else:
        op_kw["c"] = torch.as_tensor(z["in_kw_c"]) if "in_kw_c" in z else op_kw["c"]
        op = WaveInitialConditionOperator(dom, meta["receivers"], **op_kw)
    g = op.grid
    assert g.n_abs == P.n_abs and g.padded_shape == P.padded
    assert op.substeps == P.substeps and op.n_steps == P.n_steps and op.dt == P.dt
    ri, rw = op._rec
    assert np.array_equal(ri.numpy(), P.rec_idx) and np.array_equal(rw.numpy(), P.rec_w)


def test_point_weights_edge_clamp_and_outside_error():
    g = WaveGrid(Domain((4, 5), ((0.0, 4.0), (0.0, 5.0))), absorb_width=1.0)
    assert g.padded_shape == (6, 7) and g.lo_pad == (-1.0, -1.0)
    # the last cell centre: i0 clamped to m-2, weights (0, 1)
    idx, w = g.point_weights([[4.5, 5.5]])
    assert idx.tolist() == [[4 * 7 + 5, 4 * 7 + 6, 5 * 7 + 5, 5 * 7 + 6]]
    assert np.allclose(w.numpy(), [[0, 0, 0, 1]])
    with pytest.raises(ConfigError, match="outside"):
        g.point_weights([[-1.0, 0.0]])
    assert g.nearest_index([[0.1, 0.1]]).tolist() == [1 * 7 + 1]


def test_host_profiles_match_reference_formulas():
    """The 1-D profiles handed to the kernels recombine into nefi's full-grid coefficients."""
    dom = Domain((12, 10), ((0.0, 1.2), (0.0, 1.0)))
    for absorbing in ("sponge", "pml", "cerjan"):
        op = WaveOperator(dom, [[0.6, 0.5]], [[0.3, 0.4]], n_t=10, dt_obs=0.02, f0=5.0, c_max=2.0,
                          absorbing=absorbing, absorb_width=0.3)
        hp = op._host_plan(1)
        P = oracle_problem({"kind": "fwi", "shape": (12, 10), "extent": ((0.0, 1.2), (0.0, 1.0)),
                            "sources": [[0.6, 0.5]], "receivers": [[0.3, 0.4]],
                            "kw": dict(n_t=10, dt_obs=0.02, f0=5.0, c_max=2.0,
                                       absorbing=absorbing, absorb_width=0.3)}, None)
        if absorbing in ("sponge", "pml"):
            half = hp.damp[1][:, None] + hp.damp[2][None, :]
            assert np.allclose((1.0 / (1.0 + half)).numpy(), P.a, rtol=1e-15)
            assert np.allclose((1.0 - half).numpy(), P.b, rtol=1e-15)
        else:
            tap = torch.minimum(hp.damp[1][:, None], hp.damp[2][None, :])
            assert np.array_equal(tap.numpy(), P.taper)
        if absorbing == "pml":
            zz = op.dt**2 * hp.pml_zc[1][:, None] * hp.pml_zc[2][None, :]
            assert np.allclose(zz.numpy(), P.zz, rtol=1e-14, atol=0)
            for k, ax in ((1, 0), (2, 1)):
                other = hp.pml_zc[3 - k]
                shape = [1, 1]
                shape[1 - ax] = -1
                beta = (other.view(shape) - hp.pml_zf[k].view([-1 if i == ax else 1
                                                                 for i in range(2)])) \
                    * hp.pml_kappa[k].view([-1 if i == ax else 1 for i in range(2)])
                assert np.allclose(np.broadcast_to(beta.numpy(), P.beta[ax].shape), P.beta[ax],
                                   rtol=1e-13, atol=1e-300)
                alpha = hp.pml_alpha[k].view([-1 if i == ax else 1 for i in range(2)])
                assert np.allclose(np.broadcast_to(alpha.numpy(), P.alpha[ax].shape),
                                   P.alpha[ax], rtol=1e-15)


def test_injection_tables_are_the_transpose_of_the_receiver_stencils():
    from nefi_amd.physics._engine import injection_tables

    idx = torch.tensor([[3, 4, 8, 9], [4, 5, 9, 10], [20, 21, 25, 26]])
    w = torch.rand(3, 4, dtype=torch.float64)
    first, count, cell, off, rec, ww = injection_tables(idx, w, [(0, 2), (2, 1), (0, 2)])
    assert first.tolist() == [0, 6, 0] and count.tolist() == [6, 4, 6]
    dense = np.zeros((2, 30))
    for u in range(6):
        for e in range(int(off[u]), int(off[u + 1])):
            dense[int(rec[e]), int(cell[u])] += float(ww[e])
    want = np.zeros((2, 30))
    for r in range(2):
        for k in range(4):
            want[r, int(idx[r, k])] += float(w[r, k])
    assert np.allclose(dense, want)


# ---- API and errors (wave.py:369-378, 561-569, 711-712) -------------------------------------
def test_api_constants_errors_and_registry():
    assert ABSORBING == ("pml", "sponge", "cerjan", "none")
    assert set(GRAD_MODES) >= {"checkpoint", "autograd"}
    dom = Domain((16, 16), ((0.0, 1.0), (0.0, 1.0)))
    kw = dict(n_t=20, dt_obs=0.01, f0=5.0, c_max=2.5)
    with pytest.raises(ConfigError):
        WaveOperator(dom, [[0.5, 0.5]], [[0.2, 0.2]], absorbing="foo", **kw)
    with pytest.raises(ConfigError):
        WaveOperator(dom, [[0.5, 0.5]], [[0.2, 0.2]], grad_mode="foo", **kw)
    with pytest.raises(ConfigError):
        WaveOperator(dom, [[0.5, 0.5]], [[0.2, 0.2]], absorb_R=1.5, **kw)
    with pytest.raises(ConfigError):
        WaveOperator(Domain((4, 4, 4), ((0, 1),) * 3), [[0.5] * 3], [[0.2] * 3], **kw)
    with pytest.raises(ShapeError):
        WaveOperator(dom, [[0.5, 0.5, 0.5]], [[0.2, 0.2]], **kw)
    op = WaveOperator(dom, [[0.5, 0.5], [0.4, 0.4]], [[0.2, 0.2]] * 3, **kw)
    assert op.output_shape((16, 16)) == (2, 3, 20)
    assert op.fidelity_tag == "wave-leapfrog-o4-pml" and op.traceable is False
    assert op.record == [j * op.substeps for j in range(20)]
    coarse = op.at_resolution((8, 8))
    assert coarse is op.at_resolution((8, 8)) and coarse.fidelity_tag == op.fidelity_tag
    with pytest.raises(ShapeError):
        op._prepare(torch.ones(8, 8), None)
    with pytest.raises(OperatorError, match="exceeds"):
        op._check_speed(torch.full((16, 16), 3.0))
    with pytest.raises(OperatorError, match="positive"):
        op._check_speed(torch.zeros(16, 16))
    assert get("operator", "wave") is WaveOperator
    assert isinstance(build("operator", {"type": "wave", "domain": dom, "sources": [[0.5, 0.5]],
                                         "receivers": [[0.2, 0.2]], **kw}), WaveOperator)
    ic = WaveInitialConditionOperator(dom, [[0.2, 0.2]], n_t=10, dt_obs=0.01, c=1.0)
    assert ic.homogeneity == 1.0 and ic.output_shape((16, 16)) == (1, 10)


def test_ricker_stable_dt_and_kernel():
    t = torch.linspace(0, 1, 2001, dtype=torch.float64)
    w = ricker(t, 5.0)
    assert abs(float(w.max()) - 1.0) < 1e-6 and abs(float(t[w.argmax()]) - 0.24) < 1e-3
    assert math.isclose(stable_dt(1.0, (0.1, 0.1), order=2, courant=1.0), 0.1 / math.sqrt(2))
    assert math.isclose(stable_dt(1.0, (0.1, 0.1), order=4, courant=1.0), 0.1 * math.sqrt(3 / 8))
    k = laplacian_kernel((0.5, 0.25), 4, dtype=torch.float64)[0, 0]
    assert k.shape == (5, 5) and math.isclose(float(k[2, 2]), -2.5 / 0.25 - 2.5 / 0.0625)
    assert math.isclose(float(k[0, 2]), -1 / 12 / 0.25) and math.isclose(float(k[2, 3]), 4 / 3 / 0.0625)


# ---- generic time stepping (timestep.py:78-169) --------------------------------------------
def test_run_timestepping_record_and_checkpoint_gradients():
    k = torch.tensor([0.5, 1.0, 2.0], dtype=torch.float64, requires_grad=True)

    def step(s, p, n, dt):
        return (s[0] - dt * p[0] * s[0],)

    u0 = torch.ones(3, dtype=torch.float64)
    (uT,), obs = ts.run_timestepping(step, u0, (k,), 100, 0.01, lambda s, n: s[0].clone(),
                                     record=[0, 50, 100])
    assert len(obs) == 3 and torch.allclose(obs[0], u0)
    assert torch.allclose(uT.detach(), (1 - 0.01 * k.detach()) ** 100, rtol=1e-12)
    g1 = torch.autograd.grad(sum(o.sum() for o in obs), k)[0]
    _, obs2 = ts.run_timestepping(step, u0, (k,), 100, 0.01, lambda s, n: s[0].clone(),
                                  record=[0, 50, 100], checkpoint_every=7)
    g2 = torch.autograd.grad(sum(o.sum() for o in obs2), k)[0]
    assert torch.allclose(g1, g2, rtol=1e-12, atol=0)
    assert ts.auto_checkpoint_every(100) == 10
    with pytest.raises(ConfigError):
        ts.run_timestepping(step, u0, (k,), 5, 0.1, record=[7])
    assert math.isclose(ts.stable_dt_diffusion(1.0, (0.1, 0.1), safety=1.0), 0.1**2 / 4)


def test_shipped_tile_table_holds_valid_choices():
    from nefi_amd.physics import autotune as at

    table = at._shipped()
    assert table, "the shipped tile table is missing or unreadable"
    for key, choice in table.items():
        version, device, shape, shots, order, absorbing = key.split("|")
        assert version == "v3" and ":" not in device and "/" in device
        assert shots.startswith("S") and order in ("o2", "o4") and absorbing[0] == "a"
        assert choice["xcd_map"] in (0, 1)
        for kind, name in ((0, "step"), (1, "adjoint")):
            c = choice[name]
            cands = at.candidates(kind, int(shots[1:]), absorbing == "a3")
            assert (c["rows"], c["shots_per_group"], c["rows_per_iter"], c["lds"],
                    c["depth"]) in cands, (key, name)


def test_snapshot_interval_counts_the_held_keep_buffers(monkeypatch):
    """The gradient's snapshot interval falls back to ceil(sqrt(N)) only when a keep buffer of
    K + 1 wavefields does not fit in 60 % of the available memory, and the workspace's own keep
    buffers (reused by the next gradient) count as available. The 64-shot 1064^2 workload with
    K = 200 (58.25 GB per keep buffer) and four buffers held leaves ~40 GB free: without the
    held bytes the next gradient would silently snapshot every 64 steps."""
    from types import SimpleNamespace

    from nefi_amd.physics._engine import grad_block

    plan = SimpleNamespace(device=torch.device("cuda", 0), keep_slot=1064 * 1064 * 64, esize=4)
    keep = 201 * plan.keep_slot * plan.esize
    free = 40 * 10**9
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda device=None: (free, 288 * 10**9))
    assert grad_block(plan, 4000, "checkpoint", 200) == 64  # nothing held: does not fit
    assert grad_block(plan, 4000, "checkpoint", 200, held=4 * keep) == 200
    free = 20 * 10**9
    assert grad_block(plan, 4000, "checkpoint", 200, held=keep) == 64  # 0.6 x 78 GB < 58 GB
    free = 285 * 10**9
    assert grad_block(plan, 4000, "checkpoint", 200) == 200
    assert grad_block(plan, 4000, "autograd", None, held=0) == 64  # store-all does not fit


def test_kept_blocks_fall_back_to_the_one_stream_sweep_when_two_buffers_do_not_fit(monkeypatch):
    """The overlapped reverse sweep needs two keep buffers. With 58.25 GB buffers (64 shots of
    1064^2, K = 200) and ~100 GB free, one fits and two do not: the forward sweep keeps one
    block and the reverse sweep runs on one stream (it used to allocate a second buffer and
    run out of memory). Buffers the workspace already holds count as available."""
    from types import SimpleNamespace

    from nefi_amd.physics._engine import Propagator

    monkeypatch.delenv("NEFI_AMD_OVERLAP", raising=False)
    monkeypatch.delenv("NEFI_AMD_KEEP_BLOCKS", raising=False)
    plan = SimpleNamespace(device=torch.device("cuda", 0), keep_slot=1064 * 1064 * 64, esize=4,
                           dtype=torch.float32, tiles=lambda kind: (16, 8),
                           option=lambda opt: 1)
    prop = Propagator.__new__(Propagator)
    prop.plan = plan
    size = 201 * plan.keep_slot
    gb = 10**9
    free = [0]
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda device=None: (free[0], 288 * gb))

    def kept(free_gb, held=()):
        free[0] = int(free_gb * gb)
        ws = SimpleNamespace(plan=plan, bufs={k: SimpleNamespace(numel=lambda: size)
                                              for k in held})
        return prop._kept_blocks(ws, 200, 20)

    assert kept(100) == 1 and kept(116) == 1  # one buffer fits, two do not
    assert kept(150) == 2
    assert kept(285) == 4
    assert kept(20, held=("keep", "keep2")) == 2  # the held pair is reused
    assert kept(20, held=("keep",)) == 1
    assert kept(30, held=("keep", "keep2", "keep3", "keep4")) == 4


def test_errors_are_nefis_own_classes_when_nefi_is_importable(tmp_path):
    """With nefi importable (the plugin case) the wave path raises nefi's classes, so nefi's
    ``except NefiError`` (cli.py) catches them: checked in a subprocess against a stand-in
    ``nefi.errors`` module with nefi's class names. Standalone, the local hierarchy is used."""
    import subprocess
    import sys

    pkg = tmp_path / "nefi"
    pkg.mkdir()
    (pkg / "__init__.py").write_text("")
    (pkg / "errors.py").write_text(
        "class NefiError(Exception): pass\n"
        "class ShapeError(NefiError, ValueError): pass\n"
        "class ConfigError(NefiError, ValueError): pass\n"
        "class RegistryError(NefiError, KeyError): pass\n"
        "class SolverError(NefiError, RuntimeError): pass\n"
        "class OperatorError(NefiError, RuntimeError): pass\n")
    code = (
        "import sys, torch; sys.path[:0] = [sys.argv[1], sys.argv[2]]\n"
        "import nefi.errors as ne, nefi_amd.errors as e\n"
        "from nefi_amd.domain import Domain\n"
        "from nefi_amd.physics.wave import WaveOperator\n"
        "assert e.BOUND_TO_NEFI and e.OperatorError is ne.OperatorError\n"
        "assert e.ConfigError is ne.ConfigError and issubclass(e.NativeError, ne.NefiError)\n"
        "op = WaveOperator(Domain((8, 8), ((0., 1.), (0., 1.))), [[.5, .5]], [[.2, .2]],"
        " n_t=4, dt_obs=0.01, f0=5.0, c_max=2.0, absorbing='sponge', absorb_width=0.25)\n"
        "try:\n"
        "    op({'c': torch.full((8, 8), 3.0)})\n"
        "except ne.NefiError as err:\n"
        "    assert 'exceeds' in str(err); print('caught', type(err).__name__)\n"
    )
    root = str(Path(__file__).resolve().parent.parent)
    r = subprocess.run([sys.executable, "-c", code, str(tmp_path), root], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "caught OperatorError" in r.stdout
    import nefi_amd.errors as local

    assert not local.BOUND_TO_NEFI and issubclass(local.OperatorError, local.NefiError)


def test_domain_contract():
    from nefi_amd.domain import Domain
    from nefi_amd.errors import ShapeError

    d = Domain((4, 5), ((0.0, 2.0), (-1.0, 1.5)))
    assert d.ndim == 2 and d.shape == (4, 5) and d.spacing() == (0.5, 0.5)
    assert d.spacing((8, 10)) == (0.25, 0.25)
    assert d.at((8, 10)) == Domain((8, 10), d.extent) and hash(d.at((4, 5))) == hash(d)
    x = d.physical_coords(dtype=torch.float64)
    assert x.shape == (4, 5, 2)
    assert torch.allclose(x[:, 0, 0], torch.tensor([0.25, 0.75, 1.25, 1.75], dtype=torch.float64))
    assert torch.allclose(x[0, :, 1], torch.tensor([-0.75, -0.25, 0.25, 0.75, 1.25],
                                                   dtype=torch.float64))
    u = d.coords(dtype=torch.float64)
    assert torch.allclose(u[:, 0, 0], torch.tensor([-0.75, -0.25, 0.25, 0.75], dtype=torch.float64))
    assert Domain.unit((8,)).extent == ((0.0, 1.0),)
    with pytest.raises(ShapeError):
        Domain((4,), ((1.0, 0.0),))
    with pytest.raises(ShapeError):
        Domain((4, 4), ((0.0, 1.0),))
    with pytest.raises(ShapeError):
        d.at((4,))
    with pytest.raises(AttributeError):
        d.shape = (1, 1)


def test_sensor_bandwidth_attributes_and_resolution_copies():
    from nefi_amd.instances.photoacoustic3d import SensorBandwidth
    from nefi_amd.operators.base import Operator

    class Inner(Operator):
        primary, homogeneity = "p0", 1.0

        def __init__(self, n):
            super().__init__()
            self.n = n
            self.fidelity_tag = "wave-leapfrog-o2-sponge"

        def at_resolution(self, shape):
            return self if tuple(shape) == (self.n,) else Inner(shape[0])

        def output_shape(self, shape):
            return (3, 7)

    op = SensorBandwidth(Inner(16), 0.01, 12.5)
    assert op.fidelity_tag == "wave-leapfrog-o2-sponge-lp12.5"
    assert op.primary == "p0" and op.homogeneity == 1.0 and op.required_fields() == ("p0",)
    assert op.output_shape((16,)) == (3, 7) and not op.traceable
    assert op.at_resolution((16,)) is op
    half = op.at_resolution((8,))
    assert half is op.at_resolution([8]) and half.inner.n == 8
    assert half.fidelity_tag == op.fidelity_tag and half.f_cutoff == 12.5
    ideal = SensorBandwidth(Inner(16), 0.01, None)
    assert ideal.fidelity_tag == "wave-leapfrog-o2-sponge" and ideal.f_cutoff is None
    x = torch.ones(2, 5)
    assert ideal._filter(x) is x


def _blocked_lds_phases(waves, region_rows, R, levels, shots, quiet):
    """The LDS accesses of one k4_steps / k4_queue workgroup (nefi_wave_2d.hip, k4_tile) as
    barrier intervals: per interval, per wave, the (buffer, row, float4 column) sets read and
    written, plus the elements an asynchronous region fill (quiet tiles) may write until the next
    shot's full wait. Row numbers are LDS rows (R zero rows, then the region's computed rows,
    then R zero rows); float4 columns 0 and 65 are the margins, 1..64 lanes 0..63's cells."""
    mr = -(-region_rows // waves)
    crow = waves * mr
    lrows = crow + 2 * R
    lds_w4 = 66  # float4 per LDS row (kLW / 4)
    threads = 64 * waves

    def row(b, r, cols=range(lds_w4)):  # the float4 elements of LDS row r of buffer b
        return {(b, r, c) for c in cols}

    def tile_elems(b, wave):  # a wave's own region rows, lane l at float4 l + 1 (no margins)
        out = set()
        for k in range(mr):
            out |= row(b, R + wave * mr + k, range(1, 65))
        return out

    def zeroing(wave):  # zero_margin_rows: the R rows on either side, then thread r's margins
        out = set()
        for b in (0, 1):
            for i in range(64 * wave, 64 * (wave + 1)):
                j = i
                while j < R * lds_w4 * 4:  # float by float, threads strided
                    out |= {(b, j // (lds_w4 * 4), (j // 4) % lds_w4),
                            (b, lrows - R + j // (lds_w4 * 4), (j // 4) % lds_w4)}
                    j += threads
                r = R + i
                while r < lrows - R:
                    out |= {(b, r, 0), (b, r, lds_w4 - 1)}
                    r += threads
        return out

    buf = lambda s: (s % 2) if quiet else 0  # noqa: E731
    # the zeroing, and for quiet tiles the first shot's fill issued behind it, before the first
    # barrier
    phases = [([set() for _ in range(waves)], [zeroing(w) for w in range(waves)],
               set().union(*(tile_elems(0, w) for w in range(waves))) if quiet else None)]
    for s in range(shots):
        b = buf(s)
        if not quiet:  # synchronous store of the region after a barrier
            phases.append(([set() for _ in range(waves)],
                           [tile_elems(b, w) for w in range(waves)], None))
        # after the barrier: cur rows, then level 1 reads (windows), and for quiet tiles the
        # next shot's fill issued into the other buffer
        nxt = None
        if quiet and s + 1 < shots:
            nxt = set().union(*(tile_elems(buf(s + 1), w) for w in range(waves)))
        for lvl in range(1, levels + 1):
            reads = []
            for w in range(waves):
                c0 = w * mr
                rr = set()
                for k in range(mr if lvl == 1 else 0):
                    rr |= row(b, R + c0 + k)
                for j in range(mr + 2 * R):  # the window rows, x-neighbours included
                    rr |= row(b, c0 + j)
                reads.append(rr)
            phases.append((reads, [set() for _ in range(waves)], nxt))
            if lvl < levels:  # level barrier, write the new level, barrier, tap reads
                writes = []
                for w in range(waves):
                    ww = set()
                    for k in range(mr):  # lane l writes float4 l + 1 of its rows
                        ww |= row(b, R + w * mr + k, range(1, 65))
                    writes.append(ww)
                phases.append(([set() for _ in range(waves)], writes, nxt))
                taps = set()
                for r in range(R * levels, crow - R * levels):
                    taps |= row(b, R + r)
                phases.append(([set(taps) for _ in range(waves)], [set() for _ in range(waves)],
                               nxt))
        # the next shot starts with a full wait for its fill and a barrier
    return phases


@pytest.mark.parametrize("waves", [4, 8, 12, 16])
@pytest.mark.parametrize("shots", [1, 2, 3, 4])
@pytest.mark.parametrize("quiet", [False, True])
def test_blocked_forward_lds_protocol_has_no_races(waves, shots, quiet):
    """Between two barriers of the temporally blocked forward kernel no wavefront writes an LDS
    row another wavefront reads or writes, and a quiet tile's asynchronous fill of the next
    shot's region never lands in the buffer being read: for any number of wavefronts and shots
    per workgroup (32-row tiles, depth 4, order 4: 48 region rows), including the 4-wavefront,
    2-shot workgroup that once gave wrong results (the protocol is the same source for every
    wave count, so that failure did not come from the barrier and level ordering)."""
    R, T = 2, 4
    for reads, writes, fill in _blocked_lds_phases(waves, 32 + 2 * R * T, R, T, shots, quiet):
        for w in range(waves):
            for v in range(waves):
                if v != w:
                    assert not (writes[w] & (reads[v] | writes[v])), (w, v)
            if fill:
                assert not (fill & (reads[w] | writes[w]))


def test_blocked_forward_lds_protocol_check_finds_a_missing_barrier():
    """The same check on a model with the barrier between a level's reads and the next level's
    writes removed (the two intervals merged) reports the race."""
    phases = _blocked_lds_phases(8, 48, 2, 4, 2, False)
    i = next(k for k, (r, w, _) in enumerate(phases) if any(r) and not any(w)
             and k + 1 < len(phases) and any(phases[k + 1][1]))
    reads = phases[i][0]
    writes = phases[i + 1][1]
    assert any(writes[w] & reads[v] for w in range(8) for v in range(8) if v != w)
