"""Strict-build parity where the earlier cases do not reach:

  * the 64-bit-offset kernel instances.  A level switches to them once its
    population buffer reaches 4 GiB (an FP32 level of about 38 M cells, the
    size of the finest level of the cube Re 1e5 workload); lb_config's
    wide_offsets forces them on small cases, so they are compared with the
    oracle like the 32-bit ones: FP32, FP16 and BF16 storage, the thermal
    DDF, and the batched engine;
  * 4 and 5 levels (the recursion of solver_control.jl:6-98 at depth 4 and
    5, ghost parity and temporal weights at lshift 3 and 4, the leaf-only
    density store), in FP32 and FP16 storage;
  * an engine mix of 2-, 3-, 4- and 5-level cases, each compared with its
    own oracle run (cases aligned on their finest level: a 2-level case
    advances 8 coarse steps per engine step of the 5-level one).

Bars: bit-identical populations, density and velocity on every level (no
wall model, so no transcendental function is involved).  The fast build's
wide instances equal its 32-bit ones bit for bit (same arithmetic, only the
address type differs) and stay within the fast-build bar of the oracle.
"""
import functools

import numpy as np
import pytest

from openludwig_amd import _native as N
from openludwig_amd import cases
from openludwig_amd.engine import Engine
from openludwig_amd.solver import Solver, build_case, ramp_velocity
from oracle import Oracle
from parity_helpers import compare, perturbed_state, run_pair


def _assert_bitexact(res, what):
    print(what, res)
    for k, (mx, nz, rel) in res.items():
        assert mx == 0.0, f"{what} {k}: max |diff| {mx} over {nz} entries"


@functools.lru_cache(maxsize=None)
def _case(res, levels, ddf=False, wall_model=False):
    cfg, mesh = cases.tiny_sphere(res, levels, ddf=ddf, wall_model=wall_model)
    if ddf:
        cfg = cfg.replace(t_inlet=1.05, t_initial=1.0)
    grids, params, mesh = build_case(cfg, mesh)
    assert len(grids) == levels, f"case ({res}, {levels}) built {len(grids)} levels"
    return cfg, grids, params, mesh


@pytest.mark.gpu
@pytest.mark.parametrize("levels,storage", [(2, 0), (3, 0), (3, 1), (3, 2)])
def test_wide_offsets_strict_bitexact(hip_device, levels, storage):
    cfg, grids, params, mesh = _case(8 if levels == 2 else 16, levels)
    print(f"wide offsets, {levels} levels, storage {storage}", flush=True)
    s, o = run_pair(cfg, grids, params, mesh, steps=6, strict=True, batch=3, storage=storage, wide_offsets=True)
    _assert_bitexact(compare(s, o, grids), f"wide {levels}L st{storage}")


@pytest.mark.gpu
@pytest.mark.parametrize("storage", [0, 1, 2])
def test_wide_offsets_thermal_strict_bitexact(hip_device, storage):
    cfg, grids, params, mesh = _case(8, 2, ddf=True)
    print(f"wide offsets, thermal DDF, storage {storage}", flush=True)
    s, o = run_pair(cfg, grids, params, mesh, steps=8, strict=True, storage=storage, wide_offsets=True)
    _assert_bitexact(compare(s, o, grids, fields=("f", "rho", "vel", "g", "T")), f"wide thermal st{storage}")


@pytest.mark.gpu
@pytest.mark.parametrize("storage", [0, 1])
def test_wide_offsets_fast_build_equals_narrow(hip_device, storage):
    """The benchmark's kernels: fast build with the wall model, 3 levels.
    Wide and 32-bit offsets give the same bits; both within 5e-5 of the
    FP32 oracle (FP32 storage)."""
    cfg, grids, params, mesh = _case(16, 3, wall_model=True)
    print(f"fast build, wide vs narrow, storage {storage}", flush=True)
    out = []
    for wide in (False, True):
        s, o = run_pair(cfg, grids, params, mesh, steps=8, strict=False, batch=4, storage=storage,
                        wide_offsets=wide)
        out.append({(g.level_id, f): s.download(g.level_id, f) for g in grids for f in ("f", "rho", "vel")})
        if storage == 0:
            res = compare(s, o, grids)
            print("wide" if wide else "narrow", res)
            for k, (mx, nz, rel) in res.items():
                assert mx <= 5e-5, f"{k}: max |diff| {mx}"
        s.close()
        o.close()
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), f"{k}: wide and narrow instances differ"


def _batched_pair(cfg, grids, params, mesh, steps, storage, **kw):
    """GPU (strict) and oracle from the same perturbed state, stepped in the
    main loop's batches of gpu_async_depth steps with the inlet velocity at
    each batch end (main.jl:214-218), as the engine steps a case."""
    s = Solver(cfg, grids, params, mesh, strict_fp=True, storage=storage, **kw)
    o = Oracle(cfg, grids, params, "f32", storage=storage)
    s.init_equilibrium()
    o.init_equilibrium()
    for g, f in zip(grids, perturbed_state(grids)):
        s.upload(g.level_id, "f", f)
        o.put(g.level_id, "f", f.astype(np.float64))
    b = cfg.gpu_async_depth
    for t in range(1, steps + 1, b):
        n = min(b, steps - t + 1)
        u = ramp_velocity(cfg.u_lattice, t + n - 1, cfg.ramp_steps)
        s.execute_timestep_batch(t, n, u)
        o.step_batch(t, n, u)
    s.synchronize()
    return s, o


@pytest.mark.gpu
@pytest.mark.parametrize("storage", [0, 1])
def test_strict_bitexact_four_levels(hip_device, storage):
    cfg, grids, params, mesh = _case(32, 4)
    print(f"4 levels {[g.n_cells for g in grids]} storage {storage}", flush=True)
    s, o = _batched_pair(cfg, grids, params, mesh, steps=3, storage=storage)
    _assert_bitexact(compare(s, o, grids), f"4L st{storage}")


@pytest.mark.gpu
@pytest.mark.parametrize("storage", [0, 1])
def test_strict_bitexact_five_levels(hip_device, storage):
    cfg, grids, params, mesh = _case(64, 5)
    print(f"5 levels {[g.n_cells for g in grids]} storage {storage}", flush=True)
    s, o = _batched_pair(cfg, grids, params, mesh, steps=2, storage=storage)
    _assert_bitexact(compare(s, o, grids), f"5L st{storage}")


@pytest.mark.gpu
def test_engine_mix_two_to_five_levels_vs_oracle(hip_device):
    """One strict FP32 engine holding a 2-, 3-, 4- and 5-level case (the
    4-level one with forced 64-bit offsets, so the depth it shares runs the
    wide instances for every case there) for two engine steps; every case
    equals its own oracle run bit for bit."""
    specs = [((8, 2), 16, False), ((16, 3), 8, False), ((32, 4), 4, True), ((64, 5), 2, False)]
    e = Engine(storage=N.LB_STORAGE_FP32, strict_fp=True)
    runs = []
    for (res, levels), n, wide in specs:
        cfg, grids, params, mesh = _case(res, levels)
        s = Solver(cfg, grids, params, mesh, strict_fp=True, wide_offsets=wide)
        s.init_equilibrium()
        for g, f in zip(grids, perturbed_state(grids)):
            s.upload(g.level_id, "f", f)
        e.add(s, n)
        runs.append((s, cfg, grids, params, mesh, n))
    print("engine mix attached", flush=True)
    e.step(1)
    e.step(1)
    e.synchronize()
    for s, cfg, grids, params, mesh, n in runs:
        assert e.progress(s) == n, f"{len(grids)}-level case: {e.progress(s)} of {n} coarse steps"
        e.remove(s)
    e.close()
    for s, cfg, grids, params, mesh, n in runs:
        o = Oracle(cfg, grids, params, "f32")
        o.init_equilibrium()
        for g, f in zip(grids, perturbed_state(grids)):
            o.put(g.level_id, "f", f.astype(np.float64))
        b = cfg.gpu_async_depth
        for t in range(1, n + 1, b):
            m = min(b, n - t + 1)
            o.step_batch(t, m, ramp_velocity(cfg.u_lattice, t + m - 1, cfg.ramp_steps))
        _assert_bitexact(compare(s, o, grids), f"engine {len(grids)}L")
        o.close()
        s.close()


@pytest.mark.gpu
def test_dense_bouzidi_q_map_strict_bitexact(hip_device):
    """The Bouzidi q map passed dense, in the form the reference holds it
    (level.bouzidi_q_map, (8,8,8,nb,27) in the caller's block order, read
    while the blocks are stored in the device's order), gives the oracle's
    results bit for bit, as the per-cell list does."""
    cfg, mesh = cases.tiny_sphere(8, 2)
    grids, params, mesh = build_case(cfg, mesh)
    assert grids[-1].n_boundary_cells > 0
    s, o = run_pair(cfg, grids, params, mesh, steps=12, strict=True, dense_q=True)
    _assert_bitexact(compare(s, o, grids), "dense q map")
