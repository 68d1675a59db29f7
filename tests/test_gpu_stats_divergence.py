"""FP64 diagnostics and divergence handling.

  * lb_stats (compute_flow_stats, diagnostics.jl:43-93, plus the global mass
    check) against numpy's FP64 sums over the oracle's fields after the same
    steps (the strict build equals them bit for bit): sum rho, kinetic
    energy, rho min/max, v_max over fluid cells, fluid count, the population
    sum, and the count of non-finite populations after a NaN is planted.
  * lb_check_finite: silent on a healthy case, LB_ENONFINITE once a planted
    NaN population has reached the density and velocity (Solver.check_finite
    raises DivergenceError; Solver.run checks at every diagnostic step).  A
    NaN in a cell's rest population needs two steps: in the first its own
    density is clamped by max(rho, 0.01) (fmaxf keeps the number; see the
    note on non-finite inputs in the parity notes) and only its populations
    turn NaN; in the second the neighbours pull them.
  * lb_engine_check_finite: in an engine holding three cases, the one with a
    planted NaN is reported and removed; the other two go on and end
    bit-identical to their runs alone.
"""
import numpy as np
import pytest

from openludwig_amd import _native as N
from openludwig_amd import cases
from openludwig_amd.engine import Engine
from openludwig_amd.solver import DivergenceError, Solver, build_case
from parity_helpers import run_pair


def _stats_numpy(f, rho, vel, obst):
    fluid = ~obst.ravel().astype(bool)        # (n_b, 512) -> the reference's cell order
    r = rho[fluid]
    v = vel.reshape(3, -1)[:, fluid]
    u2 = (v * v).sum(axis=0)
    return {"rho_sum": r.sum(), "kinetic_energy": (0.5 * r * u2).sum(), "rho_min": r.min(), "rho_max": r.max(),
            "v_max": np.sqrt(u2).max(), "fluid_cells": int(fluid.sum()), "mass": f.sum(),
            "nonfinite": int(np.count_nonzero(~np.isfinite(f)))}


def _fluid_slot(g, i):
    """Index (in the reference layout of f) of the rest population of the
    i-th fluid cell of level g."""
    return int(np.flatnonzero(~g.obstacle.astype(bool))[i]) + 13 * g.n_cells


@pytest.mark.gpu
def test_stats_match_numpy_over_oracle_fields(hip_device):
    cfg, mesh = cases.tiny_sphere(8, 2)
    grids, params, mesh = build_case(cfg, mesh)
    s, o = run_pair(cfg, grids, params, mesh, steps=8, strict=True)
    for g in grids:
        st = s.stats(g.level_id)
        f = o.get(g.level_id, "f")
        rho = o.get(g.level_id, "rho")
        vel = o.get(g.level_id, "vel")
        ref = _stats_numpy(f, rho, vel, g.obstacle)
        print(g.level_id, st, ref)
        assert st["fluid_cells"] == ref["fluid_cells"] and st["nonfinite"] == 0
        for k in ("rho_min", "rho_max", "v_max"):
            assert st[k] == ref[k], k
        for k in ("rho_sum", "kinetic_energy", "mass"):
            assert abs(st[k] - ref[k]) <= 1e-12 * abs(ref[k]), (k, st[k], ref[k])
    # a planted NaN population is counted; the density and velocity are not yet touched
    g = grids[-1]
    f = s.download(g.level_id, "f")
    f[12345] = np.nan
    s.upload(g.level_id, "f", f)
    assert s.stats(g.level_id)["nonfinite"] == 1
    s.check_finite()


@pytest.mark.gpu
def test_check_finite_raises_after_a_step(hip_device):
    cfg, mesh = cases.tiny_sphere(8, 2)
    grids, params, mesh = build_case(cfg, mesh)
    s = Solver(cfg, grids, params, mesh)
    s.init_equilibrium()
    s.execute_timestep_batch(1, 2, 0.01)
    s.check_finite()
    g = grids[0]
    f = s.download(g.level_id, "f")
    f[_fluid_slot(g, 777)] = np.nan
    s.upload(g.level_id, "f", f)
    s.execute_timestep_batch(3, 2, 0.01)
    with pytest.raises(DivergenceError) as ei:
        s.check_finite()
    assert ei.value.count > 0
    # the main loop stops at the next diagnostic step
    s2 = Solver(cfg, grids, params, mesh)
    s2.init_equilibrium()
    s2.upload(g.level_id, "f", f)
    with pytest.raises(DivergenceError):
        s2.run(16, diag_freq=8)


def _fields(s, grids):
    return {(g.level_id, f): s.download(g.level_id, f) for g in grids for f in ("f", "rho", "vel")}


@pytest.mark.gpu
def test_engine_evicts_diverged_case(hip_device):
    specs = [(8, 2), (16, 3), (8, 3)]
    built = []
    for res, lv in specs:
        cfg, mesh = cases.tiny_sphere(res, lv)
        cfg = cfg.replace(ramp_steps=12)
        built.append((cfg,) + build_case(cfg, mesh))
    n = 16
    alone = []
    for i in (0, 2):
        cfg, grids, params, mesh = built[i]
        s = Solver(cfg, grids, params, mesh, storage=N.LB_STORAGE_FP16)
        s.init_equilibrium()
        e = Engine(storage=N.LB_STORAGE_FP16)
        e.add(s, n)
        e.run_to_completion()
        e.remove(s)
        alone.append(_fields(s, grids))
        s.close()
        e.close()
    e = Engine(storage=N.LB_STORAGE_FP16)
    sol = []
    for cfg, grids, params, mesh in built:
        s = Solver(cfg, grids, params, mesh, storage=N.LB_STORAGE_FP16)
        s.init_equilibrium()
        sol.append(s)
    g = built[1][1][0]
    f = sol[1].download(g.level_id, "f")
    f[_fluid_slot(g, 4321)] = np.nan
    sol[1].upload(g.level_id, "f", f)
    for s in sol:
        e.add(s, n)
    assert e.check_finite() == []  # nothing has been stepped yet
    e.step(2)
    bad = e.check_finite()
    assert bad == [sol[1]]
    e.remove(sol[1])
    e.run_to_completion()
    assert e.check_finite() == []
    for k, i in enumerate((0, 2)):
        assert e.progress(sol[i]) == n
        e.remove(sol[i])
        got = _fields(sol[i], built[i][1])
        for key in got:
            assert np.array_equal(got[key], alone[k][key]), f"case {i} {key} differs from its run alone"
