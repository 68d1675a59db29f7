"""Batched engine (lb_engine_*): a case stepped through the engine equals the
same case stepped by lb_step_batch, bit for bit, and a case's results are
the same alone and in a mix of cases with other level counts, sizes and step
counts (with and without the captured graph, under each work-group
schedule, the persistent stream-collide kernel included)."""
import numpy as np
import pytest

from openludwig_amd import _native as N
from openludwig_amd import cases
from openludwig_amd.engine import Engine
from openludwig_amd.solver import Solver, build_case


def _case(res, levels, **kw):
    cfg, mesh = cases.tiny_sphere(res, levels, **kw)
    cfg = cfg.replace(ramp_steps=12)
    grids, params, mesh = build_case(cfg, mesh)
    return cfg, grids, params, mesh


def _fields(s, grids):
    return {(g.level_id, f): s.download(g.level_id, f) for g in grids for f in ("f", "rho", "vel")}


def _assert_same(a, b, what):
    for k in a:
        d = np.abs(a[k].astype(np.float64) - b[k].astype(np.float64))
        assert d.max() == 0.0, f"{what} {k}: max |diff| {d.max()} over {np.count_nonzero(d)}"


@pytest.mark.gpu
@pytest.mark.parametrize("storage,sched", [(N.LB_STORAGE_FP32, N.LB_SCHED_LINEAR),
                                           (N.LB_STORAGE_FP16, N.LB_SCHED_LINEAR),
                                           (N.LB_STORAGE_FP32, N.LB_SCHED_PERSISTENT),
                                           (N.LB_STORAGE_FP16, N.LB_SCHED_PERSISTENT)])
def test_engine_matches_step_batch(hip_device, storage, sched):
    cfg, grids, params, mesh = _case(8, 2)
    n = 20  # two full batches of 8 and a partial one
    ref = Solver(cfg, grids, params, mesh, storage=storage)
    ref.init_equilibrium()
    ref.run(n)
    a = _fields(ref, grids)
    fa = ref.compute_aerodynamics()
    s = Solver(cfg, grids, params, mesh, storage=storage)
    s.init_equilibrium()
    e = Engine(storage=storage)
    e.set_options(graph=True, schedule=sched, waves=5)
    e.add(s, n)
    e.run_to_completion()
    assert e.progress(s) == n
    e.remove(s)
    _assert_same(a, _fields(s, grids), "engine vs step_batch")
    assert s.compute_aerodynamics()["Cd"] == fa["Cd"]


@pytest.mark.gpu
@pytest.mark.parametrize("graph,sched,waves", [(True, N.LB_SCHED_XCD, 5), (False, N.LB_SCHED_LINEAR, 6),
                                               (True, N.LB_SCHED_PERSISTENT, 5)])
def test_engine_alone_vs_mixed(hip_device, graph, sched, waves):
    specs = [((8, 2), 24), ((16, 3), 12), ((8, 3), 16)]
    built = [(_case(*rl), n) for rl, n in specs]
    alone = []
    for (cfg, grids, params, mesh), n in built:
        s = Solver(cfg, grids, params, mesh, storage=N.LB_STORAGE_FP16)
        s.init_equilibrium()
        e = Engine(storage=N.LB_STORAGE_FP16)
        e.set_options(graph=graph, schedule=sched, waves=waves)
        e.add(s, n)
        e.run_to_completion()
        e.remove(s)
        alone.append(_fields(s, grids))
        s.close()
        e.close()
    e = Engine(storage=N.LB_STORAGE_FP16)
    e.set_options(graph=graph, schedule=sched, waves=waves)
    sol = []
    for (cfg, grids, params, mesh), n in built:
        s = Solver(cfg, grids, params, mesh, storage=N.LB_STORAGE_FP16)
        s.init_equilibrium()
        sol.append(s)
    # staggered admission: the third case joins after two engine steps
    e.add(sol[0], built[0][1])
    e.add(sol[1], built[1][1])
    e.step(2)
    e.add(sol[2], built[2][1])
    e.run_to_completion()
    for i, s in enumerate(sol):
        assert e.progress(s) == built[i][1]
        e.remove(s)
        _assert_same(alone[i], _fields(s, built[i][0][1]), f"case {i} mixed vs alone")


@pytest.mark.gpu
@pytest.mark.parametrize("storage", [N.LB_STORAGE_FP32, N.LB_STORAGE_FP16])
def test_footprint_is_exact(hip_device, storage):
    from openludwig_amd.engine import ramp_table
    from openludwig_amd.solver import footprint
    cfg, grids, params, mesh = _case(16, 3)
    s = Solver(cfg, grids, params, mesh, storage=storage)
    assert footprint(cfg, grids, params, mesh, storage) == s.device_bytes
    s.init_equilibrium()
    e = Engine(storage=storage)
    e.add(s, 16)
    assert footprint(cfg, grids, params, mesh, storage, len(ramp_table(cfg, 16))) == s.device_bytes


@pytest.mark.gpu
def test_autotune_keeps_results(hip_device):
    """The autotuner's trial steps are real steps under different launch
    options; the case ends bit-identical to an untuned run."""
    cfg, grids, params, mesh = _case(16, 3)
    n = 64
    ref = Solver(cfg, grids, params, mesh, storage=N.LB_STORAGE_FP16)
    ref.init_equilibrium()
    ref.run(n)
    a = _fields(ref, grids)
    s = Solver(cfg, grids, params, mesh, storage=N.LB_STORAGE_FP16)
    s.init_equilibrium()
    e = Engine(storage=N.LB_STORAGE_FP16)
    e.add(s, n)
    res = e.autotune(trial_steps=1)
    assert len(res["trials"]) >= 3 and res["options"] in [{k: t[k] for k in ("graph", "schedule", "waves")}
                                                          for t in res["trials"]]
    e.run_to_completion()
    e.remove(s)
    _assert_same(a, _fields(s, grids), "autotuned engine vs step_batch")
