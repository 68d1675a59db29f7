"""Compressible thermal form on a wing, at a small size: the extruded
NACA 0012 at alpha 2 deg with the thermal DDF and the O(u^4) equilibrium,
2 levels at a resolution the CPU oracle runs in seconds.  The fast GPU build
against the FP64 oracle after 96 coarse steps (FP32 and FP16 storage): Cl and Cd within
0.5 % and the temperature field within 1e-4 RMS (FP16 5e-4); the strict build, with the
wall model off (its powf/logf differ from glibc's by an ulp), equals the FP32
mirror bit for bit on the thermal populations, temperature and density."""
import numpy as np
import pytest

from openludwig_amd import cases
from openludwig_amd.solver import Solver, build_case, ramp_velocity
from oracle import Oracle, oracle_forces


def _wing(**kw):
    cfg, mesh = cases.naca_wing_thermal(12, 2, n_chord=40, n_span=20)
    cfg = cfg.replace(ramp_steps=16, min_coarse_blocks=1, **kw)
    grids, params, mesh = build_case(cfg, mesh)
    return cfg, grids, params, mesh


def _run(s, cfg, n):
    for t in range(1, n + 1, 8):
        u = ramp_velocity(cfg.u_lattice, t + 7, cfg.ramp_steps)
        s.step_batch(t, 8, u) if hasattr(s, "step_batch") else s.execute_timestep_batch(t, 8, u)


@pytest.mark.gpu
def test_wing_thermal_strict_bitexact(hip_device):
    cfg, grids, params, mesh = _wing(wall_model_enabled=False)
    assert cfg.ddf_enabled and len(grids) == 2 and grids[-1].n_boundary_cells > 0
    s = Solver(cfg, grids, params, mesh, strict_fp=True)
    o = Oracle(cfg, grids, params, "f32")
    s.init_equilibrium()
    o.init_equilibrium()
    _run(s, cfg, 16)
    _run(o, cfg, 16)
    s.synchronize()
    for g in grids:
        for fld in ("g", "T", "rho"):
            a = s.download(g.level_id, fld).astype(np.float64)
            b = o.get(g.level_id, fld)
            d = np.abs(a - b)
            assert d.max() == 0.0, f"level {g.level_id} {fld}: max |diff| {d.max()}"


@pytest.mark.gpu
@pytest.mark.parametrize("storage", [0, 1])
def test_wing_thermal_fast_vs_fp64(hip_device, storage):
    """FP32 and FP16 offset storage (f and g) against the FP64 oracle.  Cd and
    Cl within 0.5 % of max(|Cd|, |Cl|) for both; temperature RMS error below
    1e-4 (FP32) and 5e-4 (FP16, whose stored deviation g - w_k carries 11
    significant bits; measured 1.8e-4)."""
    cfg, grids, params, mesh = _wing()
    n = 96
    s = Solver(cfg, grids, params, mesh, storage=storage)
    o = Oracle(cfg, grids, params, "f64")
    s.init_equilibrium()
    o.init_equilibrium()
    _run(s, cfg, n)
    _run(o, cfg, n)
    s.synchronize()
    fg = s.compute_aerodynamics()
    f64 = oracle_forces(o, cfg, params, mesh, grids[-1].dx)
    trms = max(float(np.sqrt(np.mean((s.download(g.level_id, "T").astype(np.float64) - o.get(g.level_id, "T")) ** 2)))
               for g in grids)
    print("storage", storage, "Cd", fg["Cd"], f64["Cd"], "Cl", fg["Cl"], f64["Cl"], "T rms", trms)
    scale = max(abs(f64["Cd"]), abs(f64["Cl"]))
    assert abs(f64["Cd"]) > 1e-3
    assert abs(fg["Cd"] - f64["Cd"]) <= 5e-3 * scale
    assert abs(fg["Cl"] - f64["Cl"]) <= 5e-3 * scale
    assert trms < (1e-4 if storage == 0 else 5e-4)
