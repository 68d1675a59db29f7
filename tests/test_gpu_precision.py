"""Precision form on the two bodies it is specified for: the 49,152-facet cube
and the subdivision-6 icosphere (81,920 facets), both at Re 1e5 with WALE +
WMLES, at a reduced resolution (3 levels, ~1.3 M cells) that the FP64 oracle
steps in seconds.  Each body runs 160 coarse steps (ramp 16; the flow has
passed the body) in the FP64 oracle and on the GPU (fast build) with FP32,
FP16 and BF16 population storage.

Reported per storage: Cd, Cl and the u-field error (RMS over all cells of all
levels, relative to u_in) against the FP64 oracle.  Bars: FP32 and FP16 Cd
within 0.5 % of the FP64 value (the plain form's bar) and u-field RMS below
1e-3 u_in.  BF16 keeps 8 significant bits of f - w and does not meet the
0.5 % bar (measured: cube 1.35 %, sphere 4.1 %; FP16 0.04 %, FP32 0.001 %);
it is held only to a sanity bound, Cd within 6 % and u-field RMS below 1e-2
u_in, and FP16 is the storage of the serving form.
"""
from __future__ import annotations

import numpy as np
import pytest

from openludwig_amd import _native as N
from openludwig_amd import cases
from openludwig_amd.solver import Solver, build_case, ramp_velocity
from oracle import Oracle, oracle_forces

STEPS = 160
BARS = {"fp32": (5e-3, 1e-3), "fp16": (5e-3, 1e-3), "bf16": (6e-2, 1e-2)}
STORAGE = {"fp32": N.LB_STORAGE_FP32, "fp16": N.LB_STORAGE_FP16, "bf16": N.LB_STORAGE_BF16}


def _case(body: str):
    if body == "cube_49k":
        cfg, mesh = cases.cube_re1e5(resolution=24, num_levels=3)
    else:
        cfg, mesh = cases.sphere_re1e5(resolution=24, num_levels=3)
    cfg = cfg.replace(domain_upstream=1.5, domain_downstream=3.0, domain_lateral=1.5, domain_height=1.5,
                      min_coarse_blocks=1, wake_length=1.0, ramp_steps=16)
    return cfg, mesh


def _run(stepper, cfg):
    for t in range(1, STEPS + 1, 8):
        stepper(t, 8, ramp_velocity(cfg.u_lattice, t + 7, cfg.ramp_steps))


@pytest.mark.gpu
@pytest.mark.parametrize("body", ["cube_49k", "sphere_82k"])
def test_precision_form_vs_fp64_oracle(hip_device, body):
    cfg, mesh = _case(body)
    grids, params, mesh = build_case(cfg, mesh)
    o64 = Oracle(cfg, grids, params, "f64")
    o64.init_equilibrium()
    _run(o64.step_batch, cfg)
    f64 = oracle_forces(o64, cfg, params, mesh, grids[-1].dx)
    u64 = [o64.get(g.level_id, "vel") for g in grids]
    o64.close()
    u_in = cfg.u_lattice
    rows = []
    for name, st in STORAGE.items():
        s = Solver(cfg, grids, params, mesh, strict_fp=False, storage=st)
        try:
            s.init_equilibrium()
            _run(s.execute_timestep_batch, cfg)
            s.synchronize()
            fg = s.compute_aerodynamics()
            d2 = sum(float(np.sum((s.download(g.level_id, "vel").astype(np.float64) - b) ** 2))
                     for g, b in zip(grids, u64))
        finally:
            s.close()
        n = sum(b.size for b in u64)
        rms = np.sqrt(d2 / n) / u_in
        dcd = abs(fg["Cd"] - f64["Cd"]) / abs(f64["Cd"])
        rows.append((name, fg["Cd"], fg["Cl"], dcd, rms))
    print(f"\n{body}: {sum(g.n_cells for g in grids)} cells, {mesh.n} facets, {STEPS} coarse steps; "
          f"FP64 oracle Cd {f64['Cd']:.6f} Cl {f64['Cl']:.6f}")
    for name, cd, cl, dcd, rms in rows:
        print(f"  {name}: Cd {cd:.6f} ({dcd * 100:.3f} %)  Cl {cl:.6f}  u-field RMS {rms:.2e} u_in")
    assert abs(f64["Cd"]) > 0.05
    for name, cd, cl, dcd, rms in rows:
        bar_cd, bar_u = BARS[name]
        assert dcd <= bar_cd, f"{name}: Cd {cd} vs FP64 {f64['Cd']} ({dcd * 100:.3f} %)"
        assert abs(cl - f64["Cl"]) <= bar_cd * max(abs(f64["Cd"]), abs(f64["Cl"])), f"{name}: Cl {cl}"
        assert rms <= bar_u, f"{name}: u-field RMS {rms}"
