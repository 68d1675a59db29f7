"""The case runner end to end on the GPU: a case folder (config.yaml + the
reference's cube STL fixture) in, convergence.csv / forces.csv / surface and
flow files out, with the same coefficients as Solver.run on the same case, and
run_all_cases recording a missing folder as failed while running the rest."""
from __future__ import annotations

import shutil
from pathlib import Path

import numpy as np
import pytest

from openludwig_amd import outputs as O
from openludwig_amd.config import load_case_configuration
from openludwig_amd.runner import load_case_meshes, run_all_cases, solve_case
from openludwig_amd.solver import Solver, build_case

GOLDEN = Path(__file__).parent / "golden"

CONFIG = """\
basic:
  stl_file: cube1m.stl
  stl_scale: 1.0
  surface_resolution: 24
  num_levels: 2
  reference_area_of_full_model: 1.0
  reference_chord: 1.0
  reference_length_for_meshing: 1.0
  fluid:
    density: 1.225
    kinematic_viscosity: 1.5e-5
  flow:
    velocity: 0.015
  simulation:
    steps: 48
    ramp_steps: 16
    output_freq: 32
    output_dir: RESULTS
advanced:
  numerics:
    u_lattice: 0.05
    inlet_turbulence_intensity: 0.0
    cumulant:
      limiter: positivity
  domain:
    upstream: 1.0
    downstream: 2.0
    lateral: 1.0
    height: 1.0
  diagnostics:
    freq: 16
  forces:
    net_force_export: true
"""


def _case(tmp_path: Path, name: str = "cube_small") -> Path:
    d = tmp_path / name
    d.mkdir()
    (d / "config.yaml").write_text(CONFIG)
    shutil.copy(GOLDEN / "cube1m.stl", d / "cube1m.stl")
    return d


@pytest.mark.gpu
def test_runner_outputs_and_coefficients(hip_device, tmp_path):
    case = _case(tmp_path)
    res = solve_case(case, log=lambda *a: None)
    out = case / "RESULTS"
    assert res["status"] == "ok" and res["steps"] == 48
    head, rows = O.read_csv(out / "convergence.csv")
    assert head == O.CONVERGENCE_HEADER.split(",")
    assert [r[0] for r in rows] == ["16", "32", "48"]
    head, frows = O.read_csv(out / "forces.csv")
    assert head == O.FORCES_HEADER.split(",") and [r[0] for r in frows] == ["16", "32", "48"]
    # output step 32 and the final step 48 (48 is not a multiple of 32)
    for step in (32, 48):
        assert (out / f"flow_{step:06d}.vtu").is_file()
        s = O.read_vtu(out / f"surface_{step:06d}.vtu")
        b = O.read_surface_pressure_bin(out / f"surface_{step:06d}.lbmp")
        assert s["n_cells"] == 12 and b["pressure"].shape == (12,)
        assert np.array_equal(s["cell_data"]["Pressure"], b["pressure"])
        assert b["names"] == ["cube1m"]
    # net per-triangle forces (closed body): their sum is the total force
    _, nrows = O.read_csv(out / "net_forces_000048.csv")
    assert len(nrows) == 12 and (out / "net_forces_000048.vtp").is_file()
    fx = sum(float(r[5]) for r in nrows)
    fx_total = float(frows[-1][3])
    assert abs(fx - fx_total) <= 1e-4 * sum(abs(float(r[5])) for r in nrows) + 1e-9
    f = O.read_vtu(out / "flow_000048.vtu")
    assert {"Density", "Velocity", "VelocityMagnitude", "Obstacle", "Level"} <= set(f["cell_data"])
    assert np.all(np.isfinite(f["cell_data"]["Density"]))

    # the same case through Solver.run: identical coefficients at every diag step
    cfg = load_case_configuration(case / "config.yaml")
    mesh, _ = load_case_meshes(cfg)
    grids, params, mesh = build_case(cfg, mesh)
    s = Solver(cfg, grids, params, mesh)
    try:
        s.init_equilibrium()
        run = s.run(48)
        final = s.compute_aerodynamics()
    finally:
        s.close()
    assert [h["step"] for h in run["history"]] == [16, 32, 48]
    for h, r, c in zip(run["history"], frows, rows):
        assert "%.6f" % h["Cd"] == r[15] and "%.6f" % h["Cl"] == r[16]
        assert "%.4f" % h["Cd"] == c[6] and "%.4f" % h["Cl"] == c[7]
    assert res["Cd"] == final["Cd"] and res["Cl"] == final["Cl"]
    assert res["n_samples"] == run["n_samples"] == 2            # steps 32 and 48 are past the ramp
    assert abs(res["avg_Cd"] - run["Cd_avg"]) <= 1e-12 * max(1.0, abs(run["Cd_avg"]))


@pytest.mark.gpu
def test_run_all_cases_records_failures(hip_device, tmp_path):
    (tmp_path / "CASES").mkdir()
    _case(tmp_path / "CASES", "a")
    (tmp_path / "cases_to_run.yaml").write_text("case_folders:\n  - a\n  - missing_case\n")
    res = run_all_cases(tmp_path / "cases_to_run.yaml", steps=16, write_flow=False, log=lambda *a: None)
    assert [r["status"] for r in res] == ["ok", "failed"]
    assert "Case folder not found" in res[1]["error_msg"]
    assert (tmp_path / "CASES" / "a" / "RESULTS" / "surface_000016.lbmp").is_file()
