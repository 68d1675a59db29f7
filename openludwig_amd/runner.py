"""Case runner: a case folder's config.yaml and STL files in, the reference's
result files out.

solve_case      solve_main (src/main.jl:26-407): load the STLs (per-file
                subdivision, merge, object names from the file names), build
                the domain, initialise, run the batched main loop with the
                ramp evaluated at each batch end, and at the diagnostic and
                output steps write convergence.csv, forces.csv and the
                surface/flow files (outputs.py); the force accumulator
                averages the diagnostic steps after the ramp.
run_all_cases   run_all_cases (src/main.jl:410-483): the case_folders list of a
                cases_to_run.yaml, one case after another, a failed case
                recorded with status "failed" and the run continuing.

Command line:
    python -m openludwig_amd.runner CASES/ball1m_Re1e3 [--steps N] [--storage fp16]
    python -m openludwig_amd.runner --cases-file cases_to_run.yaml --cases-root CASES

Deliberate differences: only the files this runner writes are removed from
an existing output folder (the reference empties the folder); the flow VTU
holds each level's current velocity (the reference picks the velocity pair
half by the coarse step's parity, which on refined levels can be the previous
substep's); the Markdown analysis summaries and the VRAM report are not
written.
"""
from __future__ import annotations

import argparse
import math
import re
import time
from pathlib import Path

import numpy as np

from . import _native as N
from . import outputs as O
from .config import CaseConfig, load_case_configuration
from .geometry import load_mesh, merge_meshes
from .solver import ForceAccumulator, Solver, build_case, ramp_velocity

STORAGE = {"fp32": N.LB_STORAGE_FP32, "fp16": N.LB_STORAGE_FP16, "bf16": N.LB_STORAGE_BF16}
_OWN_FILES = re.compile(r"^(convergence\.csv|forces\.csv|(surface|flow)_\d{6}(\.vtu|\.lbmp)|"
                        r"net_forces_\d{6}\.(csv|vtp))$")


def load_case_meshes(cfg: CaseConfig):
    """setup_multilevel_domain(stl_paths) (src/geometry/domain.jl:248-294) up
    to the rotation: each STL loaded and subdivided on its own, merged, and
    every triangle named after its file (extension dropped)."""
    meshes, names = [], []
    for p in cfg.stl_files:
        if not Path(p).is_file():
            raise FileNotFoundError(f"STL file not found: {p}")
        m = load_mesh(p, cfg.stl_scale, cfg.minimum_facet_size)
        meshes.append(m)
        names += [re.sub(r"\.[sS][tT][lL]$", "", Path(p).name)] * m.n
    if not meshes:
        raise FileNotFoundError("No STL files found")
    return merge_meshes(meshes), names


def steps_for_case(cfg: CaseConfig, params) -> int:
    """STEPS, from domain_length_covered when it is set (main.jl:84-98)."""
    steps = cfg.steps
    if cfg.domain_length_covered > 0.0:
        t_phys = cfg.domain_length_covered * params.domain_size[0] / params.u_physical
        steps = int(math.ceil(t_phys / params.time_scale))
        steps = int(math.ceil(steps / cfg.diag_freq)) * cfg.diag_freq
    if steps <= 0:
        raise ValueError("STEPS must be > 0. Set either 'steps' or 'domain_length_covered' in config.")
    return steps


def _prepare_output(out: Path) -> None:
    if out.is_dir():
        for f in out.iterdir():
            if f.is_file() and _OWN_FILES.match(f.name):
                f.unlink()
    else:
        out.mkdir(parents=True)


def _surface_files(out: Path, step: int, solver: Solver, fd: dict, unrotated, object_names) -> None:
    """save_surface_vtk on the solver's (rotated) mesh with the domain offset,
    and the .lbmp of the unrotated mesh without offset (main.jl:276-287)."""
    p = solver.params
    maps = fd["maps"]
    multi = object_names if len(set(object_names)) > 1 else None
    O.save_surface_vtk(out / f"surface_{step:06d}", solver.mesh.triangles, solver.mesh.normals, solver.mesh.areas,
                       maps[0], (maps[1], maps[2], maps[3]), offset=p.mesh_offset, object_names=multi)
    O.export_surface_pressure_bin(out / f"surface_{step:06d}.lbmp", unrotated.triangles, unrotated.normals, maps[0],
                                  (maps[1], maps[2], maps[3]), p.rho_physical, p.u_physical, object_names=object_names)


def _flow_file(out: Path, step: int, solver: Solver) -> None:
    levels = []
    for g in solver.grids:
        L = {"coords": g.coords, "dx": g.dx, "rho": solver.download(g.level_id, "rho"),
             "vel": solver.download(g.level_id, "vel").reshape(3, -1).T, "obstacle": g.obstacle.reshape(-1)}
        if solver.cfg.ddf_enabled:
            L["T"] = solver.download(g.level_id, "T")
        levels.append(L)
    O.export_flow_vtu(out / f"flow_{step:06d}", levels, solver.cfg)


def solve_case(case, out_dir: str | Path | None = None, storage: str = "fp32", steps: int | None = None,
               device: int = 0, write_flow: bool = True, log=print) -> dict:
    """Run one case folder (or a loaded CaseConfig) and return its result
    record (build_case_result's fields that this runner has)."""
    cfg = case if isinstance(case, CaseConfig) else load_case_configuration(Path(case) / "config.yaml")
    case_name = Path(cfg.case_dir).name if cfg.case_dir else "case"
    out = Path(out_dir) if out_dir is not None else Path(cfg.case_dir) / cfg.output_dir
    _prepare_output(out)
    conv_csv, force_csv = out / "convergence.csv", out / "forces.csv"
    O.write_convergence_header(conv_csv)
    if cfg.forces_enabled:
        O.write_force_csv_header(force_csv)
    t0 = time.time()
    unrotated, object_names = load_case_meshes(cfg)
    grids, params, mesh = build_case(cfg, unrotated, gpu_device=device)
    n_steps = steps if steps is not None else steps_for_case(cfg, params)
    total_cells = sum(g.n_cells for g in grids)
    log(f"[{case_name}] {len(grids)} levels, {total_cells / 1e6:.2f} M cells, {n_steps} steps, "
        f"tau {', '.join(f'{t:.6f}' for t in params.tau_levels)}")
    solver = Solver(cfg, grids, params, mesh if cfg.forces_enabled else None, device=device,
                    storage=STORAGE[storage])
    acc = ForceAccumulator()
    extremes = {k: [math.inf, -math.inf] for k in ("Cd", "Cl", "Cs")}
    fd = None
    try:
        solver.init_equilibrium()
        t_sim = time.time()
        last_diag = t_sim
        dom_x = params.domain_size[0]
        batch = cfg.gpu_async_depth
        t = 1
        while t <= n_steps:
            batch_end = min(t + batch - 1, n_steps)
            actual = batch_end - t + 1
            u_curr = ramp_velocity(cfg.u_lattice, batch_end, cfg.ramp_steps)
            solver.execute_timestep_batch(t, actual, u_curr)
            if batch_end % cfg.diag_freq < actual or batch_end == n_steps:
                diag_step = (batch_end // cfg.diag_freq) * cfg.diag_freq
                if t <= diag_step <= batch_end:
                    st = solver.stats(1)
                    now = time.time()
                    mlups = total_cells * cfg.diag_freq / ((now - last_diag) * 1e6)
                    last_diag = now
                    time_phys = float(diag_step) * params.time_scale
                    cd = cl = cs = "N/A"
                    if cfg.forces_enabled:
                        fd = solver.compute_aerodynamics()
                        if diag_step > cfg.ramp_steps:
                            acc.add(fd)
                            for k in extremes:
                                extremes[k] = [min(extremes[k][0], fd[k]), max(extremes[k][1], fd[k])]
                        cd, cl, cs = ("%.4f" % fd["Cd"], "%.4f" % fd["Cl"], "%.4f" % fd["Cs"])
                        O.append_force_csv(force_csv, diag_step, time_phys, fd, u_curr)
                    travel = 100.0 * (time_phys * params.u_physical) / dom_x if dom_x > 0 else 0.0
                    O.append_convergence(conv_csv, diag_step, time_phys, travel, u_curr, st["rho_min"], mlups,
                                         cd, cl, cs)
                    log(f"{diag_step:8d} | {time_phys:10.4f} | {travel:6.1f}% | {u_curr:.4f} | "
                        f"{st['rho_min']:.4f} | {mlups:6.1f} | {cd:>8s} | {cl:>8s} | {cs:>8s}")
            if batch_end % cfg.output_freq < actual:
                out_step = (batch_end // cfg.output_freq) * cfg.output_freq
                if t <= out_step <= batch_end:
                    _outputs(out, out_step, solver, cfg, acc, extremes, unrotated, object_names, write_flow)
            t = batch_end + 1
        last_out = (n_steps // cfg.output_freq) * cfg.output_freq
        if last_out != n_steps or n_steps < cfg.output_freq:
            fd = _outputs(out, n_steps, solver, cfg, None, None, unrotated, object_names, write_flow) or fd
        solver.synchronize()
        wall = time.time() - t_sim
    finally:
        solver.close()
    res = {"case_name": case_name, "status": "ok", "re_number": params.re_number, "u_physical": params.u_physical,
           "alpha_deg": cfg.alpha_deg, "beta_deg": cfg.beta_deg, "total_cells": total_cells,
           "wall_time": wall, "setup_time": t_sim - t0, "mlups": total_cells * n_steps / (wall * 1e6),
           "steps": n_steps, "n_samples": acc.n_samples, "ma_physical": params.u_physical / 343.0,
           "compressibility_correction": cfg.compressibility_correction, "output_dir": str(out),
           "storage": storage}
    if fd is not None:
        res.update({k: fd[k] for k in ("Cd", "Cl", "Cs", "Cmx", "Cmy", "Cmz")})
        q_inf = 0.5 * params.rho_physical * params.u_physical ** 2
        f_ref = q_inf * params.reference_area
        res["Cdp"] = fd["Fx_pressure"] / f_ref if f_ref > 1e-10 else 0.0
        res["Cdv"] = fd["Fx_viscous"] / f_ref if f_ref > 1e-10 else 0.0
    res.update(averaged_coefficients(acc, params))
    for k, (lo, hi) in extremes.items():
        res[f"min_{k}"], res[f"max_{k}"] = (lo, hi) if acc.n_samples else (0.0, 0.0)
    return res


def _outputs(out, step, solver, cfg, acc, extremes, unrotated, object_names, write_flow):
    """The output-step block of the main loop (main.jl:262-290)."""
    if write_flow:
        _flow_file(out, step, solver)
    if not cfg.forces_enabled:
        return None
    fd = solver.compute_aerodynamics(maps=True)
    if acc is not None and step != (step // cfg.diag_freq) * cfg.diag_freq and step > cfg.ramp_steps:
        acc.add(fd)
        for k in extremes:
            extremes[k] = [min(extremes[k][0], fd[k]), max(extremes[k][1], fd[k])]
    _surface_files(out, step, solver, fd, unrotated, object_names)
    if cfg.net_force_export:
        # compute_net_aerodynamics! + the two net-force exports (main.jl:289-307)
        net = solver.compute_net_aerodynamics()
        m = solver.mesh
        O.export_net_triangle_forces_csv(out / f"net_forces_{step:06d}.csv", net, unrotated.centers,
                                         alpha_deg=cfg.alpha_deg, beta_deg=cfg.beta_deg, object_names=object_names)
        O.export_net_triangle_forces_vtk(out / f"net_forces_{step:06d}", net, m.centers, m.normals, m.areas,
                                         offset=solver.params.mesh_offset)
    return fd


def averaged_coefficients(acc: ForceAccumulator, params) -> dict:
    """print_averaged_summary's averages (src/forces/surface.jl:58-120)."""
    if acc.n_samples == 0:
        return {f"avg_{k}": 0.0 for k in ("Cd", "Cl", "Cs", "Cdp", "Cdv", "Cmx", "Cmy", "Cmz")}
    a = acc.sums / acc.n_samples       # Fx_p Fy_p Fz_p Fx_v Fy_v Fz_v Mx My Mz
    q_inf = 0.5 * params.rho_physical * params.u_physical ** 2
    f_ref = q_inf * params.reference_area
    m_ref = f_ref * params.reference_chord
    fx, fy, fz = a[0] + a[3], a[1] + a[4], a[2] + a[5]
    ok, okm = f_ref > 1e-10, m_ref > 1e-10
    return {"avg_Cd": fx / f_ref if ok else 0.0, "avg_Cl": fz / f_ref if ok else 0.0,
            "avg_Cs": fy / f_ref if ok else 0.0, "avg_Cdp": a[0] / f_ref if ok else 0.0,
            "avg_Cdv": a[3] / f_ref if ok else 0.0, "avg_Cmx": a[6] / m_ref if okm else 0.0,
            "avg_Cmy": a[7] / m_ref if okm else 0.0, "avg_Cmz": a[8] / m_ref if okm else 0.0}


def run_all_cases(cases_file: str | Path, cases_root: str | Path | None = None, log=print, **kw) -> list:
    """run_all_cases (src/main.jl:410-483)."""
    import yaml
    cases_file = Path(cases_file)
    if not cases_file.is_file():
        raise FileNotFoundError("cases_to_run.yaml not found!")
    names = yaml.safe_load(cases_file.read_text())["case_folders"]
    root = Path(cases_root) if cases_root is not None else cases_file.parent / "CASES"
    results = []
    for i, name in enumerate(names, 1):
        log(f">>> CASE {i}/{len(names)}: {name}")
        try:
            folder = root / name
            if not folder.is_dir():
                raise FileNotFoundError(f"Case folder not found: {folder}")
            results.append(solve_case(folder, log=log, **kw))
        except Exception as e:  # a failed case is recorded and the loop goes on
            log(f"!!! ERROR: {e}")
            results.append({"case_name": name, "status": "failed", "error_msg": str(e)})
    return results


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("case", nargs="?", help="case folder holding config.yaml")
    ap.add_argument("--cases-file", help="cases_to_run.yaml (runs every listed case)")
    ap.add_argument("--cases-root", help="folder holding the case folders (default: CASES beside the file)")
    ap.add_argument("--out", help="output folder (default: the case's basic.simulation.output_dir)")
    ap.add_argument("--steps", type=int, help="override STEPS")
    ap.add_argument("--storage", choices=sorted(STORAGE), default="fp32")
    ap.add_argument("--no-flow", action="store_true", help="skip the flow VTU files")
    a = ap.parse_args(argv)
    kw = dict(storage=a.storage, steps=a.steps, write_flow=not a.no_flow)
    if a.cases_file:
        res = run_all_cases(a.cases_file, a.cases_root, **kw)
    elif a.case:
        res = [solve_case(a.case, out_dir=a.out, **kw)]
    else:
        ap.error("give a case folder or --cases-file")
    for r in res:
        print(r)
    return 0 if all(r.get("status") == "ok" for r in res) else 1


if __name__ == "__main__":
    raise SystemExit(main())
