"""Writes sphere_re1e3_fp64_2000steps.json: the FP64 oracle's force history of the
plain CPU-runnable sphere case (D = 1 m icosphere subdiv 4, Re 1e3, coarse
64x32x32, 3 levels, 2000 coarse steps, ramp 500, batches of 8, diagnostics
every 100 steps), stepped with the main loop's cadence (main.jl:208-260: the
inlet velocity of each batch evaluated at its last step, forces at the batch
that contains a diagnostic step, the average over diagnostic steps after the
ramp).  The GPU test compares the library's run of the same case with it.

    python tests/golden/make_sphere_re1e3_fp64.py [--steps N]      (about 30 min on 8 cores)
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

from openludwig_amd import cases  # noqa: E402
from openludwig_amd.solver import build_case, ramp_velocity  # noqa: E402
from oracle import Oracle, oracle_forces  # noqa: E402


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--out", default=str(Path(__file__).with_name("sphere_re1e3_fp64_2000steps.json")))
    a = ap.parse_args()
    cfg, mesh = cases.sphere_re1e3()
    grids, params, mesh = build_case(cfg, mesh)
    o = Oracle(cfg, grids, params, "f64")
    o.init_equilibrium()
    hist, post = [], []
    t, t0 = 1, time.time()
    while t <= a.steps:
        be = min(t + cfg.gpu_async_depth - 1, a.steps)
        u = ramp_velocity(cfg.u_lattice, be, cfg.ramp_steps)
        o.step_batch(t, be - t + 1, u)
        if be % cfg.diag_freq < be - t + 1 or be == a.steps:
            ds = (be // cfg.diag_freq) * cfg.diag_freq
            if t <= ds <= be:
                f = oracle_forces(o, cfg, params, mesh, grids[-1].dx)
                hist.append({"step": ds, "batch_end": be, "Cd": f["Cd"], "Cl": f["Cl"], "Cs": f["Cs"]})
                if ds > cfg.ramp_steps:
                    post.append(f)
                print(f"step {ds} Cd {f['Cd']:.6f} Cl {f['Cl']:.6f} ({time.time() - t0:.0f} s)", flush=True)
        t = be + 1
    o.close()
    q_inf = 0.5 * params.rho_physical * params.u_physical ** 2
    f_ref = q_inf * params.reference_area
    avg = {k: sum(f[k] for f in post) / len(post) for k in ("Fx", "Fy", "Fz")} if post else {}
    out = {"case": "sphere_re1e3 (icosphere subdiv 4, Re 1e3, 3 levels, coarse 64x32x32)",
           "precision": "f64", "steps": a.steps, "ramp_steps": cfg.ramp_steps, "diag_freq": cfg.diag_freq,
           "batch": cfg.gpu_async_depth, "cells_per_level": [g.n_cells for g in grids], "history": hist,
           "n_samples": len(post),
           "Cd_avg": avg["Fx"] / f_ref if post else None, "Cl_avg": avg["Fz"] / f_ref if post else None,
           "seconds": round(time.time() - t0, 1)}
    Path(a.out).write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", a.out)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
