#!/usr/bin/env python3
"""Benchmark of the lattice-Boltzmann time step on one MI355X.

Workload: the plain FP32 drop-in case -- a 1 m cube (12-facet STL refined to
49,152 triangles) at Re 1e5, cumulant + WALE + WMLES, Bouzidi links on the
finest level, 4 levels of 2:1 refinement, 52.5 M active cells in 8^3 blocks,
inlet u = 0.05 (lattice).  Synthetic: the geometry is generated, the flow
starts from rest as the reference's init_eq! does.

--workload naca-thermal measures the compressible thermal form instead: an
extruded NACA 0012 wing at alpha 2 deg with the thermal DDF, 29.4 M cells.
--workload sphere is the small CPU-runnable case: a 1 m sphere at Re 1e3,
coarse 64x32x32, 3 levels, 1.34 M cells, 2,000 coarse steps, no wall model
(249 B per cell update in FP32); with --cpu-steps -1 the C oracle runs the
whole case on the host as well, force samples included.

A "step" is one coarse time step: the full multilevel recursion (15
substeps over 4 levels), issued in batches of 8 like execute_timestep_batch!.
The run is the workload's stated length (cube: 2,500 coarse steps = 20,000
finest steps) with an inflow ramp of --ramp coarse steps (500): a spin-up
brings the flow to coarse step total - warmup - steps, then the warm-up
steps and the timed steps follow, so the timed window is the developed flow
at the end of the run.  Forces are sampled every --diag coarse steps after
the ramp, as the main loop does (main.jl:236-243), outside the timed window;
Cd/Cl are their averages (accumulate_forces!, surface.jl:36-56).
value = MLUPS_true (every substep's cell updates counted) over the timed
steps, with the populations resident in HBM.  roofline = the finest level's
stream-collide kernel (the dominant one): algorithmic bytes per launch
(253 B per cell update: 2x27x4 B populations + 33 B macroscopic/flags + 4 B
wall distance; 145 B with --storage fp16/bf16, populations as 16-bit offsets
f - w; on the finest level the 4 B density store happens on one substep of
the 8 per coarse step, so its launches average 249.5 and 141.5 B) / its
average launch time from HIP events on the library's stream.
cpu_baseline = the C oracle (the restated reference algorithm) on the host
cores over --cpu-steps coarse steps of the same workload (default one; -1:
the whole run with the force samples, whose averaged Cd/Cl it reports).  serving = the serving
comparison of openludwig_amd.serving.compare (FP16 continuous batching vs
plain FP32 sequential) on the serving request set: 64 requests of 1-40 M
cells, 2-5 levels.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# bytes per cell update with the wall model: 2 x 27 populations + 33 B of
# macroscopic fields/flags + 4 B wall distance
B_ALG_WMLES = {"fp32": 2 * 27 * 4 + 37, "fp16": 2 * 27 * 2 + 37, "bf16": 2 * 27 * 2 + 37}


def leaf_bytes(b: float, n_levels: int) -> float:
    """Average bytes per cell update of a finest-level launch: the 4-byte
    density (or temperature) store happens on the last of the level's
    2^(n_levels - 1) substeps per coarse step only."""
    return b - 4.0 * (1.0 - 1.0 / 2 ** (n_levels - 1))


STORAGE_CODE = {"fp32": 0, "fp16": 1, "bf16": 2}
HBM_PEAK_GBS = 8000.0       # MI355X HBM3E


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(cfg, grids, params, mesh, n_steps: int, diag: int) -> dict:
    """The CPU baseline: the C oracle (FP32 mirror of the restated reference
    algorithm, OpenMP over blocks) on the host cores over coarse steps
    1..n_steps of the case, with the GPU run's cadence -- batches of
    gpu_async_depth at the inlet velocity of each batch end, force samples
    every `diag` coarse steps after the ramp, whose time is left out of the
    rate as the GPU's timed window holds none.  Returns the cpu_baseline
    object of the JSON line (with the averaged Cd/Cl when samples were taken)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    from oracle import Oracle, oracle_forces

    from openludwig_amd.solver import ForceAccumulator, averaged_coefficients, mlups, ramp_velocity
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    batch = cfg.gpu_async_depth
    n_cells = [g.n_cells for g in grids]
    o = Oracle(cfg, grids, params, "f32")
    o.init_equilibrium()
    acc = ForceAccumulator()
    tcpu = 0.0
    t = 1
    while t <= n_steps:
        be = min(t + batch - 1, n_steps)
        tc = time.perf_counter()
        o.step_batch(t, be - t + 1, ramp_velocity(cfg.u_lattice, be, cfg.ramp_steps))
        tcpu += time.perf_counter() - tc
        d = (be // diag) * diag
        if t <= d <= be and d > cfg.ramp_steps:
            acc.add(oracle_forces(o, cfg, params, mesh, grids[-1].dx))
        if n_steps > batch and t <= d <= be:
            log(f"cpu baseline: coarse step {be} of {n_steps}, {tcpu:.1f}s")
        t = be + 1
    o.close()
    c_true, _ = mlups(grids, n_steps, tcpu)
    cpu = {"value": round(c_true, 2), "unit": "MLUPS", "cores": threads, "kind": "port",
           "sample": f"{n_steps} coarse step(s) of the same {sum(n_cells) / 1e6:.1f}M-cell, {len(grids)}-level "
                     f"workload ({sum(c * 2 ** i for i, c in enumerate(n_cells)) * n_steps / 1e6:.0f}M cell "
                     f"updates), C oracle (FP32 mirror, OpenMP)", "seconds": round(tcpu, 2),
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
    if acc.n_samples:
        cc = averaged_coefficients(params, acc.sums, acc.n_samples)
        cpu.update({"Cd_avg": float(f"{cc['Cd_avg']:.6g}"), "Cl_avg": float(f"{cc['Cl_avg']:.6g}"),
                    "force_samples": acc.n_samples})
    log(f"cpu baseline {c_true:.1f} MLUPS on {threads} threads ({tcpu:.1f}s)")
    return cpu


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--workload", default="cube", choices=["cube", "sphere", "naca-thermal"],
                    help="cube: plain drop-in case (52.5M cells); sphere: the small CPU-runnable case "
                         "(Re 1e3, 1.34M cells); naca-thermal: compressible thermal wing (29.4M cells)")
    ap.add_argument("--resolution", type=int, default=0,
                    help="0: the workload's own (cube 160, sphere 16, wing 288)")
    ap.add_argument("--strict", action="store_true", help="use the no-FMA build")
    ap.add_argument("--storage", default="fp32", choices=sorted(B_ALG_WMLES),
                    help="population storage (fp16/bf16: offsets f - w, FP32 arithmetic)")
    ap.add_argument("--cpu-steps", type=int, default=1,
                    help="coarse steps timed for the CPU baseline (0: skip; -1: the whole run, with the force "
                         "samples)")
    ap.add_argument("--probe-gib", type=float, default=2.0,
                    help="buffer size of the HBM bandwidth probe (0: skip)")
    ap.add_argument("--serving-requests", type=int, default=64,
                    help="requests in the serving comparison, FP16 batched vs plain FP32 (0: skip)")
    ap.add_argument("--serving-cells", type=float, nargs=2, default=(1e6, 40e6),
                    help="active-cell range of the serving requests")
    ap.add_argument("--total", type=int, default=-1,
                    help="coarse steps of the whole run (-1: the workload's stated length; the spin-up is "
                         "total - warmup - steps, 0 or less: none)")
    ap.add_argument("--ramp", type=int, default=500, help="inflow ramp, coarse steps")
    ap.add_argument("--diag", type=int, default=100, help="force sampling cadence after the ramp, coarse steps")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU form of this "
              f"workload is a design (block-partitioned levels with halo exchange), not a measured path.",
              file=sys.stderr)
        return 2

    from openludwig_amd import cases
    from openludwig_amd.solver import (DivergenceError, ForceAccumulator, Solver, averaged_coefficients, build_case,
                                       mlups, ramp_velocity)

    t0 = time.time()
    if args.workload == "cube":
        args.resolution = args.resolution or 160
        cfg, mesh = cases.cube_re1e5(args.resolution)
        wname = "cube Re 1e5, 49,152-tri cube, cumulant+WALE+WMLES, Bouzidi finest level"
    elif args.workload == "sphere":
        args.resolution = args.resolution or 16
        cfg, mesh = cases.sphere_re1e3(args.resolution)
        wname = "sphere D 1 m Re 1e3, 5,120-tri icosphere, cumulant+WALE, Bouzidi finest level"
    else:
        args.resolution = args.resolution or 288
        cfg, mesh = cases.naca_wing_thermal(args.resolution)
        wname = ("NACA 0012 wing (chord 1, span 2, 40,400 tris, alpha 2), cumulant+WALE+WMLES, thermal DDF + "
                 "O(u^4) equilibrium, u_lat 0.15, Bouzidi finest level")
    grids, params, mesh = build_case(cfg, mesh)
    n_cells = [g.n_cells for g in grids]
    log(f"setup {time.time() - t0:.1f}s: levels {len(grids)}, cells {[f'{c / 1e6:.2f}M' for c in n_cells]}, "
        f"bouzidi cells {grids[-1].n_boundary_cells}, tau {params.tau_levels}")

    total = cfg.steps if args.total < 0 else args.total
    total = max(total, args.warmup + args.steps)
    spinup = total - args.warmup - args.steps
    cfg = cfg.replace(ramp_steps=args.ramp, steps=total)
    solver = Solver(cfg, grids, params, mesh, strict_fp=args.strict, storage=STORAGE_CODE[args.storage])
    # without the wall model the kernel reads no wall distance (4 B less)
    b_alg = B_ALG_WMLES[args.storage] - (0 if cfg.wall_model_enabled else 4)
    solver.init_equilibrium()
    mass0 = solver.mass()
    batch = cfg.gpu_async_depth
    acc = ForceAccumulator()

    def run(t_first: int, n: int, sample: bool) -> None:
        """Coarse steps t_first .. t_first+n-1 in batches of gpu_async_depth
        (inlet velocity at each batch end, main.jl:214-218); with `sample`,
        the forces after every batch that holds a multiple of the sampling
        cadence past the ramp (main.jl:226-243) and the divergence check."""
        t = t_first
        end = t_first + n - 1
        last = time.perf_counter()
        while t <= end:
            be = min(t + batch - 1, end)
            solver.execute_timestep_batch(t, be - t + 1, ramp_velocity(cfg.u_lattice, be, cfg.ramp_steps))
            d = (be // args.diag) * args.diag
            if sample and t <= d <= be and d > cfg.ramp_steps:
                solver.check_finite()
                acc.add(solver.compute_aerodynamics())
            if sample and time.perf_counter() - last > 30.0:
                solver.synchronize()
                last = time.perf_counter()
                log(f"spin-up: coarse step {be} of {end}")
            t = be + 1

    def sync() -> None:
        solver.synchronize()

    ts = time.perf_counter()
    run(1, spinup + args.warmup, True)
    sync()
    t_spin = time.perf_counter() - ts
    log(f"spin-up and warm-up: {spinup + args.warmup} coarse steps in {t_spin:.1f}s, {acc.n_samples} force samples")
    finest = grids[-1].level_id
    solver.profile(True)
    for g in grids:
        solver.profile_read(g.level_id)
        solver.profile_read(g.level_id, 1)
    sync()
    t_start = time.perf_counter()
    run(1 + spinup + args.warmup, args.steps, False)
    sync()
    elapsed = time.perf_counter() - t_start
    kern = {g.level_id: solver.profile_read(g.level_id) for g in grids}
    kern_t = {g.level_id: solver.profile_read(g.level_id, 1) for g in grids}
    solver.profile(False)
    try:
        solver.check_finite()
    except DivergenceError as exc:
        log(str(exc))
        return 3
    t_end = spinup + args.warmup + args.steps
    # the sample the timed window holds (its last step is a multiple of the cadence by construction)
    d_last = (t_end // args.diag) * args.diag
    if d_last > spinup + args.warmup and d_last > cfg.ramp_steps:
        acc.add(solver.compute_aerodynamics())
    forces = solver.compute_aerodynamics()
    coeffs = averaged_coefficients(params, acc.sums, acc.n_samples)
    lstats = [solver.stats(g.level_id) for g in grids]
    mass1 = solver.mass()
    diagnostics = {"rho_min": round(min(s_["rho_min"] for s_ in lstats), 6),
                   "rho_max": round(max(s_["rho_max"] for s_ in lstats), 6),
                   "v_max": round(max(s_["v_max"] for s_ in lstats), 6),
                   "mass_rel_change": float(f"{(mass1 - mass0) / mass0:.4g}")}

    # HBM bytes per launch of the same kernel, from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE passes (scaled by a copy of known size)
    traffic = None
    prof = ROOT / "profiles" / ("collide_cube_re1e5_52M.json" if args.storage == "fp32"
                                else f"collide_cube_re1e5_52M_{args.storage}.json")
    if prof.exists() and not args.strict and args.workload == "cube" and args.resolution == 160:
        pj = json.loads(prof.read_text())
        if pj.get("cells") == grids[-1].n_cells:
            traffic = round(pj["read_bytes_per_launch"] + pj["write_bytes_per_launch"])

    m_true, m_ref = mlups(grids, args.steps, elapsed)
    thermal = None
    if cfg.ddf_enabled:
        # thermal DDF kernel: 2x27 populations (4 or 2 B) + u 12 + T 4 + obstacle 1 + sponge 4
        tms, tn = kern_t[finest]
        t_avg = tms / max(tn, 1)
        tb = leaf_bytes(2 * 27 * (4 if args.storage == "fp32" else 2) + 21, len(grids))
        thermal = {"bytes_per_lu": tb, "finest_avg_launch_ms": round(t_avg, 4),
                   "finest_GBps": round(grids[-1].n_cells * tb / (t_avg * 1e-3) / 1e9, 1)}
    ms_k, n_k = kern[finest]
    avg_ms = ms_k / max(n_k, 1)
    b_lev = {g.level_id: (leaf_bytes(b_alg, len(grids)) if g.level_id == finest else b_alg) for g in grids}
    bytes_launch = round(grids[-1].n_cells * b_lev[finest])
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    per_level = {str(lid): {"launches": n, "avg_ms": round(ms / max(n, 1), 4), "bytes_per_lu": b_lev[lid],
                            "GBps": round(grids[lid - 1].n_cells * b_lev[lid] / (ms / max(n, 1) * 1e-3) / 1e9, 1)
                            if n else None} for lid, (ms, n) in kern.items()}
    step_bytes = sum(g.n_cells * 2 ** (g.level_id - 1) * b_lev[g.level_id] for g in grids)
    step_lu = sum(g.n_cells * 2 ** (g.level_id - 1) for g in grids)
    log(f"timed {args.steps} coarse steps in {elapsed:.3f}s: MLUPS_true {m_true:.0f}, MLUPS_ref {m_ref:.0f}; "
        f"finest collide {avg_ms:.3f} ms/launch = {achieved:.0f} GB/s; Cd avg {coeffs.get('Cd_avg', float('nan')):.4f} "
        f"over {acc.n_samples} samples, Cd at the end {forces['Cd']:.4f}")

    # practical HBM ceiling of plain streaming kernels on this device
    probe = None
    if args.probe_gib > 0:
        import ctypes as C
        import numpy as np
        from openludwig_amd import _native as N
        pb = np.zeros(5)
        rc = N.hip_lib().lb_bandwidth_probe(0, int(args.probe_gib * (1 << 30)), 10, N.ptr(pb, C.c_double))
        if rc == 0:
            probe = {"copy": round(pb[0], 1), "read": round(pb[1], 1), "write": round(pb[2], 1),
                     "copy_nontemporal": round(pb[3], 1), "copy_nontemporal_grid_stride": round(pb[4], 1),
                     "unit": "GB/s", "buffer_gib": args.probe_gib}
            log(f"bandwidth probe: copy {pb[0]:.0f}, read {pb[1]:.0f}, write {pb[2]:.0f}, "
                f"copy NT {pb[3]:.0f}, copy NT grid-stride {pb[4]:.0f} GB/s")

    cpu = None
    n_cpu = t_end if args.cpu_steps < 0 else args.cpu_steps
    if n_cpu > 0:
        del solver
        cpu = cpu_baseline(cfg, grids, params, mesh, n_cpu, args.diag)

    serving = None
    if args.serving_requests > 0:
        from openludwig_amd import serving as S
        if "solver" in locals():
            del solver
        serving = S.compare(args.serving_requests, cells=tuple(args.serving_cells), log=log)

    line = {
        "metric": "MLUPS & % HBM roofline; served cases/h + p99 latency vs plain FP32",
        "value": round(m_true, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.storage == "fp32" else f"{args.storage}-storage/fp32-compute",
        "data": "synthetic",
        "config": {"workload": f"{wname}, {len(grids)} levels, {sum(n_cells) / 1e6:.1f}M cells, "
                               f"resolution {args.resolution}",
                   "cells_per_level": n_cells, "lu_per_step": sum(c * 2 ** i for i, c in enumerate(n_cells)),
                   "mlups_ref": round(m_ref, 1), "build": "strict" if args.strict else "fast",
                   "population_storage": args.storage, "bytes_per_lu": round(step_bytes / step_lu, 2),
                   "hbm_roofline_pct_step": round(m_true * 1e6 * step_bytes / step_lu / (HBM_PEAK_GBS * 1e9) * 100, 2),
                   "per_level_collide": per_level,
                   "coarse_steps_total": t_end, "timed_window": [t_end - args.steps + 1, t_end],
                   "ramp_steps": cfg.ramp_steps, "force_samples_every": args.diag,
                   "Cd_avg": float(f"{coeffs['Cd_avg']:.6g}") if acc.n_samples else None,
                   "Cl_avg": float(f"{coeffs['Cl_avg']:.6g}") if acc.n_samples else None,
                   "force_samples": acc.n_samples,
                   "Cd_end": float(f"{forces['Cd']:.6g}"), "Cl_end": float(f"{forces['Cl']:.6g}"),
                   "diagnostics_fp64": diagnostics, "thermal": thermal},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "collide_kernel (finest level)", "bytes_per_launch": bytes_launch,
                     "avg_launch_ms": round(avg_ms, 4),
                     "traffic_GBps": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "probe": probe},
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
