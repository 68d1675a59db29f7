"""Serving of case requests on one GPU: continuous batching against a plain
sequential loop.

A request names a body (sphere, cube or ellipsoid), a Reynolds number, a
level count and a grid resolution, and asks for the force coefficients after
a fixed amount of simulated time.  Two ways to serve a stream of requests:

run_sequential   the reference's way (run_all_cases, main.jl:425-469): one
                 case at a time -- set up, run through execute_timestep_batch!
                 in batches of gpu_async_depth, compute forces, free.
run_batched      continuous batching: set-up (block topology on the CPU;
                 voxelization, wall distances and Bouzidi Q-map on the GPU;
                 then device upload) runs on worker threads as requests
                 arrive; a case is admitted when its exact device footprint (lb_footprint) fits
                 the HBM budget; admitted cases are attached to one Engine and
                 advance together in batched launches replayed as a hipGraph;
                 finished cases are detached, their forces computed and their
                 memory released while the others keep running.

Each served result carries the force coefficients averaged over the samples
after the inflow ramp (accumulate_forces!, surface.jl:36-56), taken every
DIAG coarse steps: by Solver.run in the plain form and inside the engine step
in the batched form.  A case that diverges (non-finite density or velocity at
a sample or host check) is stopped and reported with status "diverged"; the
others go on unaffected.

Latency = finish - arrival.  Every per-case result equals the case run alone
(the engine's per-case sequence does not depend on the mix).
"""
from __future__ import annotations

import ctypes as C
import math
import threading
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .config import CaseConfig, f32
from .domain import compute_domain_from_mesh, setup_multilevel_domain
from .engine import Engine, ramp_table
from .geometry import cube, ellipsoid, icosphere, mesh_from_triangles, rotate_mesh, subdivide_mesh
from .solver import DivergenceError, Solver, build_case, footprint

NU_AIR = 1.5e-5
DIAG = 8  # force-sampling cadence of served cases, coarse steps (a multiple of the ramp batch)


@dataclass
class Request:
    rid: int
    body: str                      # "sphere", "cube" or "ellipsoid"
    axes: tuple = (1.0, 1.0, 1.0)  # ellipsoid full lengths (x = 1)
    reynolds: float = 1e3
    levels: int = 3
    resolution: int = 16
    finest_steps: int = 256        # simulated time, in finest-level steps
    arrival: float = 0.0           # seconds after the start of the run
    cells: int = 0                 # active cells (from the block topology)

    def coarse_steps(self, levels: int) -> int:
        n = max(1, self.finest_steps >> (levels - 1))
        return int(math.ceil(n / 8.0) * 8)


@dataclass
class Result:
    rid: int
    arrival: float
    start: float
    finish: float
    levels: int
    cells: int
    coarse_steps: int
    lu: int
    cd: float                      # averaged over the post-ramp samples
    cl: float
    device_bytes: int = 0
    n_samples: int = 0
    status: str = "ok"             # "ok" or "diverged"
    mass_drift: float = 0.0        # relative change of the FP64 population sum over the run

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def request_case(req: Request) -> tuple:
    """Case configuration and mesh of a request.  Bodies are 1 m long in x;
    the wall model (WMLES) is on from Re 1e5."""
    if req.body == "sphere":
        tris, area = icosphere(4, 0.5), math.pi / 4.0
    elif req.body == "cube":
        tris, area = cube(1.0), 1.0
    elif req.body == "ellipsoid":
        tris = ellipsoid(4, req.axes)
        area = math.pi / 4.0 * req.axes[1] * req.axes[2]
    else:
        raise ValueError(f"unknown body {req.body}")
    cfg = CaseConfig(
        surface_resolution=int(req.resolution), num_levels=int(req.levels), reference_area_full_model=area,
        reference_chord=1.0, reference_length_for_meshing=1.0, reference_dimension="x",
        fluid_density=1.225, kinematic_viscosity=NU_AIR, flow_velocity=float(req.reynolds) * NU_AIR,
        steps=0, ramp_steps=max(8, req.coarse_steps(req.levels) // 2), output_freq=10 ** 9,
        u_lattice=f32(0.05), c_wale=f32(0.5), tau_min=f32(0.501), inlet_turbulence_intensity=f32(0.0),
        nu_sgs_background=f32(0.0), limiter="positivity", lambda_param=f32(0.1667),
        wall_model_enabled=req.reynolds >= 1e5, domain_upstream=3.0, domain_downstream=6.0, domain_lateral=2.5,
        domain_height=2.5, sponge_thickness=f32(0.10), refinement_margin=2, wake_enabled=True,
        wake_length=1.0, wake_width_factor=1.0, wake_height_factor=1.0, boundary_method="bouzidi",
        bouzidi_levels=1, q_min_threshold=f32(0.001), forces_enabled=True, moment_center=(0.25, 0.0, 0.0),
        diag_freq=10 ** 9, min_coarse_blocks=1,
    ).validate()
    mesh = mesh_from_triangles(tris)
    if req.body == "cube":
        mesh = subdivide_mesh(mesh, 0.1)
    return cfg, mesh


def topology_cells(req: Request) -> tuple:
    """(active cells, levels) of a request from its block topology alone."""
    cfg, mesh = request_case(req)
    mesh = rotate_mesh(mesh, cfg.alpha_deg, cfg.beta_deg)
    params = compute_domain_from_mesh(cfg, mesh.min_bounds, mesh.max_bounds)
    nbs = setup_multilevel_domain(cfg, mesh, params, topology_only=True)
    return sum(nbs) * 512, len(nbs)


def size_request(req: Request, target_cells: float, iters: int = 5) -> Request:
    """Pick a resolution whose active cell count is close to the target:
    secant steps on log(cells) against log(resolution), from resolution 16."""
    levels = req.levels
    res, best = 16, None
    pts = []
    for _ in range(iters):
        req.resolution, req.levels = res, levels
        n, lv = topology_cells(req)
        if best is None or abs(math.log(n / target_cells)) < abs(math.log(best[1] / target_cells)):
            best = (res, n, lv)
        pts.append((math.log(res), math.log(n)))
        if abs(math.log(n / target_cells)) < 0.1:
            break
        slope = 2.5
        if len(pts) >= 2 and abs(pts[-1][0] - pts[-2][0]) > 1e-9:
            slope = min(max((pts[-1][1] - pts[-2][1]) / (pts[-1][0] - pts[-2][0]), 0.5), 3.0)
        nxt = int(round(math.exp(pts[-1][0] + (math.log(target_cells) - pts[-1][1]) / slope)))
        nxt = min(max(nxt, 4), 512)
        if nxt == res:
            break
        res = nxt
    req.resolution, req.cells, req.levels = best
    return req


def make_requests(n: int, rate: float, cells: tuple = (1e6, 40e6), levels: tuple = (2, 5),
                  reynolds: tuple = (1e3, 1e7), finest_steps: int = 256, seed_arrivals: int = 0,
                  seed_bodies: int = 1) -> list:
    """The serving workload: Poisson arrivals at `rate` per second (seed 0);
    bodies uniform over sphere/cube/ellipsoid with ellipsoid axis ratios
    U[0.5, 2] on y and z; Re log-uniform; levels uniform; target active
    cells log-uniform, reached through the resolution (seed 1)."""
    ra = np.random.default_rng(seed_arrivals)
    rb = np.random.default_rng(seed_bodies)
    gaps = ra.exponential(1.0 / rate, size=n) if rate > 0 else np.zeros(n)
    gaps[0] = 0.0
    arrivals = np.cumsum(gaps)
    todo = []
    for i in range(n):
        body = ["sphere", "cube", "ellipsoid"][int(rb.integers(0, 3))]
        axes = (1.0, float(rb.uniform(0.5, 2.0)), float(rb.uniform(0.5, 2.0)))
        re = float(np.exp(rb.uniform(np.log(reynolds[0]), np.log(reynolds[1]))))
        lv = int(rb.integers(levels[0], levels[1] + 1))
        target = float(np.exp(rb.uniform(np.log(cells[0]), np.log(cells[1]))))
        req = Request(rid=i, body=body, axes=axes, reynolds=re, levels=lv, finest_steps=finest_steps,
                      arrival=float(arrivals[i]))
        todo.append((req, target))
    with ThreadPoolExecutor(max_workers=8) as pool:
        return list(pool.map(lambda rt: size_request(*rt), todo))


def _lu(grids, coarse_steps: int) -> int:
    return sum(g.n_cells * 2 ** (g.level_id - 1) for g in grids) * coarse_steps


def device_memory(device: int = 0) -> tuple:
    """(free, total) bytes of the HIP device."""
    free, total = C.c_int64(), C.c_int64()
    if N.hip_lib().lb_device_memory(device, C.byref(free), C.byref(total)) != 0:
        raise RuntimeError("lb_device_memory failed")
    return free.value, total.value


def _wait_until(t0: float, t: float) -> None:
    d = t0 + t - time.perf_counter()
    if d > 0:
        time.sleep(d)


def run_sequential(requests: list, storage: int = N.LB_STORAGE_FP32, strict_fp: bool = False, log=None,
                   gpu_setup: bool = False) -> list:
    """One case at a time, in arrival order (run_all_cases): set up, run the
    main loop with force samples every DIAG coarse steps, free.  With
    gpu_setup the voxelization, wall distances and Q-map run on the GPU (the
    batched form's set-up path), else on the CPU as the reference does."""
    t0 = time.perf_counter()
    out = []
    for req in sorted(requests, key=lambda r: r.arrival):
        _wait_until(t0, req.arrival)
        start = time.perf_counter() - t0
        cfg, mesh = request_case(req)
        grids, params, mesh = build_case(cfg, mesh, gpu_device=0 if gpu_setup else None)
        n = req.coarse_steps(len(grids))
        s = Solver(cfg, grids, params, mesh, strict_fp=strict_fp, storage=storage)
        s.init_equilibrium()
        m0 = s.mass()
        status = "ok"
        try:
            res = s.run(n, diag_freq=DIAG)
            cd, cl, ns = res.get("Cd_avg", math.nan), res.get("Cl_avg", math.nan), res["n_samples"]
        except DivergenceError:
            status, cd, cl, ns = "diverged", math.nan, math.nan, 0
        drift = (s.mass() - m0) / m0
        nbytes = s.device_bytes
        s.close()
        out.append(Result(req.rid, req.arrival, start, time.perf_counter() - t0, len(grids),
                          sum(g.n_cells for g in grids), n, _lu(grids, n), cd, cl, nbytes, ns, status, drift))
        if log:
            log(f"  sequential: request {req.rid} {status} at {out[-1].finish:.1f}s "
                f"({out[-1].cells / 1e6:.1f}M cells)")
    return sorted(out, key=lambda r: r.rid)


@dataclass
class _Prepared:
    req: Request
    solver: Solver
    grids: list
    n_steps: int
    nbytes: int
    start: float
    mass0: float = 0.0


@dataclass
class _Budget:
    limit: int
    used: int = 0
    cv: threading.Condition = field(default_factory=threading.Condition)

    peak: int = 0

    def acquire(self, n: int) -> None:
        with self.cv:
            if n > self.limit:
                raise MemoryError(f"case needs {n} bytes, budget is {self.limit}")
            while self.used + n > self.limit:
                self.cv.wait()
            self.used += n
            self.peak = max(self.peak, self.used)

    def release(self, n: int) -> None:
        with self.cv:
            self.used -= n
            self.cv.notify_all()


_TUNED: dict = {}


def run_batched(requests: list, storage: int = N.LB_STORAGE_FP16, strict_fp: bool = False,
                hbm_budget: int | None = None, reserve: int = 8 << 30, setup_workers: int = 4,
                chunk: int = 4, options: dict | None = None, tune: bool = True, log=None,
                gpu_setup: bool = True, info: dict | None = None, prepared_hook=None) -> list:
    """Continuous batching over one Engine (see the module docstring).
    Launch options: `options` if given, else autotuned once per (storage,
    deepest level count) on the first cases admitted, and reused.  With
    gpu_setup, the voxelization, wall distances and Bouzidi Q-map of each case
    run on the GPU (the same bits as the CPU restatement the plain form uses).
    Admitted cases sample their forces inside the engine step every DIAG
    coarse steps after their ramp; after every chunk of engine steps the
    divergence check runs over all cases and evicts any that diverged.
    `info` (if given) receives the budget and the peak admitted bytes;
    `prepared_hook(solver, request)` runs on each case before it is attached."""
    if hbm_budget is None:
        free, _total = device_memory()
        hbm_budget = free - reserve
    budget = _Budget(hbm_budget)
    engine = Engine(storage=storage, strict_fp=strict_fp)
    if options:
        engine.set_options(**options)
    t0 = time.perf_counter()
    ready: deque = deque()
    lock = threading.Lock()
    errors: list = []

    def prepare(req: Request) -> None:
        try:
            start = time.perf_counter() - t0
            cfg, mesh = request_case(req)
            grids, params, mesh = build_case(cfg, mesh, gpu_device=0 if gpu_setup else None)
            n = req.coarse_steps(len(grids))
            nbytes = footprint(cfg, grids, params, mesh, storage, len(ramp_table(cfg, n)))
            budget.acquire(nbytes)
            s = Solver(cfg, grids, params, mesh, strict_fp=strict_fp, storage=storage)
            s.init_equilibrium()
            if prepared_hook is not None:
                prepared_hook(s, req)
            m0 = s.mass()
            with lock:
                ready.append(_Prepared(req, s, grids, n, nbytes, start, m0))
        except BaseException as exc:  # surfaced by the serving loop
            with lock:
                errors.append(exc)

    pending = sorted(requests, key=lambda r: r.arrival)
    active: dict = {}
    out = []
    # where the loop's wall time goes: engine steps (enqueue + wait), the
    # per-chunk checks and completions, tuning, and waiting with nothing admitted
    tm = {"steps_s": 0.0, "checks_and_completions_s": 0.0, "tune_s": 0.0, "idle_s": 0.0, "chunks": 0}
    with ThreadPoolExecutor(max_workers=setup_workers) as pool:
        while len(out) < len(requests):
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                pool.submit(prepare, pending.pop(0))
            with lock:
                if errors:
                    raise errors[0]
                while ready:
                    p = ready.popleft()
                    engine.add(p.solver, p.n_steps)
                    engine.sample_forces(p.solver, DIAG)
                    active[p.solver] = p
            if active and not options:
                key = (storage, max(len(p.grids) for p in active.values()))
                ta = time.perf_counter()
                if key not in _TUNED and tune:
                    _TUNED[key] = engine.autotune()["options"]
                if key in _TUNED and engine.options != _TUNED[key]:
                    engine.set_options(**_TUNED[key])
                tm["tune_s"] += time.perf_counter() - ta
            if active:
                ta = time.perf_counter()
                engine.step(chunk)
                engine.synchronize()
                tb = time.perf_counter()
                tm["steps_s"] += tb - ta
                tm["chunks"] += 1
                diverged = set(engine.check_finite())
                for s in [s for s in active if s in diverged or engine.done(s)]:
                    p = active.pop(s)
                    if s in diverged:
                        status, cd, cl, ns = "diverged", math.nan, math.nan, 0
                    else:
                        fr = engine.case_forces(s)
                        status, ns = "ok", fr["n_samples"]
                        cd, cl = fr.get("Cd_avg", math.nan), fr.get("Cl_avg", math.nan)
                    engine.remove(s)
                    drift = (s.mass() - p.mass0) / p.mass0
                    nbytes = s.device_bytes
                    s.close()
                    budget.release(p.nbytes)
                    out.append(Result(p.req.rid, p.req.arrival, p.start, time.perf_counter() - t0,
                                      len(p.grids), sum(g.n_cells for g in p.grids), p.n_steps,
                                      _lu(p.grids, p.n_steps), cd, cl, nbytes, ns, status, drift))
                    if log:
                        log(f"  batched: request {p.req.rid} {status} at {out[-1].finish:.1f}s "
                            f"({out[-1].cells / 1e6:.1f}M cells, {len(active)} still running)")
                tm["checks_and_completions_s"] += time.perf_counter() - tb
            else:
                ta = time.perf_counter()
                nxt = pending[0].arrival - now if pending else 0.002
                time.sleep(min(max(nxt, 0.0), 0.002))
                tm["idle_s"] += time.perf_counter() - ta
    engine.close()
    if info is not None:
        info.update({"budget_bytes": int(budget.limit), "peak_admitted_bytes": int(budget.peak),
                     "loop": {k: (round(v, 3) if isinstance(v, float) else v) for k, v in tm.items()}})
    return sorted(out, key=lambda r: r.rid)


def summary(results: list) -> dict:
    lat = np.array([r.latency for r in results])
    span = max(r.finish for r in results) - min(r.arrival for r in results)
    lu = sum(r.lu for r in results)
    return {"n": len(results), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "span_s": float(span), "cases_per_hour": len(results) / span * 3600.0,
            "mlups_true": lu / span / 1e6}


def compare(n: int = 64, cells: tuple = (1e6, 40e6), finest_steps: int = 1024, load: float = 1.5,
            setup_workers: int = 4, log=print, verbose: bool = False, same_setup: bool = True) -> dict:
    """The serving benchmark: plain FP32 sequential against FP16 continuous
    batching on the same requests.  Capacity: cases/h of each form with all
    requests present at t = 0 (closed runs); the plain form once with the
    reference's CPU set-up and, with same_setup, once with the GPU set-up the
    batched form uses, so that the ratio can be read with the set-up path
    held equal.  Latency: p50/p99 of both forms under Poisson arrivals at
    `load` times the plain (CPU set-up) form's capacity.  Accuracy: the
    post-ramp averaged Cd/Cl of every request, FP16 batched against plain
    FP32."""
    reqs = make_requests(n, rate=0.0, cells=cells, finest_steps=finest_steps)
    log(f"serving: {n} requests, {sum(r.cells for r in reqs) / 1e6:.0f}M cells in total")
    plog = log if verbose else None
    closed = run_sequential(reqs, storage=N.LB_STORAGE_FP32, log=plog)
    cs = summary(closed)
    mu = n / cs["span_s"]
    log(f"serving: plain FP32 closed run {cs['span_s']:.1f}s for {n} cases ({cs['cases_per_hour']:.0f} cases/h)")
    cg = None
    if same_setup:
        closed_g = run_sequential(reqs, storage=N.LB_STORAGE_FP32, log=plog, gpu_setup=True)
        cg = summary(closed_g)
        log(f"serving: plain FP32 closed run, GPU set-up, {cg['span_s']:.1f}s ({cg['cases_per_hour']:.0f} cases/h)")
    info: dict = {}
    closed_b = run_batched(reqs, storage=N.LB_STORAGE_FP16, setup_workers=setup_workers, log=plog, info=info)
    cb = summary(closed_b)
    log(f"serving: FP16 batched closed run {cb['span_s']:.1f}s ({cb['cases_per_hour']:.0f} cases/h)")
    gaps = np.random.default_rng(0).exponential(1.0 / (load * mu), size=n)
    gaps[0] = 0.0
    for r, t in zip(reqs, np.cumsum(gaps)):
        r.arrival = float(t)
    seq = run_sequential(reqs, storage=N.LB_STORAGE_FP32, log=plog)
    bat = run_batched(reqs, storage=N.LB_STORAGE_FP16, setup_workers=setup_workers, log=plog)
    ss, bs = summary(seq), summary(bat)
    ok = [(a, b) for a, b in zip(closed, closed_b) if a.status == "ok" and b.status == "ok"]
    dcd = [abs(a.cd - b.cd) / max(abs(a.cd), 1e-12) for a, b in ok]
    log(f"serving: sequential FP32 p50 {ss['p50_s']:.2f}s p99 {ss['p99_s']:.2f}s {ss['cases_per_hour']:.0f} cases/h; "
        f"batched FP16 p50 {bs['p50_s']:.2f}s p99 {bs['p99_s']:.2f}s {bs['cases_per_hour']:.0f} cases/h")
    cap = {"plain_fp32_sequential": round(cs["cases_per_hour"], 1),
           "fp16_batched": round(cb["cases_per_hour"], 1),
           "ratio": round(cb["cases_per_hour"] / cs["cases_per_hour"], 3)}
    if cg is not None:
        cap["plain_fp32_sequential_gpu_setup"] = round(cg["cases_per_hour"], 1)
        cap["ratio_same_setup"] = round(cb["cases_per_hour"] / cg["cases_per_hour"], 3)
    return {"requests": n, "cells_range": list(cells), "finest_steps": finest_steps, "diag_every": DIAG,
            "total_cells": int(sum(r.cells for r in reqs)),
            "capacity_cases_per_hour": cap,
            "admission": {"budget_gb": round(info.get("budget_bytes", 0) / 1e9, 1),
                          "peak_admitted_gb": round(info.get("peak_admitted_bytes", 0) / 1e9, 2)},
            "batched_closed_loop_time": info.get("loop"),
            "arrival_rate_per_s": round(load * mu, 4), "load_vs_plain_throughput": load,
            "plain_fp32_sequential": {k: round(v, 3) for k, v in ss.items()},
            "fp16_batched": {k: round(v, 3) for k, v in bs.items()},
            "cases_per_hour_ratio": round(bs["cases_per_hour"] / ss["cases_per_hour"], 3),
            "p99_ratio": round(ss["p99_s"] / bs["p99_s"], 3),
            "diverged": {"plain_fp32": sum(r.status != "ok" for r in closed),
                         "fp16_batched": sum(r.status != "ok" for r in closed_b)},
            "cd_avg_rel_diff_fp16_vs_fp32": {"max": float(max(dcd)) if dcd else None,
                                             "median": float(np.median(dcd)) if dcd else None,
                                             "within_0.5pct": int(sum(d <= 5e-3 for d in dcd)),
                                             "compared": len(dcd)},
            "per_request": [{"rid": a.rid, "body": r.body, "re": round(r.reynolds, 1), "cells": a.cells,
                             "levels": a.levels, "coarse_steps": a.coarse_steps, "samples": a.n_samples,
                             "cd_avg_fp32": round(a.cd, 6), "cd_avg_fp16": round(b.cd, 6),
                             "cl_avg_fp32": round(a.cl, 6), "cl_avg_fp16": round(b.cl, 6),
                             "status": [a.status, b.status],
                             "mass_drift_fp32": float(f"{a.mass_drift:.3g}"),
                             "mass_drift_fp16": float(f"{b.mass_drift:.3g}"),
                             "latency_fp32_s": round(x.latency, 3), "latency_fp16_s": round(y.latency, 3)}
                            for a, b, x, y, r in zip(closed, closed_b, seq, bat, sorted(reqs, key=lambda q: q.rid))]}


def main() -> int:
    import argparse
    import json
    ap = argparse.ArgumentParser(description="serving benchmark: FP16 continuous batching vs plain FP32")
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--min-cells", type=float, default=1e6)
    ap.add_argument("--max-cells", type=float, default=40e6)
    ap.add_argument("--finest-steps", type=int, default=1024)
    ap.add_argument("--load", type=float, default=1.5)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--no-same-setup", action="store_true", help="skip the plain form's GPU set-up run")
    a = ap.parse_args()
    res = compare(a.requests, (a.min_cells, a.max_cells), a.finest_steps, a.load, a.workers,
                  log=lambda m: print(m, flush=True), verbose=True, same_setup=not a.no_same_setup)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
