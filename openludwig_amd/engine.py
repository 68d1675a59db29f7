"""Batched stepping of many cases on one GPU (lb_engine_* of libludwig_hip).

The reference runs cases one after another (run_all_cases, main.jl:425-469),
each through execute_timestep_batch! (solver_control.jl:199-219).  Here
several cases -- each a Solver built as usual -- are attached to one engine
and advanced together: every launch covers the same recursion depth of all
cases over a concatenated block list, and one engine step is replayed as a
captured hipGraph.  A case's coarse steps, seeds and per-batch inlet
velocities are exactly those of Solver.run, so its results do not depend on
the other cases sharing the engine.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .solver import LudwigError, Solver, averaged_coefficients, ramp_velocity


def ramp_table(cfg, n_steps: int) -> np.ndarray:
    """Inlet velocity of each batch of cfg.gpu_async_depth coarse steps,
    evaluated at the batch end as main.jl:214-218 does."""
    b = cfg.gpu_async_depth
    n_batches = (n_steps + b - 1) // b
    return np.array([ramp_velocity(cfg.u_lattice, min((i + 1) * b, n_steps), cfg.ramp_steps)
                     for i in range(n_batches)], dtype=np.float32)


def default_waves(storage: int) -> int:
    """The interior stream-collide kernel's default occupancy, as the library
    picks it: 5 waves per SIMD for FP32 storage, 6 for 16-bit storage."""
    return 5 if storage == N.LB_STORAGE_FP32 else 6


class Engine:
    def __init__(self, storage: int = N.LB_STORAGE_FP32, strict_fp: bool = False, device: int = 0):
        self.lib = N.hip_lib()
        self._e = C.c_void_p()
        rc = self.lib.lb_engine_create(C.byref(self._e), device, storage, int(strict_fp))
        if rc != 0:
            raise LudwigError(f"lb_engine_create failed ({rc})")
        self.storage = storage
        self.cases: dict = {}
        self.options = {"graph": True, "schedule": N.LB_SCHED_LINEAR, "waves": default_waves(storage)}

    def close(self) -> None:
        if self._e:
            for s in list(self.cases):
                s.engine = None
            self.cases.clear()
            self.lib.lb_engine_destroy(self._e)
            self._e = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self.lib.lb_engine_last_error(self._e).decode(errors="replace")
            raise LudwigError(f"{what} failed ({rc}): {msg}")

    def add(self, solver: Solver, n_steps: int) -> None:
        """Attach an initialised case for coarse steps 1..n_steps."""
        u = ramp_table(solver.cfg, n_steps)
        self._check(self.lib.lb_engine_add(self._e, solver._ctx, int(n_steps), N.ptr(u, C.c_float), len(u),
                                           int(solver.cfg.gpu_async_depth), C.byref(solver.step_params)),
                    "lb_engine_add")
        self.cases[solver] = n_steps
        solver.engine = self

    def remove(self, solver: Solver) -> None:
        self._check(self.lib.lb_engine_remove(self._e, solver._ctx), "lb_engine_remove")
        solver.t_next = self.cases.pop(solver) + 1
        solver.engine = None

    def step(self, n: int = 1) -> None:
        self._check(self.lib.lb_engine_step(self._e, int(n)), "lb_engine_step")

    def synchronize(self) -> None:
        self._check(self.lib.lb_engine_synchronize(self._e), "lb_engine_synchronize")

    def progress(self, solver: Solver) -> int:
        t = C.c_int64()
        self._check(self.lib.lb_engine_case_progress(self._e, solver._ctx, C.byref(t)), "lb_engine_case_progress")
        return t.value

    def done(self, solver: Solver) -> bool:
        return self.progress(solver) >= self.cases[solver]

    def sample_forces(self, solver: Solver, diag_freq: int, ramp_steps: int | None = None) -> None:
        """Sample the case's forces inside the engine step at every coarse
        step t with t % diag_freq == 0 and t > ramp_steps (default: the
        case's ramp), as Solver.run does at its diagnostic steps."""
        ramp = solver.cfg.ramp_steps if ramp_steps is None else ramp_steps
        self._check(self.lib.lb_engine_sample_forces(self._e, solver._ctx, C.byref(solver._force_params()),
                                                     int(diag_freq), int(ramp)), "lb_engine_sample_forces")

    def case_forces(self, solver: Solver) -> dict:
        """The averaged coefficients of the case's samples so far
        ({"n_samples", "Cd_avg", "Cl_avg"}, as Solver.run returns them)."""
        sums = np.zeros(9)
        n = C.c_int64()
        self._check(self.lib.lb_engine_case_forces(self._e, solver._ctx, N.ptr(sums, C.c_double), C.byref(n)),
                    "lb_engine_case_forces")
        return averaged_coefficients(solver.params, sums, n.value)

    def check_finite(self) -> list:
        """Divergence check over every attached case (waits for the enqueued
        work): the solvers whose density or velocity hold non-finite values."""
        cap = max(1, len(self.cases))
        arr = (C.c_void_p * cap)()
        n = C.c_int32()
        rc = self.lib.lb_engine_check_finite(self._e, arr, cap, C.byref(n))
        if rc not in (0, N.LB_ENONFINITE):
            self._check(rc, "lb_engine_check_finite")
        bad = {arr[i] for i in range(min(n.value, cap))}
        return [s for s in self.cases if s._ctx.value in bad]

    def set_options(self, graph: bool = True, schedule: int = N.LB_SCHED_LINEAR, waves: int | None = None) -> None:
        """Launch options (lb_engine_set_options); waves None: the storage's
        default occupancy (5 waves per SIMD for FP32, 6 for 16-bit)."""
        if waves is None:
            waves = default_waves(self.storage)
        self._check(self.lib.lb_engine_set_options(self._e, int(graph), int(schedule), int(waves)),
                    "lb_engine_set_options")
        self.options = {"graph": bool(graph), "schedule": int(schedule), "waves": int(waves)}

    def autotune(self, trial_steps: int = 2, candidates=None, margin: float = 0.03) -> dict:
        """Pick the launch options (graph replay, work-group schedule,
        interior-kernel occupancy) by timing each on the attached cases.
        Every option computes the same results, so the trial steps are real
        steps of the cases.  The current options are timed first and kept
        unless another candidate is faster by more than `margin`: trials on
        a small mix are noisy, and the choice is reused for the mixes that
        follow.  The default candidates leave out LB_SCHED_PERSISTENT, which
        computes the same and measured 20-25 % slower on every mix tried (pass
        it in `candidates` to include it).  Returns the chosen options and
        the timings."""
        import itertools
        import time
        if candidates is None:
            candidates = [dict(graph=g, schedule=sc, waves=w) for g, sc, w in
                          itertools.product((True, False), (N.LB_SCHED_LINEAR, N.LB_SCHED_XCD), (5, 6))]
        start = dict(self.options)
        candidates = [start] + [c for c in candidates if c != start]
        timings = []
        for opt in candidates:
            if not any(not self.done(s) for s in self.cases):
                break
            self.set_options(**opt)
            self.step(1)  # graph capture / warm-up
            self.synchronize()
            t = time.perf_counter()
            self.step(trial_steps)
            self.synchronize()
            timings.append(((time.perf_counter() - t) / trial_steps, opt))
        best = start
        if timings:
            t_start = timings[0][0] if timings[0][1] == start else float("inf")
            t_min, o_min = min(timings, key=lambda x: x[0])
            if t_min < t_start * (1.0 - margin):
                best = o_min
        self.set_options(**best)
        return {"options": best, "trials": [{"s_per_step": round(t, 6), **o} for t, o in timings]}

    def run_to_completion(self) -> None:
        """Step until every attached case has done its coarse steps."""
        while any(not self.done(s) for s in self.cases):
            self.step(1)
        self.synchronize()
