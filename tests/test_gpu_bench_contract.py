"""bench.py's output line: one JSON object with the fields the benchmark
contract names, a roofline object whose fraction is its achieved rate over
the HBM peak, and the CPU baseline of the oracle.  The GPU test runs the
cube workload at a small resolution (a few million cells, 2 timed coarse
steps) in a child process: FP32 after a short spin-up with force samples,
and FP16 storage (its dtype names the 16-bit storage); the CPU tests check
the refusal of N > 1 and the CPU baseline leg (bench.cpu_baseline) on a
small case: its fields, its force cadence, and its averaged Cd/Cl against
the FP64 oracle run with the main loop's cadence written out here."""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], cwd=ROOT, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout.strip() == ""


def test_cpu_baseline_leg_against_fp64_oracle():
    sys.path.insert(0, str(ROOT))
    sys.path.insert(0, str(ROOT / "oracle"))
    import bench
    from oracle import Oracle, oracle_forces

    from openludwig_amd import cases
    from openludwig_amd.solver import build_case

    cfg, mesh = cases.tiny_sphere(8, 2)
    cfg = cfg.replace(ramp_steps=4)
    grids, params, mesh = build_case(cfg, mesh)
    n, diag = 24, 8
    cpu = bench.cpu_baseline(cfg, grids, params, mesh, n, diag)
    for key in ("value", "unit", "cores", "kind", "sample", "seconds", "cpu_model", "host_cpus"):
        assert key in cpu
    assert cpu["kind"] == "port" and cpu["unit"] == "MLUPS" and cpu["value"] > 0 and cpu["cores"] >= 1
    # main.jl:214-243: batches of gpu_async_depth (8) at the inlet velocity
    # of the batch end, a force sample after every batch holding a multiple
    # of the cadence past the ramp: steps 8, 16, 24.  Every batch ends past
    # the 4-step ramp, so every batch runs at the full inlet velocity.
    assert cfg.gpu_async_depth == 8
    assert cpu["force_samples"] == 3
    o = Oracle(cfg, grids, params, "f64")
    o.init_equilibrium()
    cd, cl = [], []
    for t in range(1, n + 1, 8):
        o.step_batch(t, 8, float(cfg.u_lattice))
        f = oracle_forces(o, cfg, params, mesh, grids[-1].dx)
        cd.append(f["Cd"])
        cl.append(f["Cl"])
    o.close()
    cd64, cl64 = sum(cd) / 3, sum(cl) / 3
    assert abs(cpu["Cd_avg"] - cd64) <= 5e-3 * abs(cd64), (cpu["Cd_avg"], cd64)
    assert abs(cpu["Cl_avg"] - cl64) <= 5e-3 * abs(cd64), (cpu["Cl_avg"], cl64)
    one = bench.cpu_baseline(cfg, grids, params, mesh, 1, diag)
    assert "force_samples" not in one and "1 coarse step(s)" in one["sample"]


@pytest.mark.gpu
@pytest.mark.parametrize("storage", ["fp32", "fp16"])
def test_bench_json_line_contract(storage):
    extra = ["--total", "40", "--ramp", "8", "--diag", "8"] if storage == "fp32" else ["--total", "0"]
    r = subprocess.run([sys.executable, "bench.py", "--resolution", "48", "--steps", "2", "--warmup", "1",
                        "--cpu-steps", "1", "--serving-requests", "0", "--probe-gib", "0", "--storage", storage]
                       + extra, cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, r.stdout
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["higher_is_better"] is True and d["unit"] == "MLUPS" and d["value"] > 0
    assert d["dtype"] == ("fp32" if storage == "fp32" else "fp16-storage/fp32-compute")
    assert "workload" in d["config"] and d["config"]["population_storage"] == storage
    c = d["config"]
    if storage == "fp32":
        # coarse steps 1..40: samples after the ramp at 16, 24, 32 (spin-up) and 40 (end of the timed window)
        assert c["timed_window"] == [39, 40] and c["force_samples"] == 4 and c["Cd_avg"] is not None
    else:
        assert c["timed_window"] == [2, 3] and c["force_samples"] == 0
    assert c["diagnostics_fp64"]["v_max"] > 0.0
    assert abs(d["ms_per_step"] - 1e3 * d["config"]["lu_per_step"] / (d["value"] * 1e6)) <= 0.01 * d["ms_per_step"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert 0.0 < rf["frac"] < 1.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 1e-3
    assert abs(rf["achieved"] - rf["bytes_per_launch"] / (rf["avg_launch_ms"] * 1e-3) / 1e9) <= 0.01 * rf["achieved"]
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["value"] > 0 and cb["cores"] >= 1 and cb["unit"] == "MLUPS"
    assert cb["cpu_model"] != "unknown"
