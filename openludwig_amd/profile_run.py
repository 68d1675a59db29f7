"""Workload driver for rocprofv3 runs.

Runs the benchmark case for a few coarse steps (after a warm-up batch) and,
optionally, one plain device-to-device copy of a known size, so that HBM
counters (FETCH_SIZE / WRITE_SIZE) of the stream-collide kernels can be
scaled against a copy of known traffic.

    rocprofv3 --pmc FETCH_SIZE -- python3 -m openludwig_amd.profile_run --steps 2
"""
from __future__ import annotations

import argparse
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--case", default="cube", choices=["cube", "sphere"])
    ap.add_argument("--resolution", type=int, default=160)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--copy-mib", type=int, default=1024, help="size of the calibration copy (0: none)")
    ap.add_argument("--strict", action="store_true")
    ap.add_argument("--storage", type=int, default=0, help="population storage: 0 FP32, 1 FP16, 2 BF16 offsets")
    ap.add_argument("--probe-gib", type=float, default=0.0, help="run the HBM bandwidth probe on this many GiB")
    args = ap.parse_args()

    from openludwig_amd import cases
    from openludwig_amd.solver import Solver, build_case, ramp_velocity

    if args.case == "cube":
        cfg, mesh = cases.cube_re1e5(args.resolution)
    else:
        cfg, mesh = cases.sphere_re1e3()
    grids, params, mesh = build_case(cfg, mesh)
    s = Solver(cfg, grids, params, mesh, strict_fp=args.strict, storage=args.storage)
    s.init_equilibrium()
    t = 1
    for n in (args.warmup, args.steps):
        if n > 0:
            s.execute_timestep_batch(t, n, ramp_velocity(cfg.u_lattice, t + n - 1, cfg.ramp_steps))
            t += n
    s.synchronize()
    print("levels", [(g.level_id, g.n_blocks, g.n_cells) for g in grids], flush=True)
    print("collide grid sizes (work-items):", [g.n_blocks * 512 for g in grids], flush=True)
    if args.probe_gib > 0:
        import ctypes as C

        import numpy as np

        from openludwig_amd import _native as N

        out = np.zeros(5)
        rc = N.hip_lib().lb_bandwidth_probe(0, int(args.probe_gib * (1 << 30)), 10, N.ptr(out, C.c_double))
        print(f"bandwidth probe rc={rc}: copy {out[0]:.0f} GB/s, read {out[1]:.0f} GB/s, write {out[2]:.0f} GB/s, "
              f"copy non-temporal {out[3]:.0f} GB/s, copy non-temporal grid-stride {out[4]:.0f} GB/s", flush=True)
    if args.copy_mib > 0:
        # one device-to-device copy through the HIP runtime (a blit kernel)
        import ctypes as C

        hip = C.CDLL("libamdhip64.so")
        n = args.copy_mib * (1 << 20)
        src, dst = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0
        assert hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
        assert hip.hipMemset(src, 1, C.c_size_t(n)) == 0
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(src)
        hip.hipFree(dst)
        print(f"copy kernel: {n} bytes read, {n} bytes written", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
