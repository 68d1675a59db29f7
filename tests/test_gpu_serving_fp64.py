"""Served accuracy against the FP64 oracle.  Two small requests (a sphere at
Re 2e3 and a cube at Re 3e5 with the wall model, 2 levels) are served by the
FP16 continuous-batching form (run_batched) and by the plain FP32 form
(run_sequential); each returns the Cd/Cl averaged over its post-ramp force
samples.  The FP64 oracle runs the same cases with the same cadence
(Solver.run's batches of gpu_async_depth coarse steps, the inlet velocity of
each batch end, samples every serving.DIAG coarse steps after the ramp) and
its averaged Cd is the reference.  Bar: both forms within 0.5 % of the FP64
Cd, and Cl within 0.5 % of |Cd| (the bodies are symmetric, Cl is near 0)."""
import pytest

from openludwig_amd import _native as N
from openludwig_amd import serving as S
from openludwig_amd.solver import build_case, ramp_velocity
from oracle import Oracle, oracle_forces

REQS = [S.Request(0, "sphere", reynolds=2e3, levels=2, resolution=8, finest_steps=256),
        S.Request(1, "cube", reynolds=3e5, levels=2, resolution=8, finest_steps=256)]


def _fp64_averaged(req):
    cfg, mesh = S.request_case(req)
    grids, params, mesh = build_case(cfg, mesh)
    n = req.coarse_steps(len(grids))
    o = Oracle(cfg, grids, params, "f64")
    o.init_equilibrium()
    b = cfg.gpu_async_depth
    cd, cl = [], []
    for t in range(1, n + 1, b):
        be = min(t + b - 1, n)
        o.step_batch(t, be - t + 1, ramp_velocity(cfg.u_lattice, be, cfg.ramp_steps))
        d = (be // S.DIAG) * S.DIAG
        if t <= d <= be and d > cfg.ramp_steps:
            f = oracle_forces(o, cfg, params, mesh, grids[-1].dx)
            cd.append(f["Cd"])
            cl.append(f["Cl"])
    o.close()
    return sum(cd) / len(cd), sum(cl) / len(cl), len(cd)


@pytest.mark.gpu
def test_served_cd_within_half_percent_of_fp64(hip_device):
    ref = {r.rid: _fp64_averaged(r) for r in REQS}
    batched = S.run_batched(REQS, storage=N.LB_STORAGE_FP16, setup_workers=2, chunk=2)
    plain = S.run_sequential(REQS, storage=N.LB_STORAGE_FP32)
    for form, res in (("fp16 batched", batched), ("fp32 plain", plain)):
        for r in res:
            cd64, cl64, n64 = ref[r.rid]
            print(form, r.rid, r.cd, cd64, r.cl, cl64, r.n_samples)
            assert r.status == "ok" and r.n_samples == n64 > 0
            assert abs(r.cd - cd64) <= 5e-3 * abs(cd64), f"{form} request {r.rid}: Cd {r.cd} vs FP64 {cd64}"
            assert abs(r.cl - cl64) <= 5e-3 * abs(cd64), f"{form} request {r.rid}: Cl {r.cl} vs FP64 {cl64}"
