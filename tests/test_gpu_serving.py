"""Continuous-batching serving: requests that arrive while others run are
admitted as memory frees up, share the engine, and each result -- the Cd and
Cl averaged over the post-ramp force samples, taken inside the engine step --
equals the same request served alone and served by the sequential loop
(Solver.run's samples at its diagnostic steps), bit for bit."""
import pytest

from openludwig_amd import _native as N
from openludwig_amd import serving as S
from openludwig_amd.solver import build_case, footprint


def _requests():
    reqs = [S.Request(0, "sphere", reynolds=2e3, levels=2, resolution=8, finest_steps=32, arrival=0.0),
            S.Request(1, "cube", reynolds=3e5, levels=3, resolution=10, finest_steps=64, arrival=0.05),
            S.Request(2, "ellipsoid", axes=(1.0, 1.5, 0.7), reynolds=5e4, levels=3, resolution=8,
                      finest_steps=64, arrival=0.1),
            S.Request(3, "sphere", reynolds=1e6, levels=2, resolution=10, finest_steps=32, arrival=0.1)]
    return reqs


@pytest.mark.gpu
def test_serving_batched_equals_alone_and_sequential(hip_device):
    reqs = _requests()
    # budget for about two cases at a time, so admission has to wait for memory
    sizes = []
    for r in reqs:
        cfg, mesh = S.request_case(r)
        grids, params, mesh = build_case(cfg, mesh)
        sizes.append(footprint(cfg, grids, params, mesh, N.LB_STORAGE_FP16, 64))
    budget = sorted(sizes)[-1] + sorted(sizes)[-2]
    mixed = S.run_batched(reqs, storage=N.LB_STORAGE_FP16, hbm_budget=budget, setup_workers=2, chunk=2)
    assert [r.rid for r in mixed] == [0, 1, 2, 3]
    for r in reqs:
        alone = S.run_batched([r], storage=N.LB_STORAGE_FP16, hbm_budget=budget, setup_workers=1, chunk=2)[0]
        m = mixed[r.rid]
        assert (m.cd, m.cl) == (alone.cd, alone.cl), f"request {r.rid}: mixed {m.cd, m.cl} alone {alone.cd, alone.cl}"
        assert m.coarse_steps == alone.coarse_steps
    seq = S.run_sequential(reqs, storage=N.LB_STORAGE_FP16)
    for a, b in zip(seq, mixed):
        assert (a.cd, a.cl) == (b.cd, b.cl), f"request {a.rid}: sequential {a.cd, a.cl} batched {b.cd, b.cl}"
    assert all(r.n_samples >= 1 and r.status == "ok" for r in mixed)
    summ = S.summary(mixed)
    assert summ["n"] == 4 and summ["p99_s"] >= summ["p50_s"] > 0
