"""The batched serving configuration at full size: 64 requests drawn by the
serving generator (bodies sphere/cube/ellipsoid, Re log-uniform [1e3, 1e7],
2-5 levels, 1-40 M active cells), all present at once, through
run_batched under an HBM budget of 60 GB -- less than the set needs, so
admission waits for memory.

  * every request is served: status ok, finite post-ramp averaged Cd/Cl
    from at least one force sample;
  * the admitted bytes never exceed the budget;
  * one request of each level count equals its run alone, bit for bit;
  * the same requests served by the plain FP32 form: the FP16-storage
    batched form's averaged Cd within the stated bar of it.
"""
import math

import pytest

from openludwig_amd import _native as N
from openludwig_amd import serving as S

BUDGET = 60 * 10 ** 9


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_serving_request_set_full_size(hip_device):
    reqs = S.make_requests(64, rate=0.0, cells=(1e6, 40e6), finest_steps=1024)
    assert len(reqs) == 64 and {r.levels for r in reqs} == {2, 3, 4, 5}
    total = sum(r.cells for r in reqs)
    print(f"{len(reqs)} requests, {total / 1e6:.0f}M cells, {min(r.cells for r in reqs) / 1e6:.2f}-"
          f"{max(r.cells for r in reqs) / 1e6:.1f}M each", flush=True)
    info: dict = {}
    res = S.run_batched(reqs, storage=N.LB_STORAGE_FP16, hbm_budget=BUDGET, info=info)
    assert [r.rid for r in res] == list(range(64))
    for r in res:
        assert r.status == "ok" and r.n_samples >= 1, (r.rid, r.status, r.n_samples)
        assert math.isfinite(r.cd) and math.isfinite(r.cl), (r.rid, r.cd, r.cl)
    print("admission", info, flush=True)
    assert 0 < info["peak_admitted_bytes"] <= info["budget_bytes"] == BUDGET
    sample = [next(r for r in reqs if r.levels == lv) for lv in (2, 3, 4, 5)]
    for r in sample:
        alone = S.run_batched([r], storage=N.LB_STORAGE_FP16, hbm_budget=BUDGET)[0]
        m = res[r.rid]
        assert (alone.cd, alone.cl) == (m.cd, m.cl), f"request {r.rid}: alone {alone.cd, alone.cl} mixed {m.cd, m.cl}"
    plain = S.run_sequential(sample, storage=N.LB_STORAGE_FP32)
    for p in plain:
        b = res[p.rid]
        d = abs(b.cd - p.cd) / abs(p.cd)
        print(f"request {p.rid} ({p.levels} levels, {p.cells / 1e6:.1f}M cells, {p.n_samples} samples): "
              f"Cd plain FP32 {p.cd:.6f}, FP16 batched {b.cd:.6f}, rel {d:.2e}", flush=True)
        assert d <= 5e-3, f"request {p.rid}: FP16 batched Cd {b.cd} vs plain FP32 {p.cd}"
