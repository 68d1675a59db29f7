"""Shared helpers of the GPU parity tests: the seeded perturbed starting state,
a GPU/oracle pair stepped through the same batches, and the field comparison."""
import numpy as np

from openludwig_amd.solver import Solver, ramp_velocity
from oracle import Oracle

def _weights():
    k = np.arange(27)
    cx, cy, cz = k % 3 - 1, (k // 3) % 3 - 1, k // 9 - 1
    d2 = cx * cx + cy * cy + cz * cz
    w = np.where(d2 == 0, np.float32(8) / np.float32(27),
                 np.where(d2 == 1, np.float32(2) / np.float32(27),
                          np.where(d2 == 2, np.float32(1) / np.float32(54), np.float32(1) / np.float32(216))))
    return w.astype(np.float32)


def perturbed_state(grids, seed=1234, amp=0.01):
    """f = w_k (1 + amp U[-1,1]) on every cell of every level (float32)."""
    rng = np.random.default_rng(seed)
    w = _weights()
    out = []
    for g in grids:
        n4 = g.n_cells
        f = (w[:, None] * (1.0 + amp * rng.uniform(-1, 1, size=(27, n4)))).astype(np.float32)
        out.append(f.ravel())
    return out


def run_pair(cfg, grids, params, mesh, steps, strict, perturb=True, batch=4, storage=0, amp=0.01, **solver_kw):
    s = Solver(cfg, grids, params, mesh, strict_fp=strict, storage=storage, **solver_kw)
    o = Oracle(cfg, grids, params, "f32", storage=storage)
    s.init_equilibrium()
    o.init_equilibrium()
    if perturb:
        for g, f in zip(grids, perturbed_state(grids, amp=amp)):
            s.upload(g.level_id, "f", f)
            o.put(g.level_id, "f", f.astype(np.float64))
    t = 1
    while t <= steps:
        n = min(batch, steps - t + 1)
        u = ramp_velocity(cfg.u_lattice, t + n - 1, cfg.ramp_steps)
        s.execute_timestep_batch(t, n, u)
        o.step_batch(t, n, u)
        t += n
    s.synchronize()
    return s, o


def compare(s, o, grids, fields=("f", "rho", "vel")):
    res = {}
    for g in grids:
        for fld in fields:
            a = s.download(g.level_id, fld).astype(np.float64)
            b = o.get(g.level_id, fld)
            assert a.shape == b.shape
            d = np.abs(a - b)
            res[(g.level_id, fld)] = (float(d.max()), int(np.count_nonzero(d)),
                                      float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)))
            assert np.all(np.isfinite(a)), f"non-finite GPU values level {g.level_id} {fld}"
    return res
