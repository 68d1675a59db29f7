"""Pins of the oracle's multilevel pieces: the 2:1 recursion schedule, the
coarse-to-fine interpolation and the Bouzidi correction (CPU only)."""
import ctypes as C

import numpy as np
import pytest

from openludwig_amd import cases
from openludwig_amd.domain import BlockLevel
from openludwig_amd.solver import build_case, ramp_velocity
from oracle import Oracle, lib


def reference_schedule(n_levels, t):
    """recursive_step! / recursive_step_temporal! (solver_control.jl:6-197):
    level l at t_sub, then its children at (2 t_sub, tw 0.0) and (2 t_sub + 1, tw 0.5)."""
    out = []

    def rec(lvl, ts, tw):
        out.append((lvl, ts, tw))
        if lvl < n_levels:
            rec(lvl + 1, 2 * ts, 0.0)
            rec(lvl + 1, 2 * ts + 1, 0.5)

    rec(1, t, 0.0)
    return out


@pytest.mark.parametrize("levels,res", [(2, 8), (3, 16)])
def test_recursion_schedule(levels, res):
    cfg, mesh = cases.tiny_sphere(res, levels)
    grids, params, mesh = build_case(cfg, mesh)
    assert len(grids) == levels
    o = Oracle(cfg, grids, params, "f32")
    o.init_equilibrium()
    L = lib("f32")
    cap = 256
    lv = (C.c_int * cap)(); ts = (C.c_longlong * cap)(); tw = (C.c_double * cap)()
    L.orc_trace_set(lv, ts, tw, cap)
    try:
        o.step_batch(5, 2, 0.01)
        n = L.orc_trace_count()
    finally:
        L.orc_trace_set(None, None, None, 0)
    got = [(lv[i], ts[i], tw[i]) for i in range(n)]
    ref = reference_schedule(levels, 5) + reference_schedule(levels, 6)
    assert got == ref
    for lvl in range(1, levels + 1):
        assert sum(1 for g in got if g[0] == lvl) == 2 * 2 ** (lvl - 1)


def _uniform_parent(o, grids, rho, u, fvals):
    g1 = grids[0]
    n4 = g1.n_cells
    f = np.repeat(np.asarray(fvals, dtype=np.float64)[:, None], n4, axis=1).ravel()
    o.put(1, "f", f)
    o.put(1, "rho", np.full(n4, rho))
    o.put(1, "vel", np.repeat(np.asarray(u, dtype=np.float64)[:, None], n4, axis=1).ravel())


def test_interpolation_uniform_equilibrium_and_ramp():
    """interpolate_with_rescaling (physics_interpolation.jl:4-127): a uniform
    equilibrium parent interpolates to that equilibrium; a parent field linear
    in the parent index P reproduces it at p = (g - 0.5)/2 -- the half-cell
    shift of the reference -- with the non-equilibrium part scaled by
    clamp((tau_f - 0.5)/(tau_c - 0.5), 0.01, 100)."""
    L = lib("f64")
    cfg, mesh = cases.tiny_sphere(8, 2)
    grids, params, mesh = build_case(cfg, mesh)
    o = Oracle(cfg, grids, params, "f64")
    o.init_equilibrium()
    rho, u = 1.02, (0.03, -0.01, 0.005)
    feq = [L.orc_feq(0, rho, *u, k) for k in range(27)]
    _uniform_parent(o, grids, rho, u, feq)
    tf = grids[1].tau
    for (gx, gy, gz) in [(5, 5, 5), (9, 12, 20), (30, 17, 8)]:
        for k in (0, 4, 13, 22):
            v = L.orc_interp_probe(o.ctx, 0, 1, gx, gy, gz, k, tf, 1.0, 1, 0)
            assert abs(v - feq[k]) < 1e-15
    # linear ramp along x in the parent's own index: f_k(P) = a + b * P_x
    g1 = grids[0]
    P_x = ((g1.coords[:, 0] - 1) * 8)[:, None] + (np.arange(512) % 8)[None, :] + 1   # 1-based parent x
    a, b = 0.02, 1e-4
    fk = (a + b * P_x).ravel()
    f = np.tile(fk, 27)
    o.put(1, "f", f)
    o.put(1, "rho", np.ones(g1.n_cells))
    o.put(1, "vel", np.zeros(3 * g1.n_cells))
    tc = grids[0].tau
    scale = min(max((tf - 0.5) / (tc - 0.5), 0.01), 100.0)
    w = [L.orc_feq(0, 1.0, 0, 0, 0, k) for k in range(27)]
    for gx in (5, 6, 7, 8, 13, 20):
        for k in (0, 13, 26):
            v = L.orc_interp_probe(o.ctx, 0, 1, gx, 9, 9, k, tf, 1.0, 1, 0)
            p = (gx - 0.5) * 0.5
            assert abs(v - (w[k] + (a + b * p - w[k]) * scale)) < 1e-15


def _strip_bouzidi(g: BlockLevel) -> BlockLevel:
    import dataclasses
    return dataclasses.replace(g, bouzidi_cells=np.zeros((0, 4), np.int32),
                               bouzidi_q=np.zeros((0, 27), np.uint16), bouzidi_tri=np.zeros((0, 27), np.int32))


def test_bouzidi_slot_and_formula():
    """bouzidi_correction_kernel_fixed! (bouzidi_kernel.jl:13-92): after one
    step, slot (x, opp k) of every link with q_min < q <= 1 holds
    2q f*(x,k) + (1-2q) f*(x-c_k,k) for q < 1/2, else f*(x,k)/(2q) +
    (2q-1)/(2q) f*(x,opp k), f* being the uncorrected post-collision values
    (own value when x - c_k has no block); every other slot is untouched."""
    cfg, mesh = cases.tiny_sphere(8, 1)
    grids, params, mesh = build_case(cfg, mesh)
    assert len(grids) == 1 and grids[0].n_boundary_cells > 0
    g = grids[0]
    rng = np.random.default_rng(4)
    k = np.arange(27)
    w = np.where((k % 3 - 1) ** 2 + ((k // 3) % 3 - 1) ** 2 + (k // 9 - 1) ** 2 == 0, 8 / 27, 0)
    o1 = Oracle(cfg, grids, params, "f64")
    o0 = Oracle(cfg, [_strip_bouzidi(g)], params, "f64")
    f0 = np.tile(np.array([o1.L.orc_feq(0, 1.0, 0, 0, 0, kk) for kk in range(27)])[:, None], (1, g.n_cells))
    f0 = (f0 * (1 + 0.01 * rng.uniform(-1, 1, f0.shape))).ravel()
    for o in (o0, o1):
        o.init_equilibrium()
        o.put(1, "f", f0)
        o.step_batch(1, 1, 0.01)
    fu = o0.get(1, "f").reshape(27, g.n_blocks, 512)
    fc = o1.get(1, "f").reshape(27, g.n_blocks, 512)
    expect = fu.copy()
    nbt = g.neighbor_table
    half = np.frombuffer(g.bouzidi_q.astype(np.uint16).tobytes(), dtype=np.float16).reshape(-1, 27).astype(np.float64)
    n_links = 0
    for i, (b, x, y, z) in enumerate(g.bouzidi_cells):
        b0 = b - 1
        c = (x - 1) + 8 * (y - 1) + 64 * (z - 1)
        for kk in range(27):
            q = float(np.float32(half[i, kk]))
            if not (q > 0.001 and q <= 1.0):
                continue
            ok = 26 - kk
            cx, cy, cz = kk % 3 - 1, (kk // 3) % 3 - 1, kk // 9 - 1
            fk = fu[kk, b0, c]
            if q < 0.5:
                nx, ny, nz = x - cx, y - cy, z - cz
                if 1 <= nx <= 8 and 1 <= ny <= 8 and 1 <= nz <= 8:
                    fff = fu[kk, b0, (nx - 1) + 8 * (ny - 1) + 64 * (nz - 1)]
                else:
                    ox = -1 if nx < 1 else (1 if nx > 8 else 0)
                    oy = -1 if ny < 1 else (1 if ny > 8 else 0)
                    oz = -1 if nz < 1 else (1 if nz > 8 else 0)
                    nb = nbt[b0, (ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)]
                    if nb > 0:
                        fff = fu[kk, nb - 1, ((nx - 1) % 8) + 8 * ((ny - 1) % 8) + 64 * ((nz - 1) % 8)]
                    else:
                        fff = fk
                expect[ok, b0, c] = 2 * q * fk + (1 - 2 * q) * fff
            else:
                expect[ok, b0, c] = fk / (2 * q) + (2 * q - 1) / (2 * q) * fu[ok, b0, c]
            n_links += 1
    assert n_links > 100
    assert np.abs(fc - expect).max() < 1e-15


def test_fp32_mirror_tracks_fp64_oracle():
    """The FP32 mirror and the FP64 oracle run the same algorithm; their
    difference after a few steps is FP32 rounding, not an algorithmic term."""
    cfg, mesh = cases.tiny_sphere(8, 2)
    grids, params, mesh = build_case(cfg, mesh)
    o32 = Oracle(cfg, grids, params, "f32")
    o64 = Oracle(cfg, grids, params, "f64")
    for o in (o32, o64):
        o.init_equilibrium()
        for t in (1, 5, 9):
            o.step_batch(t, 4, ramp_velocity(cfg.u_lattice, t + 3, cfg.ramp_steps))
    for g in grids:
        a, b = o32.get(g.level_id, "f"), o64.get(g.level_id, "f")
        assert np.abs(a - b).max() < 2e-6
        ra, rb = o32.get(g.level_id, "rho"), o64.get(g.level_id, "rho")
        assert np.abs(ra - rb).max() < 2e-6
