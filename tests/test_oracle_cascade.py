"""Pins of the oracle's pull and boundary cascade, per-order relaxation and
cumulant rebuild, wall model and thermal step (CPU).

Cascade (kernel_cumulant.jl:92-159): on a box of whole blocks,
  * rest is a fixed point with the inflow at u = 0: inlet and outlet inject
    w_k, the y/z walls reflect it (up to the 0.333333333 literal of cs2);
  * one step from a random state equals a numpy restatement of the pull --
    in-block and neighbour-block sources, inlet feq(1, u_in), outlet
    feq(1, u_in), y walls f[mirror_y k] of the cell itself, then z walls,
    in that precedence at edges and corners -- followed by the pinned
    single-cell collision;
  * the y/z walls are specular: with a perturbation that stays clear of the
    inlet and outlet, total mass and x-momentum are conserved, which a swap
    of mirror_y/mirror_z (mass) or a bounce-back (momentum) breaks.
Relaxation (kernel_cumulant.jl:441-496): at u = 0 with distinct rates for
bulk, shear, 3rd order, the diagonal and mixed 4th-order groups, every
cumulant of the output is the rate-scaled cumulant of the input index by
index, and the 5th/6th-order central moments are those of zero 5th/6th
cumulants, computed by a generic set-partition expansion (independent of
the reference's hand-expanded formulas).
Wall model (kernel_cumulant.jl:216-254, 604-623): the momentum change of a
near-wall cell is (3 - omega_1)/2 F, F from the closed-form power-law u_tau
below y+ = 11.81 and its one log-law correction above.
Thermal (kernel_thermal.jl:14-202): uniform T with uniform u along x is a
fixed point; the BGK step and the mirror walls conserve sum g; solid cells of
an isothermal wall take T_wall and return 2 w_k T_wall - g[opp k].
"""
import itertools
import math

import numpy as np
import pytest

from oracle import lib
from oracle_box import CX, CY, CZ, MIRROR_Y, MIRROR_Z, OPP, Box, feq, params, weights

W = weights()
CVEC = np.stack([CX, CY, CZ], 1).astype(np.float64)


def collide1(L, fin, tau, prm, u_in, sponge=0.0, wall=100.0):
    import ctypes as C
    fin = np.ascontiguousarray(fin, dtype=np.float64)
    un = np.zeros(3)
    out = np.zeros(27)
    mac = np.zeros(4)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    L.orc_collide_single(dp(fin), dp(un), C.cast(None, C.POINTER(C.c_double)), tau, C.byref(prm), sponge, wall,
                         u_in, dp(out), dp(mac))
    return out, mac


def random_state(shape, rng, amp=0.02):
    return W[:, None, None, None] * (1.0 + amp * rng.uniform(-1, 1, (27,) + shape))


# ---------------------------------------------------------------------------
# pull and boundary cascade

def test_rest_box_is_fixed_point_with_zero_inflow():
    b = Box(2, 1, 1, tau=0.55)
    b.init()
    b.step(params(), 0.0, 20)
    f = b.get("f")
    assert np.abs(f - W[:, None, None, None]).max() < 1e-8
    assert np.abs(b.get("vel")).max() < 1e-9
    assert np.abs(b.get("rho") - 1.0).max() < 1e-8


def restated_pull(f, u_in, L):
    """numpy restatement of kernel_cumulant.jl:92-159 on a full box."""
    _, nx, ny, nz = f.shape
    gx, gy, gz = np.meshgrid(np.arange(1, nx + 1), np.arange(1, ny + 1), np.arange(1, nz + 1), indexing="ij")
    out = np.empty_like(f)
    for k in range(27):
        sx, sy, sz = gx - CX[k], gy - CY[k], gz - CZ[k]
        inside = (sx >= 1) & (sx <= nx) & (sy >= 1) & (sy <= ny) & (sz >= 1) & (sz <= nz)
        v = np.zeros((nx, ny, nz))
        v[inside] = f[k][sx[inside] - 1, sy[inside] - 1, sz[inside] - 1]
        inl, outl = sx < 1, sx > nx
        ywall = ~inl & ~outl & ((sy < 1) | (sy > ny))
        zwall = ~inl & ~outl & ~ywall & ((sz < 1) | (sz > nz))
        v[inl] = feq(L, 1.0, (u_in, 0, 0), k)
        v[outl] = feq(L, 1.0, (u_in, 0, 0), k)
        v[ywall] = f[MIRROR_Y[k]][ywall]
        v[zwall] = f[MIRROR_Z[k]][zwall]
        out[k] = v
    return out


def test_one_step_equals_restated_pull_and_collision():
    L = lib("f64")
    rng = np.random.default_rng(21)
    b = Box(2, 1, 1, tau=0.56)
    b.init()
    f0 = random_state(b.shape, rng)
    b.put("f", f0)
    prm = params(omega_bulk=0.9, omega_3=0.8, omega_4=0.7)
    u_in = float(np.float32(0.05))          # the inflow velocity crosses the ABI as Float32
    b.step(prm, u_in, 1)
    got = b.get("f")
    pulled = restated_pull(f0, u_in, L)
    # the inlet and outlet slots really are the u_in equilibrium, not w
    assert abs(pulled[14 + 0, 0, 3, 3] - W[14]) > 1e-3       # k = 14: c = (+1, 0, 0), pulled through x-min
    nx, ny, nz = b.shape
    worst = 0.0
    for i, j, k in itertools.product(range(nx), range(ny), range(nz)):
        exp, _ = collide1(L, pulled[:, i, j, k], 0.56, prm, u_in)
        worst = max(worst, float(np.abs(exp - got[:, i, j, k]).max()))
    assert worst == 0.0, worst


@pytest.mark.parametrize("wall", ["y", "z"])
def test_mirror_walls_conserve_mass_and_tangential_momentum(wall):
    """A perturbation that touches the walls normal to one axis only: clear
    of the inlet and outlet and of the y-z edges, where the reference's
    precedence (y mirror first, then z) makes the reflection non-bijective."""
    rng = np.random.default_rng(4)
    dims = (3, 2, 1) if wall == "z" else (3, 1, 2)
    prm = params(omega_bulk=0.7, omega_3=1.3, omega_4=0.9)
    # the rest state drifts by ~1e-10 per step where the inflow meets it (the
    # 0.333333333 literal of cs2); a run from rest measures that drift, which
    # the perturbed run shares exactly as long as the perturbation stays away
    rest = Box(*dims, tau=0.6)
    rest.init()
    rest.step(prm, 0.0, 3)
    fr = rest.get("f")
    b = Box(*dims, tau=0.6)
    b.init()
    f0 = np.broadcast_to(W[:, None, None, None], (27,) + b.shape).copy()
    # x = 10..15 and the middle 4 cells of the axis whose walls stay untouched
    sel = (slice(None), slice(9, 15), slice(6, 10), slice(None)) if wall == "z" else \
          (slice(None), slice(9, 15), slice(None), slice(6, 10))
    shp = f0[sel].shape
    f0[sel] += W[:, None, None, None] * 0.03 * rng.uniform(-1, 1, shp)
    f0[sel] += 0.002 * W[:, None, None, None] * (CX + CY + CZ)[:, None, None, None]   # net momentum
    b.put("f", f0)
    w4 = np.broadcast_to(W[:, None, None, None], f0.shape)
    b.step(prm, 0.0, 3)
    f = b.get("f")
    assert np.abs(f[:, :4] - fr[:, :4]).max() == 0.0          # the inlet region is untouched
    m0 = f0.sum()
    dm = (f.sum() - f0.sum()) - (fr.sum() - w4.sum())
    assert abs(dm) < 1e-13 * m0, dm
    normal = 2 if wall == "z" else 1
    for a, c in enumerate((CX, CY, CZ)):
        c4 = c[:, None, None, None]
        dj = ((f * c4).sum() - (f0 * c4).sum()) - ((fr * c4).sum() - (w4 * c4).sum())
        if a == normal:
            assert abs(dj) > 1e-6                     # reversed at the walls
        else:
            assert abs(dj) < 1e-13 * m0, (a, dj)
    # the perturbation did reach the walls
    edge = f[:, 9:15, 6:10, 0] if wall == "z" else f[:, 9:15, 0, 6:10]
    assert np.abs(edge - W[:, None, None]).max() > 1e-5


# ---------------------------------------------------------------------------
# relaxation per index and the 5th/6th-order rebuild

def _partitions(items):
    if not items:
        yield []
        return
    first, rest = items[0], items[1:]
    for p in _partitions(rest):
        for i in range(len(p)):
            yield p[:i] + [[first] + p[i]] + p[i + 1:]
        yield [[first]] + p


def moment_from_cumulants(abc, kappa):
    """Central moment m_abc from cumulants kappa[(a,b,c)] (first-order
    cumulants zero): sum over set partitions of the product of block
    cumulants (the multivariate moment-cumulant relation)."""
    items = ["x"] * abc[0] + ["y"] * abc[1] + ["z"] * abc[2]
    tot = 0.0
    for p in _partitions(items):
        prod = 1.0
        for blk in p:
            idx = (blk.count("x"), blk.count("y"), blk.count("z"))
            prod *= kappa.get(idx, 0.0)
            if prod == 0.0:
                break
        tot += prod
    return tot


def cumulants_from_moments(m):
    """Cumulants up to 4th order from normalised central moments m[(a,b,c)]."""
    kap = {}
    for order in (2, 3, 4):
        for abc in itertools.product(range(3), repeat=3):
            if sum(abc) != order:
                continue
            rest = moment_from_cumulants(abc, kap)     # partitions into lower-order blocks only
            kap[abc] = m[abc] - rest
    return kap


def _moments(f):
    rho = f.sum()
    u = (f[:, None] * CVEC).sum(0) / rho
    m = {}
    for abc in itertools.product(range(3), repeat=3):
        m[abc] = float(np.sum(f * (CX - u[0]) ** abc[0] * (CY - u[1]) ** abc[1] * (CZ - u[2]) ** abc[2]) / rho)
    return rho, u, m


@pytest.mark.parametrize("adaptive", [0, 1])
def test_per_index_relaxation_and_rebuild(adaptive):
    L = lib("f64")
    rng = np.random.default_rng(17)
    f = W * (1 + 0.04 * rng.uniform(-1, 1, 27))
    # remove the momentum so that u = 0 (chimera around rest)
    j = (f[:, None] * CVEC).sum(0)
    f -= (W[:, None] * CVEC * 3.0) @ j
    rho, u, m = _moments(f)
    assert np.abs(u).max() < 1e-16
    # the rates cross the ABI as Float32
    tau, wb, w3, w4 = (float(np.float32(v)) for v in (0.57, 0.83, 1.31, 0.61))
    w1 = 1.0 / tau
    prm = params(omega_bulk=wb, omega_3=w3, omega_4=w4, adaptive_omega4=adaptive, lambda_param=0.1667)
    out, mac = collide1(L, f, tau, prm, 0.0)
    rho2, u2, m2 = _moments(out)
    assert abs(rho2 - rho) < 1e-15 and np.abs(u2).max() < 1e-15
    kin = cumulants_from_moments(m)
    lam = float(np.float32(0.1667))
    w4d = min(max(-(w1 - 2) ** 2 / (lam * w1 + 1e-10), 0.01), 2.0) if adaptive else w4
    exp = {}
    cs2 = 0.333333333
    tr = (1 - wb) * (kin[2, 0, 0] + kin[0, 2, 0] + kin[0, 0, 2]) + wb * 3 * cs2
    d1 = (1 - w1) * (kin[2, 0, 0] - kin[0, 2, 0])
    d2 = (1 - w1) * (kin[2, 0, 0] - kin[0, 0, 2])
    exp[2, 0, 0] = (tr + d1 + d2) / 3
    exp[0, 2, 0] = (tr - 2 * d1 + d2) / 3
    exp[0, 0, 2] = (tr + d1 - 2 * d2) / 3
    for abc in ((1, 1, 0), (1, 0, 1), (0, 1, 1)):
        exp[abc] = (1 - w1) * kin[abc]
    for abc in ((2, 1, 0), (1, 2, 0), (2, 0, 1), (0, 2, 1), (1, 0, 2), (0, 1, 2), (1, 1, 1)):
        exp[abc] = (1 - w3) * kin[abc]
    for abc in ((2, 2, 0), (2, 0, 2), (0, 2, 2)):
        exp[abc] = (1 - w4d) * kin[abc]
    for abc in ((2, 1, 1), (1, 2, 1), (1, 1, 2)):
        exp[abc] = (1 - w4) * kin[abc]
    kout = cumulants_from_moments(m2)
    for abc, v in exp.items():
        assert abs(kout[abc] - v) < 1e-14, (abc, kout[abc], v)
        if sum(abc) >= 3:
            assert abs(kin[abc]) > 1e-6   # the input really carries each cumulant
    # 5th and 6th order: the central moments of zero 5th/6th cumulants
    for abc in ((2, 2, 1), (2, 1, 2), (1, 2, 2), (2, 2, 2)):
        assert abs(m2[abc] - moment_from_cumulants(abc, exp)) < 1e-14, abc


# ---------------------------------------------------------------------------
# wall model

def _wall_force(rho, u, d, tau):
    """Closed form of kernel_cumulant.jl:216-250 in FP64."""
    um = math.sqrt(u @ u)
    nu = (tau - 0.5) / 3.0
    ut = um * (nu / (d * um + 1e-10)) ** (1.0 / 7.0) * (2.0 * 8.3) ** (-1.0 / 7.0)
    ut = max(ut, 1e-6)
    yp = ut * d / nu
    if yp > 11.81:
        up = (1.0 / 0.41) * math.log(yp) + 5.2
        if up > 0.1:
            ut = max(ut * ((um / ut) / up), 1e-6)
    tw, tr = rho * ut * ut, rho * nu * (um / d)
    fm = (tw - tr) / d if tw > tr else 0.0
    return -fm * u / um, yp


@pytest.mark.parametrize("tau,d,branch", [(0.51, 2.0, "power"), (0.52, 3.0, "power"), (0.505, 0.5, "power"),
                                          (0.501, 3.0, "log"), (0.5005, 2.0, "log"), (0.5005, 9.5, "log"),
                                          (0.51, 3.0, "none")])
def test_wall_model_force_magnitude(tau, d, branch):
    L = lib("f64")
    rho, u = 1.02, np.array([0.04, 0.01, -0.005])
    f = np.array([feq(L, rho, u, k) for k in range(27)])
    tau32 = float(np.float32(tau))
    F, yp = _wall_force(rho, u, float(np.float32(d)), tau32)
    assert (yp <= 11.81) == (branch == "power") or branch == "none"
    assert (np.abs(F).max() > 1e-7) == (branch != "none")
    prm = params(wall_model=1, omega_bulk=1.0, omega_3=1.0, omega_4=1.0)
    out, _ = collide1(L, f, tau32, prm, 0.0, wall=d)
    far, _ = collide1(L, f, tau32, prm, 0.0, wall=100.0)
    dj = (out[:, None] * CVEC).sum(0) - (far[:, None] * CVEC).sum(0)
    w1 = 1.0 / tau32
    # momentum sums of O(0.04) round at ~1e-17; a wrong factor moves F by percents
    assert np.abs(dj - 0.5 * (3.0 - w1) * F).max() < 1e-8 * np.abs(F).max() + 1e-15, (dj, F, yp)


# ---------------------------------------------------------------------------
# thermal step

def _thermal_box(**kw):
    return Box(2, 1, 1, tau=0.6, tau_g=0.7, thermal=True, **kw)


def test_thermal_uniform_state_is_fixed_point():
    L = lib("f64")
    b = _thermal_box()
    T0, u0 = float(np.float32(1.07)), float(np.float32(0.03))     # both cross the ABI as Float32
    b.init(T0)
    # flow and temperature at the inflow state everywhere
    f = np.array([feq(L, 1.0, (u0, 0, 0), k) for k in range(27)])
    b.put("f", np.broadcast_to(f[:, None, None, None], (27,) + b.shape))
    b.put("vel", np.stack([np.full(b.shape, u0), np.zeros(b.shape), np.zeros(b.shape)]))
    g = W * T0 * (1.0 + 3.0 * CX * u0)    # calculate_thermal_equilibrium (physics_utils.jl:78-82)
    b.put("g", np.broadcast_to(g[:, None, None, None], (27,) + b.shape))
    prm = params(ddf=1, T_inlet=T0)
    b.step(prm, u0, 4)
    assert np.abs(b.get("T") - T0).max() < 1e-8
    assert np.abs(b.get("g") - g[:, None, None, None]).max() < 1e-8


@pytest.mark.parametrize("wall", ["y", "z"])
def test_thermal_bgk_and_walls_conserve_sum_g(wall):
    """sum g over the box is conserved by the BGK step and the mirror walls
    (a perturbation of g that reaches the walls of one axis, clear of the
    inflow, the outflow and the y-z edges; the rest-state drift of a run
    from rest is subtracted as in the flow case)."""
    rng = np.random.default_rng(8)
    dims = (3, 2, 1) if wall == "z" else (3, 1, 2)
    prm = params(ddf=1, T_inlet=1.0)
    rest = Box(*dims, tau=0.6, tau_g=0.65, thermal=True)
    rest.init(1.0)
    rest.step(prm, 0.0, 3)
    gr = rest.get("g")
    b = Box(*dims, tau=0.6, tau_g=0.65, thermal=True)
    b.init(1.0)
    g0 = np.broadcast_to(W[:, None, None, None], (27,) + b.shape).copy()
    sel = (slice(None), slice(9, 15), slice(6, 10), slice(None)) if wall == "z" else \
          (slice(None), slice(9, 15), slice(None), slice(6, 10))
    g0[sel] *= 1.0 + 0.05 * rng.uniform(-1, 1, g0[sel].shape)
    b.put("g", g0)
    b.step(prm, 0.0, 3)
    g = b.get("g")
    ds = (g.sum() - g0.sum()) - (gr.sum() - np.broadcast_to(W[:, None, None, None], g.shape).sum())
    assert abs(ds) < 1e-13 * g0.sum(), ds
    assert np.abs(g - g0).max() > 1e-4
    edge = g[:, 9:15, 6:10, 0] if wall == "z" else g[:, 9:15, 0, 6:10]
    assert np.abs(edge - W[:, None, None]).max() > 1e-6


def test_isothermal_wall_cells():
    T0, Tw = 1.0, 1.25
    solid = np.zeros((16, 8, 8), dtype=bool)
    solid[6:10, 2:6, 2:6] = True
    b = Box(2, 1, 1, tau=0.6, tau_g=0.7, thermal=True, obstacle=solid)
    b.init(T0)
    prm = params(ddf=1, T_inlet=T0, T_wall=Tw, wall_bc_adiabatic=0)
    b.step(prm, 0.0, 1)
    T = b.get("T")[0]
    g = b.get("g")
    assert np.all(T[solid] == Tw)
    # solid cells return 2 w_k T_wall - g_in[opp k] (g_in = w T0 at rest)
    exp = 2 * W * Tw - W[OPP] * T0
    assert np.abs(g[:, solid] - exp[:, None]).max() < 1e-15
    b.step(prm, 0.0, 20)
    T = b.get("T")[0]
    near = np.zeros_like(solid)
    near[5, 3:5, 3:5] = True                 # fluid cells facing the solid block
    assert np.all(T[near] > T0 + 1e-3)
    # adiabatic: the solid cell keeps the sum of what it pulls (T0 at rest, up
    # to the ~1e-10 velocities the cs2 literal leaves in the rest state)
    b2 = Box(2, 1, 1, tau=0.6, tau_g=0.7, thermal=True, obstacle=solid)
    b2.init(T0)
    b2.step(params(ddf=1, T_inlet=T0, T_wall=Tw, wall_bc_adiabatic=1), 0.0, 3)
    assert np.abs(b2.get("T")[0] - T0).max() < 1e-8


# ---------------------------------------------------------------------------
# compute_stress_from_cell (surface.jl:143-190) on a linear shear layer

def test_wall_stress_of_linear_shear_layer():
    """A flat wall at y = 4 (cells below solid) under u = (g (y - 4),
    b (y - 4), 0) and uniform density rho: the probe one and a half cells
    out (the first with full fluid support, surface.jl:224-346) sees
    u_t = (1.5 g, 0, 0) at wall distance 1.5, so the stress is
    tau_x = rho nu_mol g * stress_scale (the normal part b is removed),
    and p = (rho - 1)/3 * pressure_scale, with forces -p n A and tau A."""
    import ctypes as C
    L = lib("f64")
    solid = np.zeros((16, 16, 16), dtype=bool)
    solid[:, :4, :] = True
    tau = float(np.float32(0.53))
    b = Box(2, 2, 2, tau=tau, obstacle=solid)
    b.init()
    gam, bet, rho = 0.004, 0.001, 1.03
    yc = np.arange(16) + 0.5
    h = np.where(yc > 4.0, yc - 4.0, 0.0)[None, :, None] * np.ones((16, 16, 16))
    b.put("vel", np.stack([gam * h, bet * h, np.zeros_like(h)]))
    b.put("rho", np.full((1, 16, 16, 16), rho))
    # probe points (triangle centres) on the wall, away from block faces
    xs = np.array([5.3, 8.7, 10.1, 6.45], dtype=np.float32)
    zs = np.array([6.2, 9.9, 4.8, 11.3], dtype=np.float32)
    n = len(xs)
    cx, cy, cz = xs, np.full(n, 4.0, np.float32), zs
    nx, ny, nz = np.zeros(n, np.float32), np.ones(n, np.float32), np.zeros(n, np.float32)
    area = np.array([1.0, 0.5, 2.0, 0.25], dtype=np.float32)
    off = np.zeros(3, np.float32)
    ref = np.array([8.0, 0.0, 8.0], np.float32)
    ps, ss = 1.7, 2.3
    out = np.zeros(9)
    maps = np.zeros(4 * n, np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    L.orc_forces(b.ctx, 0, fp(cx), fp(cy), fp(cz), fp(nx), fp(ny), fp(nz), fp(area), n, fp(off), fp(ref), 1.0,
                 ps, ss, 5, 0.0, 0.0, out.ctypes.data_as(C.POINTER(C.c_double)), fp(maps))
    p, sx, sy, sz = maps.reshape(4, n).astype(np.float64)
    nu = (tau - 0.5) / 3.0
    assert np.allclose(p, (rho - 1.0) / 3.0 * ps, rtol=1e-5)
    assert np.allclose(sx, rho * nu * gam * ss, rtol=1e-5)
    assert np.abs(sy).max() < 1e-9 and np.abs(sz).max() < 1e-9
    A = area.astype(np.float64)
    assert abs(out[1] - (-(rho - 1.0) / 3.0 * ps * A.sum())) < 1e-5 * abs(out[1])      # Fy pressure = -p n_y A
    assert abs(out[3] - rho * nu * gam * ss * A.sum()) < 1e-5 * abs(out[3])              # Fx viscous
    assert out[0] == 0.0 and out[2] == 0.0 and out[5] == 0.0
    # moments about the reference point: M = sum r x dF, dF = (tau_x A, -p A, 0), r = (x - 8, 4, z - 8)
    pA = (rho - 1.0) / 3.0 * ps * A
    tA = rho * nu * gam * ss * A
    rx, rz = cx.astype(np.float64) - 8.0, cz.astype(np.float64) - 8.0
    for got, exp in ((out[6], np.sum(rz * pA)), (out[7], np.sum(rz * tA)), (out[8], np.sum(-rx * pA - 4.0 * tA))):
        assert abs(got - exp) <= 1e-5 * np.sum(np.abs(pA) + np.abs(tA)) * 8, (got, exp)
