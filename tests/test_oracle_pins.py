"""Pins of the CPU oracle (oracle/lb_oracle.c) against what the mathematics of
the reference's operators fixes.  The reference ships no tests or golden
vectors (src/ has no test/ directory); each check here is chosen so that a
plausible restatement mistake -- a dropped term, a wrong sign or index, a
transposed operand -- fails it.  All run on the CPU.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import OrcParams, lib

K = np.arange(27)
CXs, CYs, CZs = K % 3 - 1, (K // 3) % 3 - 1, K // 9 - 1
CVEC = np.stack([CXs, CYs, CZs], 1).astype(np.float64)


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def params(**kw):
    p = OrcParams()
    p.c_wale, p.nu_sgs_bg, p.inlet_turbulence = 0.0, 0.0, 0.0
    p.symmetric = p.wall_model = p.sponge_blend = 0
    p.temporal = 1
    p.omega_bulk = p.omega_3 = p.omega_4 = 1.0
    p.lambda_param = 1.0 / 6.0
    p.adaptive_omega4 = 0
    p.limiter = 0
    p.compressible = 0
    p.q_min = 0.001
    p.ddf = 0
    p.T_inlet = p.T_wall = 1.0
    p.wall_bc_adiabatic = 1
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def weights(L):
    cx = (C.c_int * 27)(); cy = (C.c_int * 27)(); cz = (C.c_int * 27)()
    w = (C.c_float * 27)() if L is lib("f32") else (C.c_double * 27)()
    opp = (C.c_int * 27)(); my = (C.c_int * 27)(); mz = (C.c_int * 27)()
    L.orc_lattice(cx, cy, cz, w, opp, my, mz)
    return (np.array(cx[:]), np.array(cy[:]), np.array(cz[:]), np.array(w[:]), np.array(opp[:]),
            np.array(my[:]), np.array(mz[:]))


def collide(L, fin, tau, prm, u_nb=(0, 0, 0), grad=None, sponge=0.0, wall=100.0, u_in=0.05):
    fin = np.ascontiguousarray(fin, dtype=np.float64)
    un = np.ascontiguousarray(u_nb, dtype=np.float64)
    g = None if grad is None else np.ascontiguousarray(np.asarray(grad, dtype=np.float64).ravel())
    out = np.zeros(27)
    mac = np.zeros(4)
    L.orc_collide_single(dp(fin), dp(un), dp(g) if g is not None else C.cast(None, C.POINTER(C.c_double)),
                         tau, C.byref(prm), sponge, wall, u_in, dp(out), dp(mac))
    return out, mac


# ---------------------------------------------------------------------------
# lattice (lattice.jl:5-56)

def test_lattice_tables_closed_form():
    cx, cy, cz, w, opp, my, mz = weights(lib("f64"))
    assert (cx == CXs).all() and (cy == CYs).all() and (cz == CZs).all()
    assert (opp == 26 - K).all() and (opp[opp] == K).all()
    assert (cx[my] == cx).all() and (cy[my] == -cy).all() and (cz[my] == cz).all()
    assert (cz[mz] == -cz).all() and (cy[mz] == cy).all()
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.abs((w[:, None] * CVEC).sum(0)).max() < 1e-16
    m2 = np.einsum("k,ka,kb->ab", w, CVEC, CVEC)
    assert np.abs(m2 - np.eye(3) / 3).max() < 1e-15
    m4 = np.einsum("k,ka,kb,kc,kd->abcd", w, CVEC, CVEC, CVEC, CVEC)
    d = np.eye(3)
    iso = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d)) / 9
    assert np.abs(m4 - iso).max() < 1e-15
    # FP32 mirror uses the Float32 literals 8f/27f ... (timestep.jl:168)
    w32 = weights(lib("f32"))[3]
    assert w32[13] == np.float32(8) / np.float32(27)
    assert w32[0] == np.float32(1) / np.float32(216)


# ---------------------------------------------------------------------------
# equilibria (physics_utils.jl:19-72)

@pytest.mark.parametrize("comp", [0, 1])
def test_equilibrium_moments(comp):
    L = lib("f64")
    rng = np.random.default_rng(3)
    for _ in range(5):
        rho = 1 + 0.1 * rng.uniform(-1, 1)
        u = 0.1 * rng.uniform(-1, 1, 3)
        f = np.array([L.orc_feq(comp, rho, *u, k) for k in range(27)])
        assert abs(f.sum() - rho) < 1e-14
        assert np.abs((f[:, None] * CVEC).sum(0) - rho * u).max() < 1e-14
        if comp == 0:
            m2 = np.einsum("k,ka,kb->ab", f, CVEC, CVEC)
            assert np.abs(m2 - rho * (np.outer(u, u) + np.eye(3) / 3)).max() < 1e-14
    # the O(u^4) terms vanish at rest
    for k in range(27):
        assert L.orc_feq(1, 1.0, 0, 0, 0, k) == L.orc_feq(0, 1.0, 0, 0, 0, k)
    # and for a non-zero u they are the Hermite 3rd/4th order terms
    u = np.array([0.05, -0.02, 0.01])
    for k in range(27):
        cu, usq = CVEC[k] @ u, u @ u
        extra = 4.5 * cu ** 3 - 4.5 * cu * usq + 3.375 * cu ** 4 - 6.75 * cu ** 2 * usq + 1.125 * usq ** 2
        w = L.orc_feq(0, 1.0, 0, 0, 0, k)
        assert abs(L.orc_feq(1, 1.0, *u, k) - L.orc_feq(0, 1.0, *u, k) - w * extra) < 1e-16


# ---------------------------------------------------------------------------
# chimera transforms (kernel_cumulant.jl:4-26, 330-371, 525-559)

def test_chimera_roundtrip_and_central_moments():
    L = lib("f64")
    rng = np.random.default_rng(7)
    f = rng.uniform(0.01, 0.3, 27)
    u = rng.uniform(-0.1, 0.1, 3)
    out = np.zeros(27)
    L.orc_chimera_roundtrip(dp(f), *u, dp(out))
    assert np.abs(out - f).max() < 1e-15
    # forward = central moments: A[a + 3b + 9c] = sum f (cx-u)^a (cy-v)^b (cz-w)^c
    km = np.zeros(27)
    L.orc_chimera_forward(dp(f), *u, dp(km))
    for a in range(3):
        for b in range(3):
            for c in range(3):
                ref = np.sum(f * (CXs - u[0]) ** a * (CYs - u[1]) ** b * (CZs - u[2]) ** c)
                assert abs(km[a + 3 * b + 9 * c] - ref) < 1e-14


# ---------------------------------------------------------------------------
# collision (kernel_cumulant.jl:190-623)

def test_conservation_mass_momentum():
    L = lib("f64")
    rng = np.random.default_rng(11)
    w = weights(L)[3]
    for _ in range(5):
        f = w * (1 + 0.05 * rng.uniform(-1, 1, 27))
        out, mac = collide(L, f, 0.6, params(omega_bulk=0.8, omega_3=0.9, omega_4=0.7))
        assert abs(out.sum() - f.sum()) < 1e-15
        assert np.abs((out[:, None] * CVEC).sum(0) - (f[:, None] * CVEC).sum(0)).max() < 1e-15
        assert abs(mac[0] - f.sum()) < 1e-15


def test_all_rates_one_gives_factorized_equilibrium():
    """With omega_1 = omega_bulk = omega_3 = omega_4 = 1 every non-conserved
    cumulant takes its equilibrium: second-order (trace 3*0.333333333,
    deviators and off-diagonals 0), higher cumulants 0.  The output is then
    rho * phi(cx;ux) phi(cy;uy) phi(cz;uz), phi(+-1) = (k2 + u^2 +- u)/2,
    phi(0) = 1 - k2 - u^2, k2 = 0.333333333 (kernel_cumulant.jl:420)."""
    L = lib("f64")
    rng = np.random.default_rng(5)
    w = weights(L)[3]
    f = w * (1 + 0.05 * rng.uniform(-1, 1, 27))
    out, mac = collide(L, f, 1.0, params())    # tau = 1 -> omega_1 = 1 (no WALE: uniform u_nb)
    rho = f.sum()
    u = (f[:, None] * CVEC).sum(0) / rho
    k2 = 0.333333333

    def phi(c, ua):
        return np.where(c == 0, 1 - k2 - ua * ua, (k2 + ua * ua + c * ua) / 2)

    ref = rho * phi(CXs, u[0]) * phi(CYs, u[1]) * phi(CZs, u[2])
    assert np.abs(out - ref).max() < 1e-14
    assert abs(mac[0] - rho) < 1e-15 and np.abs(mac[1:] - u).max() < 1e-15


def test_rest_state_near_fixed_point():
    L = lib("f64")
    w = weights(L)[3]
    out, mac = collide(L, w, 0.51, params())
    # only the cs2 = 0.333333333 literal moves the rest state, by ~1e-9
    assert np.abs(out - w).max() < 1e-9
    assert abs(out.sum() - 1.0) < 1e-15


def _omega1_from_shear(L, tau, grad, c_wale, nu_bg=0.0):
    """Feed a state with a known off-diagonal central moment at u = 0 and read
    the relaxation factor (1 - omega_1) off the post-collision one."""
    w = weights(L)[3]
    f = w.copy()
    f += 1e-3 * w * (CXs * CYs)                     # adds a k110 moment, keeps mass/momentum
    prm = params(c_wale=c_wale, nu_sgs_bg=nu_bg)
    out, _ = collide(L, f, tau, prm, grad=grad)
    kin = np.zeros(27); kout = np.zeros(27)
    L.orc_chimera_forward(dp(f), 0.0, 0.0, 0.0, dp(kin))
    L.orc_chimera_forward(dp(out), 0.0, 0.0, 0.0, dp(kout))
    return 1.0 - kout[4] / kin[4]


def wale_nu(g, c_w):
    gsq = g @ g
    sd = 0.5 * (gsq + gsq.T) - np.trace(gsq) / 3 * np.eye(3)
    s = 0.5 * (g + g.T)
    op1 = (sd * sd).sum()
    op2 = g[0, 0] ** 2 + g[1, 1] ** 2 + g[2, 2] ** 2 + 2 * (s[0, 1] ** 2 + s[0, 2] ** 2 + s[1, 2] ** 2)
    if op1 <= 1e-12:
        return 0.0
    return c_w * c_w * op1 ** 1.5 / (op2 ** 2.5 + op1 ** 1.25)


@pytest.mark.parametrize("grad", [
    np.zeros((3, 3)),
    np.array([[0, 0.01, 0], [0, 0, 0], [0, 0, 0]]),            # pure shear: WALE gives 0
    np.array([[0.004, 0, 0], [0, -0.003, 0], [0, 0, -0.001]]),   # strain
    np.array([[0.002, 0.003, -0.001], [0.001, -0.002, 0.004], [0.0, 0.002, 0.0]]),
])
def test_wale_effective_relaxation(grad):
    """omega_1 = 1 / max(0.5 + 3(nu_mol + max(nu_t, nu_bg)), 0.500001) with the
    WALE nu_t of the central-difference gradient of the previous velocity
    (kernel_cumulant.jl:264-310, physics_utils.jl:112-123)."""
    L = lib("f64")
    tau, c_w = float(np.float32(0.52)), 0.5      # tau crosses the ABI as Float32
    om = _omega1_from_shear(L, tau, grad, c_w)
    nu = (tau - 0.5) / 3 + wale_nu(np.asarray(grad, dtype=np.float64), c_w)
    assert abs(om - 1.0 / max(0.5 + 3 * nu, 0.500001)) < 1e-9


def test_adaptive_omega4_is_always_001():
    """A = -(omega_1 - 2)^2 / (lambda omega_1 + 1e-10) <= 0, so the clamp to
    [0.01, 2] yields 0.01 (kernel_cumulant.jl:317-325): adaptive on equals the
    diagonal-4th-order rate fixed at 0.01 for every tau and lambda."""
    L = lib("f64")
    rng = np.random.default_rng(2)
    w = weights(L)[3]
    f = w * (1 + 0.05 * rng.uniform(-1, 1, 27))
    for tau in (0.501, 0.55, 0.9):
        for lam in (0.1667, 1.0, 3.0):
            a, _ = collide(L, f, tau, params(adaptive_omega4=1, lambda_param=lam, omega_4=1.0))
            # emulate: diag group 0.01, mixed group 1.0 -> differs from uniform 1.0 and from uniform 0.01
            b, _ = collide(L, f, tau, params(adaptive_omega4=0, omega_4=1.0))
            c, _ = collide(L, f, tau, params(adaptive_omega4=0, omega_4=0.01))
            assert np.abs(a - b).max() > 1e-9 and np.abs(a - c).max() > 1e-9
            a2, _ = collide(L, f, tau, params(adaptive_omega4=1, lambda_param=lam * 7.3, omega_4=1.0))
            assert np.abs(a - a2).max() == 0.0


def test_sponge_blends_rho_u_only():
    L = lib("f64")
    rng = np.random.default_rng(9)
    w = weights(L)[3]
    f = w * (1 + 0.05 * rng.uniform(-1, 1, 27))
    sp, u_in = 0.3, float(np.float32(0.05))      # u_inlet crosses the ABI as Float32
    out, mac = collide(L, f, 0.6, params(), sponge=sp, u_in=u_in)
    rho = f.sum()
    u = (f[:, None] * CVEC).sum(0) / rho
    assert abs(mac[0] - (rho * (1 - sp) + sp)) < 1e-15
    assert abs(mac[1] - (u[0] * (1 - sp) + u_in * sp)) < 1e-15
    assert abs(mac[2] - u[1] * (1 - sp)) < 1e-15 and abs(mac[3] - u[2] * (1 - sp)) < 1e-15
    # the populations are not blended, so the output mass is the unblended one
    assert abs(out.sum() - rho) < 1e-14


def test_sponge_blend_distributions_moves_mass():
    """sponge_blend_distributions = 1 (kernel_cumulant.jl:204-212) blends the
    streamed populations toward feq(1, u_in) before the collision, which
    keeps their sum: the output mass is (1 - sp) rho + sp (without the blend
    the collision returns rho itself)."""
    L = lib("f64")
    rng = np.random.default_rng(19)
    w = weights(L)[3]
    f = w * (1.04 + 0.05 * rng.uniform(-1, 1, 27))
    sp, u_in = float(np.float32(0.3)), float(np.float32(0.05))
    rho = f.sum()
    plain, _ = collide(L, f, 0.6, params(), sponge=sp, u_in=u_in)
    blend, _ = collide(L, f, 0.6, params(sponge_blend=1), sponge=sp, u_in=u_in)
    assert abs(plain.sum() - rho) < 1e-14
    assert abs(blend.sum() - ((1 - sp) * rho + sp)) < 1e-14
    # the momentum is the blended mass times the blended velocity (the collision is taken around
    # u = (1 - sp) j / rho + sp u_in x, with zero first central moments)
    u = (f[:, None] * CVEC).sum(0) / rho
    ub = (1 - sp) * u + np.array([sp * u_in, 0.0, 0.0])
    jb = (blend[:, None] * CVEC).sum(0)
    assert np.abs(jb - ((1 - sp) * rho + sp) * ub).max() < 1e-14


def test_positivity_limiter_clamps_negative():
    """A strong diagonal non-equilibrium drives the unlimited output negative;
    the positivity limiter (kernel_cumulant.jl:525-559) replaces it by
    feq + alpha (f - feq) with one alpha for all 27 directions,
    alpha = min over f - feq < 0 of feq / (feq - (f - feq) + 1e-15): every
    population becomes positive, mass and momentum are kept."""
    L = lib("f64")
    w = weights(L)[3]
    f = w.copy()
    f[0] += 0.02; f[26] -= 0.02                        # strong non-equilibrium along a diagonal
    out0, _ = collide(L, f, 0.5002, params(limiter=0, omega_3=1.99, omega_4=1.99))
    out2, mac = collide(L, f, 0.5002, params(limiter=2, omega_3=1.99, omega_4=1.99))
    assert out0.min() < -1e-3                          # the limiter has something to do
    assert out2.min() > 0.0
    rho = f.sum()
    u = (f[:, None] * CVEC).sum(0) / rho
    feq = np.array([L.orc_feq(0, rho, *u, k) for k in range(27)])
    fneq = out0 - feq
    neg = (fneq < 0) & (feq > 1e-15)
    alpha = max(min(1.0, float(np.min(feq[neg] / (feq[neg] - fneq[neg] + 1e-15)))), 0.0)
    assert 0.0 < alpha < 1.0
    assert np.abs(out2 - (feq + alpha * fneq)).max() < 1e-15
    assert abs(out2.sum() - out0.sum()) < 1e-14
    assert np.abs(((out2 - out0)[:, None] * CVEC).sum(0)).max() < 1e-14


def test_wall_model_force_direction():
    """Near a wall (0 < d < 10) the WMLES force opposes the velocity and the
    Guo term changes the momentum by (1 - omega/2)*F-scaled amounts; away from
    the wall (d = 100) nothing changes (kernel_cumulant.jl:216-254, 604-623)."""
    L = lib("f64")
    w = weights(L)[3]
    u = np.array([0.04, 0.0, 0.0])
    f = np.array([L.orc_feq(0, 1.0, *u, k) for k in range(27)])
    far, _ = collide(L, f, 0.501, params(wall_model=1), wall=100.0)
    near, _ = collide(L, f, 0.501, params(wall_model=1), wall=5.0)   # y+ ~ 160: log-law branch
    jx_far = (far * CXs).sum()
    jx_near = (near * CXs).sum()
    assert jx_near < jx_far - 1e-8


def test_gradient_noise_is_fmix32():
    """gpu_hash is MurmurHash3's fmix32 finaliser (physics_utils.jl:5-16)."""
    L = lib("f32")

    def fmix(h):
        h = np.uint32(h)
        with np.errstate(over="ignore"):
            h = np.uint32((h ^ (h >> np.uint32(16))) * np.uint32(0x85EBCA6B))
            h = np.uint32((h ^ (h >> np.uint32(13))) * np.uint32(0xC2B2AE35))
        return np.uint32(h ^ (h >> np.uint32(16)))

    rng = np.random.default_rng(1)
    for _ in range(50):
        gx, gy, gz = (int(v) for v in rng.integers(0, 5000, 3))
        seed = int(rng.integers(0, 10 ** 6))
        with np.errstate(over="ignore"):
            comb = (np.uint32(gx) * np.uint32(374761393) + np.uint32(gy) * np.uint32(668265263)
                    + np.uint32(gz) * np.uint32(1274126177) + np.uint32(seed))
        h = fmix(comb)
        ref = np.float32(np.float32(h & np.uint32(0xFFFF)) / np.float32(32768.0)) - np.float32(1.0)
        assert L.orc_gradient_noise(gx, gy, gz, seed) == float(ref)


def test_ramp_velocity():
    L = lib("f32")
    from openludwig_amd.solver import ramp_velocity
    for be in (1, 8, 250, 499, 500, 501, 4000):
        a = L.orc_ramp_velocity(0.05, be, 500)
        b = ramp_velocity(0.05, be, 500)
        assert a == np.float32(b)
        ref = 0.05 * (0.5 * (1 - math.cos(math.pi * min(be, 500) / 500)))
        assert abs(a - ref) < 1e-8


def _storage_cases():
    """f values whose offsets f - w cover normals, binary16 subnormals, exact
    ties of both formats, overflow and signed zero."""
    rng = np.random.default_rng(5)
    w = np.float32(8) / np.float32(27)
    d = np.concatenate([
        rng.uniform(-0.3, 0.3, 4000), rng.uniform(-1e-4, 1e-4, 2000),   # normals and binary16 subnormals
        rng.uniform(-6e-8, 6e-8, 500),                                  # below half the smallest subnormal
        [0.0, -0.0, 6.5e4, 6.6e4, -7e4, 65519.0, 65520.0],             # overflow threshold of binary16
    ]).astype(np.float32)
    # exact ties: midpoints between neighbouring binary16 values
    h = rng.integers(0, 0x7bff, 500).astype(np.uint16)
    lo = h.view(np.float16).astype(np.float32)
    hi = (h + 1).astype(np.uint16).view(np.float16).astype(np.float32)
    d = np.concatenate([d, (lo + hi) / 2, -(lo + hi) / 2]).astype(np.float32)
    # exact bfloat16 ties: low 16 bits = 0x8000
    b = (rng.integers(0x3000, 0x4000, 500).astype(np.uint32) << 16) | 0x8000
    d = np.concatenate([d, b.view(np.float32)]).astype(np.float32)
    return w, d


def test_storage_round_fp16_matches_numpy():
    """Offset storage FP16: w + binary16(f - w), RNE; numpy's float32 ->
    float16 conversion is IEEE roundTiesToEven and is the check."""
    L = lib("f32")
    w, d = _storage_cases()
    f = (w + d).astype(np.float32)
    got = np.array([L.orc_store_round(1, 13, float(x)) for x in f])
    dd = (f - w).astype(np.float32)
    with np.errstate(over="ignore"):
        exp = (w + dd.astype(np.float16).astype(np.float32)).astype(np.float32)
    assert np.array_equal(got.astype(np.float32), exp, equal_nan=True)


def test_storage_round_bf16_matches_torch():
    """Offset storage BF16: torch's float32 -> bfloat16 cast (RNE) is the
    check."""
    torch = pytest.importorskip("torch")
    L = lib("f32")
    w, d = _storage_cases()
    f = (w + d).astype(np.float32)
    got = np.array([L.orc_store_round(2, 13, float(x)) for x in f]).astype(np.float32)
    dd = torch.from_numpy((f - w).astype(np.float32))
    exp = (w + dd.to(torch.bfloat16).to(torch.float32).numpy()).astype(np.float32)
    assert np.array_equal(got, exp, equal_nan=True)


def test_storage_round_fp32_is_identity():
    L = lib("f32")
    for x in (0.1, 0.3, -1.5, 1e-30):
        assert L.orc_store_round(0, 0, x) == float(np.float32(x))
