"""ctypes front end of the CPU oracle (lb_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker / CPU baseline.  It restates the
reference's time step (see lb_oracle.c); nothing in openludwig_amd imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BUILD = HERE / "_build"


def build() -> None:
    subprocess.run(["bash", str(HERE / "build.sh")], check=True)


class OrcParams(C.Structure):
    _fields_ = [("c_wale", C.c_float), ("nu_sgs_bg", C.c_float), ("inlet_turbulence", C.c_float),
                ("symmetric", C.c_int), ("wall_model", C.c_int), ("temporal", C.c_int), ("sponge_blend", C.c_int),
                ("omega_bulk", C.c_float), ("omega_3", C.c_float), ("omega_4", C.c_float),
                ("lambda_param", C.c_float), ("adaptive_omega4", C.c_int), ("limiter", C.c_int),
                ("compressible", C.c_int), ("q_min", C.c_float), ("ddf", C.c_int), ("T_inlet", C.c_float),
                ("T_wall", C.c_float), ("wall_bc_adiabatic", C.c_int)]


_libs: dict = {}


def lib(precision: str = "f32") -> C.CDLL:
    if precision not in _libs:
        path = BUILD / f"liblb_oracle_{precision}.so"
        if not path.exists():
            build()
        L = C.CDLL(str(path))
        vp, ip, fp_, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_level.restype = C.c_int
        L.orc_set_level.argtypes = [vp, C.c_int, C.c_int, ip, ip, ip, ip, ip, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_uint8), fp_, fp_, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                    ip, C.POINTER(C.c_int8), C.POINTER(C.c_int8), C.POINTER(C.c_int8),
                                    C.POINTER(C.c_uint16)]
        L.orc_init_equilibrium.argtypes = [vp, C.c_float]
        L.orc_step_batch.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.POINTER(OrcParams)]
        L.orc_get.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.orc_put.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp]
        L.orc_ramp_velocity.restype = C.c_float
        L.orc_ramp_velocity.argtypes = [C.c_float, C.c_int, C.c_int]
        L.orc_collide_single.argtypes = [dp, dp, dp, C.c_float, C.POINTER(OrcParams), C.c_double, C.c_double,
                                         C.c_float, dp, dp]
        L.orc_trace_set.argtypes = [ip, C.POINTER(C.c_longlong), dp, C.c_int]
        L.orc_trace_count.restype = C.c_int
        L.orc_forces.argtypes = [vp, C.c_int, fp_, fp_, fp_, fp_, fp_, fp_, fp_, C.c_longlong, fp_, fp_, C.c_float,
                                 C.c_float, C.c_float, C.c_int, C.c_float, C.c_float, dp, fp_]
        L.orc_chimera_roundtrip.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_chimera_forward.argtypes = [dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_feq.restype = C.c_double
        L.orc_feq.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_interp_probe.restype = C.c_double
        L.orc_interp_probe.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float,
                                       C.c_float, C.c_int, C.c_int]
        L.orc_set_storage.argtypes = [vp, C.c_int]
        L.orc_store_round.argtypes = [C.c_int, C.c_int, C.c_double]
        L.orc_store_round.restype = C.c_double
        L.orc_gradient_noise.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32]
        L.orc_gradient_noise.restype = C.c_float if precision == "f32" else C.c_double
        L.orc_lattice.argtypes = [ip, ip, ip, fp_ if precision == "f32" else dp, ip, ip, ip]
        _libs[precision] = L
    return _libs[precision]


def params_from_cfg(cfg) -> OrcParams:
    p = OrcParams()
    p.c_wale = cfg.c_wale
    p.nu_sgs_bg = cfg.nu_sgs_background
    p.inlet_turbulence = cfg.inlet_turbulence_intensity
    p.symmetric = int(cfg.symmetric_analysis)
    p.wall_model = int(cfg.wall_model_enabled)
    p.temporal = int(cfg.temporal_interpolation)
    p.sponge_blend = int(cfg.sponge_blend_distributions)
    p.omega_bulk, p.omega_3, p.omega_4 = cfg.omega_bulk, cfg.omega_3, cfg.omega_4
    p.lambda_param = cfg.lambda_param
    p.adaptive_omega4 = int(cfg.adaptive_omega_4)
    p.limiter = cfg.limiter_code
    p.compressible = int(cfg.compressibility_correction)
    p.q_min = cfg.q_min_threshold
    p.ddf = int(cfg.ddf_enabled)
    p.T_inlet, p.T_wall = cfg.t_inlet, cfg.t_wall
    p.wall_bc_adiabatic = int(cfg.wall_bc == "adiabatic")
    return p


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """Multilevel CPU restatement holding the same host tables as the GPU."""

    def __init__(self, cfg, grids, params, precision: str = "f32", storage: int = 0):
        """storage: population storage of the HIP library to emulate
        (0 FP32, 1 FP16 offset, 2 BF16 offset)."""
        self.L = lib(precision)
        self.cfg = cfg
        self.grids = grids
        self.prm = params_from_cfg(cfg)
        self.ctx = self.L.orc_create(len(grids), params.nx_coarse, params.ny_coarse, params.nz_coarse)
        if not self.ctx:
            raise RuntimeError("orc_create failed")
        self.t_last = {g.level_id: 0 for g in grids}
        self._keep = []
        for i, g in enumerate(grids):
            mx = np.ascontiguousarray(g.coords[:, 0], dtype=np.int32)
            my = np.ascontiguousarray(g.coords[:, 1], dtype=np.int32)
            mz = np.ascontiguousarray(g.coords[:, 2], dtype=np.int32)
            nbt = np.ascontiguousarray(g.neighbor_table.T, dtype=np.int32)
            bp = np.ascontiguousarray(np.asfortranarray(g.block_pointer).ravel(order="F"), dtype=np.int32)
            pdx, pdy, pdz = (int(v) for v in g.block_pointer.shape)
            n = g.n_boundary_cells
            cb = np.ascontiguousarray(g.bouzidi_cells[:, 0], dtype=np.int32)
            cx = np.ascontiguousarray(g.bouzidi_cells[:, 1], dtype=np.int8)
            cy = np.ascontiguousarray(g.bouzidi_cells[:, 2], dtype=np.int8)
            cz = np.ascontiguousarray(g.bouzidi_cells[:, 3], dtype=np.int8)
            q = np.ascontiguousarray(g.bouzidi_q, dtype=np.uint16)
            self._keep += [mx, my, mz, nbt, bp, cb, cx, cy, cz, q]
            rc = self.L.orc_set_level(self.ctx, i, g.n_blocks, _p(mx, C.c_int), _p(my, C.c_int), _p(mz, C.c_int),
                                      _p(nbt, C.c_int), _p(bp, C.c_int), pdx, pdy, pdz,
                                      _p(g.obstacle, C.c_uint8), _p(g.sponge, C.c_float),
                                      _p(g.wall_dist, C.c_float), g.tau, g.tau_g, int(cfg.temporal_interpolation),
                                      int(cfg.ddf_enabled), n, _p(cb, C.c_int), _p(cx, C.c_int8),
                                      _p(cy, C.c_int8), _p(cz, C.c_int8), _p(q, C.c_uint16))
            if rc != 0:
                raise RuntimeError(f"orc_set_level failed for level {i + 1}")
        if self.L.orc_set_storage(self.ctx, int(storage)) != 0:
            raise RuntimeError("orc_set_storage failed")

    def close(self):
        if self.ctx:
            self.L.orc_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_equilibrium(self):
        self.L.orc_init_equilibrium(self.ctx, float(self.cfg.t_initial))
        self.t_last = {g.level_id: 0 for g in self.grids}

    def step_batch(self, t_start: int, n: int, u_curr: float):
        self.L.orc_step_batch(self.ctx, int(t_start), int(n), float(u_curr), C.byref(self.prm))
        t_end = t_start + n - 1
        if n > 0:
            for g in self.grids:
                # last substep of level l in coarse step t: t*2^(l-1) + 2^(l-1) - 1
                s = 2 ** (g.level_id - 1)
                self.t_last[g.level_id] = t_end * s + s - 1

    def get(self, level_id: int, field: str) -> np.ndarray:
        g = self.grids[level_id - 1]
        which, n = {"f": (0, 27), "rho": (1, 1), "vel": (2, 3), "g": (3, 27), "T": (4, 1)}[field]
        out = np.zeros(g.n_cells * n)
        rc = self.L.orc_get(self.ctx, level_id - 1, which, self.t_last[level_id], _p(out, C.c_double))
        if rc != 0:
            raise RuntimeError(f"orc_get({field}) failed")
        return out

    def put(self, level_id: int, field: str, data: np.ndarray):
        which = {"f": 0, "rho": 1, "vel": 2, "g": 3, "T": 4}[field]
        a = np.ascontiguousarray(data, dtype=np.float64)
        # write both halves of the pair by writing at both parities
        for t in (0, 1):
            rc = self.L.orc_put(self.ctx, level_id - 1, which, t, _p(a, C.c_double))
            if rc != 0:
                raise RuntimeError("orc_put failed")


def oracle_forces(orc: "Oracle", cfg, params, mesh, level_dx: float):
    """compute_aerodynamics! restated (FP64 ordered sum): dict like Solver.compute_aerodynamics."""
    f32 = np.float32
    arr = [np.ascontiguousarray(a, dtype=np.float32) for a in
           (mesh.centers[:, 0], mesh.centers[:, 1], mesh.centers[:, 2], mesh.normals[:, 0], mesh.normals[:, 1],
            mesh.normals[:, 2], mesh.areas)]
    off = np.array(params.mesh_offset, dtype=np.float32)
    ref = np.array(params.moment_center, dtype=np.float32)
    ps = float(f32(params.rho_physical * params.velocity_scale * params.velocity_scale))
    out = np.zeros(9)
    maps = np.zeros(4 * mesh.n, dtype=np.float32)
    orc.L.orc_forces(orc.ctx, 0, *[_p(a, C.c_float) for a in arr], mesh.n, _p(off, C.c_float), _p(ref, C.c_float),
                     float(f32(level_dx)), ps, ps, 5, cfg.c_wale, cfg.nu_sgs_background, _p(out, C.c_double),
                     _p(maps, C.c_float))
    Fxp, Fyp, Fzp, Fxv, Fyv, Fzv, Mx, My, Mz = out
    if cfg.symmetric_analysis:
        Fxp *= 2.0; Fzp *= 2.0; Fxv *= 2.0; Fzv *= 2.0; My *= 2.0
        Fyp = Fyv = Mx = Mz = 0.0
    q_inf = 0.5 * params.rho_physical * params.u_physical ** 2
    F_ref = q_inf * params.reference_area
    M_ref = F_ref * params.reference_chord
    res = {"Fx_pressure": Fxp, "Fy_pressure": Fyp, "Fz_pressure": Fzp, "Fx_viscous": Fxv, "Fy_viscous": Fyv,
           "Fz_viscous": Fzv, "Mx": Mx, "My": My, "Mz": Mz, "Fx": Fxp + Fxv, "Fy": Fyp + Fyv, "Fz": Fzp + Fzv}
    res["Cd"] = res["Fx"] / F_ref if F_ref > 1e-10 else 0.0
    res["Cl"] = res["Fz"] / F_ref if F_ref > 1e-10 else 0.0
    res["Cs"] = res["Fy"] / F_ref if F_ref > 1e-10 else 0.0
    res["Cmx"], res["Cmy"], res["Cmz"] = ((Mx / M_ref, My / M_ref, Mz / M_ref) if M_ref > 1e-10 else (0.0, 0.0, 0.0))
    res["maps"] = maps.reshape(4, mesh.n)
    return res
