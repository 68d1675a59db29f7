"""Host driver over libludwig_hip.so: the counterpart of the Julia host calls
that this library replaces (see include/ludwig_hip.h).

``Solver`` owns one device context:
  upload_levels      adapt(backend, BlockLevel)            (main.jl:103)
  init_equilibrium   init_eq! / init_thermal_eq!           (main.jl:113-172)
  execute_timestep_batch  execute_timestep_batch!          (solver_control.jl:199)
  compute_aerodynamics    compute_aerodynamics!            (surface.jl:913)
  run                the main loop cadence                  (main.jl:208-315)
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .config import CaseConfig
from .domain import BlockLevel, DomainParameters, compute_domain_from_mesh, setup_multilevel_domain
from .geometry import SolverMesh, load_mesh, merge_meshes, rotate_mesh


class LudwigError(RuntimeError):
    pass


class DivergenceError(LudwigError):
    """A case holds non-finite density or velocity values (LB_ENONFINITE)."""

    def __init__(self, msg: str, count: int = 0):
        super().__init__(msg)
        self.count = count


def step_params(cfg: CaseConfig, params: DomainParameters | None = None) -> N.LbStepParams:
    """Scalars of the launcher (timestep.jl:40-86) from the case configuration."""
    p = N.LbStepParams()
    p.c_wale = cfg.c_wale
    p.nu_sgs_background = cfg.nu_sgs_background
    p.inlet_turbulence = cfg.inlet_turbulence_intensity
    p.symmetric = int(cfg.symmetric_analysis)
    p.wall_model = int(cfg.wall_model_enabled)
    p.sponge_blend_distributions = int(cfg.sponge_blend_distributions)
    p.omega_bulk = cfg.omega_bulk
    p.omega_3 = cfg.omega_3
    p.omega_4 = cfg.omega_4
    p.adaptive_omega4 = int(cfg.adaptive_omega_4)
    p.lambda_param = cfg.lambda_param
    p.limiter = cfg.limiter_code
    p.compressible = int(cfg.compressibility_correction)
    p.T_inlet = cfg.t_inlet
    p.T_wall = cfg.t_wall
    p.wall_bc_adiabatic = int(cfg.wall_bc == "adiabatic")
    return p


def ramp_velocity(u_target: float, batch_end: int, ramp_steps: int) -> float:
    """Inlet ramp of main.jl:217-218, evaluated in Float32."""
    if batch_end <= ramp_steps:
        a = np.float32(np.pi) * np.float32(batch_end)
        a = a / np.float32(ramp_steps)
        prog = np.float32(0.5) * (np.float32(1.0) - np.cos(a, dtype=np.float32))
        return float(np.float32(u_target) * prog)
    return float(np.float32(u_target))


def level_desc(lev: BlockLevel, keep: list, dense_q: bool = False) -> N.LbLevelDesc:
    """The lb_level_desc of a level.  dense_q passes the Bouzidi q map as the
    reference holds it (level.bouzidi_q_map, (8,8,8,nb,27) Float16 bits, 0
    where no link) instead of the per-cell list."""
    d = N.LbLevelDesc()
    nb = lev.n_blocks
    mx = np.ascontiguousarray(lev.coords[:, 0]); my = np.ascontiguousarray(lev.coords[:, 1])
    mz = np.ascontiguousarray(lev.coords[:, 2])
    nbt = np.ascontiguousarray(lev.neighbor_table.T)       # column-major (nb,27)
    bptr = np.ascontiguousarray(np.asfortranarray(lev.block_pointer).ravel(order="F"))
    keep.extend([mx, my, mz, nbt, bptr])
    d.level_id = lev.level_id
    d.n_blocks = nb
    d.map_x, d.map_y, d.map_z = N.ptr(mx, C.c_int32), N.ptr(my, C.c_int32), N.ptr(mz, C.c_int32)
    d.neighbor_table = N.ptr(nbt, C.c_int32)
    d.block_pointer = N.ptr(bptr, C.c_int32)
    d.ptr_dim_x, d.ptr_dim_y, d.ptr_dim_z = (int(v) for v in lev.block_pointer.shape)
    d.obstacle = N.ptr(lev.obstacle, C.c_uint8)
    d.sponge = N.ptr(lev.sponge, C.c_float)
    d.wall_dist = N.ptr(lev.wall_dist, C.c_float)
    d.tau = lev.tau
    d.tau_g = lev.tau_g
    d.dx = lev.dx
    n = lev.n_boundary_cells
    d.n_boundary_cells = n
    if n:
        cb = np.ascontiguousarray(lev.bouzidi_cells[:, 0].astype(np.int32))
        cx = np.ascontiguousarray(lev.bouzidi_cells[:, 1].astype(np.int8))
        cy = np.ascontiguousarray(lev.bouzidi_cells[:, 2].astype(np.int8))
        cz = np.ascontiguousarray(lev.bouzidi_cells[:, 3].astype(np.int8))
        q = np.ascontiguousarray(lev.bouzidi_q)
        keep.extend([cb, cx, cy, cz, q])
        d.bouzidi_cell_block = N.ptr(cb, C.c_int32)
        d.bouzidi_cell_x, d.bouzidi_cell_y, d.bouzidi_cell_z = (N.ptr(cx, C.c_int8), N.ptr(cy, C.c_int8),
                                                                N.ptr(cz, C.c_int8))
        if dense_q:
            dense = np.zeros((27, nb, 512), dtype=np.uint16)
            cells = lev.bouzidi_cells.astype(np.int64)
            cl = (cells[:, 1] - 1) + 8 * (cells[:, 2] - 1) + 64 * (cells[:, 3] - 1)
            dense[:, cells[:, 0] - 1, cl] = q.T
            keep.append(dense)
            d.bouzidi_q_dense = N.ptr(dense, C.c_uint16)
        else:
            d.bouzidi_q_cells = N.ptr(q, C.c_uint16)
    return d


@dataclass
class ForceAccumulator:
    """ForceAccumulator / accumulate_forces! (surface.jl:6-56)."""
    sums: np.ndarray = field(default_factory=lambda: np.zeros(9))
    cd: list = field(default_factory=list)
    n_samples: int = 0

    def add(self, f: dict) -> None:
        self.sums += np.array([f[k] for k in N.FORCE_KEYS[:9]])
        self.cd.append(f["Cd"])
        self.n_samples += 1


def averaged_coefficients(params: DomainParameters, sums: np.ndarray, n_samples: int) -> dict:
    """Time-averaged coefficients from accumulated force samples
    (accumulate_forces!, surface.jl:36-56): sums of Fp xyz, Fv xyz, M xyz."""
    res = {"n_samples": int(n_samples)}
    if n_samples:
        q_inf = 0.5 * params.rho_physical * params.u_physical ** 2
        F_ref = q_inf * params.reference_area
        avg = np.asarray(sums, dtype=np.float64) / n_samples
        res["Cd_avg"] = (avg[0] + avg[3]) / F_ref
        res["Cl_avg"] = (avg[2] + avg[5]) / F_ref
    return res


def lb_config(cfg: CaseConfig, grids: list, params: DomainParameters, storage: int, strict_fp: bool,
              wide_offsets: bool = False) -> N.LbConfig:
    c = N.LbConfig()
    c.n_levels = len(grids)
    c.domain_nx, c.domain_ny, c.domain_nz = params.nx_coarse, params.ny_coarse, params.nz_coarse
    c.temporal_interpolation = int(cfg.temporal_interpolation)
    c.ddf_enabled = int(cfg.ddf_enabled)
    c.q_min_threshold = cfg.q_min_threshold
    c.storage = storage
    c.strict_fp = int(strict_fp)
    c.wide_offsets = int(wide_offsets)
    return c


def footprint(cfg: CaseConfig, grids: list, params: DomainParameters, mesh: SolverMesh | None,
              storage: int = N.LB_STORAGE_FP32, n_ramp_batches: int = 0) -> int:
    """Exact device bytes of a case (lb_footprint), computed without a GPU."""
    lib = N.hip_lib()
    keep: list = []
    descs = (N.LbLevelDesc * len(grids))(*[level_desc(g, keep) for g in grids])
    n_tri = mesh.n if (mesh is not None and cfg.forces_enabled) else 0
    out = C.c_int64()
    rc = lib.lb_footprint(C.byref(lb_config(cfg, grids, params, storage, False)), descs, len(grids), n_tri,
                          int(n_ramp_batches), C.byref(out))
    if rc != 0:
        raise LudwigError(f"lb_footprint failed ({rc})")
    return out.value


class Solver:
    def __init__(self, cfg: CaseConfig, grids: list, params: DomainParameters, mesh: SolverMesh | None = None,
                 device: int = 0, strict_fp: bool = False, storage: int = N.LB_STORAGE_FP32,
                 wide_offsets: bool = False, dense_q: bool = False):
        self.cfg = cfg
        self.grids = grids
        self.params = params
        self.mesh = mesh
        self.lib = N.hip_lib()
        self._ctx = C.c_void_p()
        self.engine = None  # set while attached to a batched Engine
        c = lb_config(cfg, grids, params, storage, strict_fp, wide_offsets)
        rc = self.lib.lb_create(C.byref(self._ctx), device, C.byref(c))
        if rc != 0:
            raise LudwigError(f"lb_create failed ({rc})")
        self.step_params = step_params(cfg, params)
        self.t_next = 1
        keep: list = []
        for lev in grids:
            d = level_desc(lev, keep, dense_q)
            self._check(self.lib.lb_level_upload(self._ctx, C.byref(d)), "lb_level_upload")
        if mesh is not None and cfg.forces_enabled:
            arr = [np.ascontiguousarray(a, dtype=np.float32) for a in
                   (mesh.centers[:, 0], mesh.centers[:, 1], mesh.centers[:, 2], mesh.normals[:, 0],
                    mesh.normals[:, 1], mesh.normals[:, 2], mesh.areas)]
            self._check(self.lib.lb_mesh_upload(self._ctx, mesh.n, *[N.ptr(a, C.c_float) for a in arr]),
                        "lb_mesh_upload")

    # -- lifecycle -------------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            if self.engine is not None:
                self.engine.remove(self)
            self.lib.lb_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc != 0:
            msg = self.lib.lb_last_error(self._ctx).decode(errors="replace")
            raise LudwigError(f"{what} failed ({rc}): {msg}")

    @property
    def device_bytes(self) -> int:
        return int(self.lib.lb_device_bytes(self._ctx))

    # -- the replaced host calls ----------------------------------------------
    def init_equilibrium(self) -> None:
        self._check(self.lib.lb_init_equilibrium(self._ctx, float(self.cfg.t_initial)), "lb_init_equilibrium")
        self.t_next = 1

    def execute_timestep_batch(self, t_start: int, batch_size: int, u_curr: float) -> None:
        self._check(self.lib.lb_step_batch(self._ctx, int(t_start), int(batch_size), float(u_curr),
                                           C.byref(self.step_params)), "lb_step_batch")
        self.t_next = t_start + batch_size

    def synchronize(self) -> None:
        self._check(self.lib.lb_synchronize(self._ctx), "lb_synchronize")

    def _force_params(self) -> N.LbForceParams:
        p = self.params
        fp = N.LbForceParams()
        for i in range(3):
            fp.mesh_offset[i] = p.mesh_offset[i]
            fp.moment_center[i] = p.moment_center[i]
        fp.velocity_scale = p.velocity_scale
        fp.rho_physical = p.rho_physical
        fp.rho_ref = p.rho_physical
        fp.u_ref = p.u_physical
        fp.area_ref = p.reference_area
        fp.chord_ref = p.reference_chord
        fp.c_wale = self.cfg.c_wale
        fp.nu_sgs_background = self.cfg.nu_sgs_background
        fp.search_radius = 5
        fp.symmetric = int(self.cfg.symmetric_analysis)
        return fp

    def compute_aerodynamics(self, maps: bool = False) -> dict:
        """compute_aerodynamics! (surface.jl:913-919) on the finest level."""
        fp = self._force_params()
        out = np.zeros(N.LB_F_N)
        mp = [np.zeros(self.mesh.n, np.float32) for _ in range(4)] if maps else [None] * 4
        self._check(self.lib.lb_forces(self._ctx, C.byref(fp), N.ptr(out, C.c_double),
                                       *[N.ptr(m, C.c_float) for m in mp]), "lb_forces")
        res = dict(zip(N.FORCE_KEYS, (float(v) for v in out)))
        if maps:
            res["maps"] = mp
        return res

    def surface_maps(self, side: int = 1) -> list:
        """Per-triangle [p, tau_x, tau_y, tau_z] (Float32) of the plus (+1) or
        minus (-1) side of the surface (lb_surface_maps)."""
        mp = [np.zeros(self.mesh.n, np.float32) for _ in range(4)]
        self._check(self.lib.lb_surface_maps(self._ctx, C.byref(self._force_params()), int(side),
                                             *[N.ptr(m, C.c_float) for m in mp]), "lb_surface_maps")
        return mp

    def compute_net_aerodynamics(self, closed_body: bool = True) -> tuple:
        """compute_net_aerodynamics! (surface.jl:980-1035): per-triangle net
        force F = (-p_net n + tau_net) A with p_net = p_plus - p_minus, in
        Float32 as compute_net_triangle_forces_kernel! (surface.jl:665-700).
        A closed body (the reference's default) has no fluid inside, so the
        minus side is zero; an open surface probes both sides."""
        plus = self.surface_maps(1)
        minus = [np.zeros_like(m) for m in plus] if closed_body else self.surface_maps(-1)
        f32 = np.float32
        nx, ny, nz = (self.mesh.normals[:, i].astype(f32) for i in range(3))
        area = self.mesh.areas.astype(f32)
        p_net = plus[0] - minus[0]
        s_net = [plus[i] - minus[i] for i in (1, 2, 3)]
        out = []
        for n_c, s_c in zip((nx, ny, nz), s_net):
            out.append(((-p_net) * n_c) * area + s_c * area)
        return tuple(out)

    def download(self, level_id: int, field: str) -> np.ndarray:
        lev = self.grids[level_id - 1]
        n4 = lev.n_cells
        code, n = {"f": (N.LB_FIELD_F, 27), "rho": (N.LB_FIELD_RHO, 1), "vel": (N.LB_FIELD_VEL, 3),
                   "g": (N.LB_FIELD_G, 27), "T": (N.LB_FIELD_T, 1)}[field]
        out = np.zeros(n4 * n, dtype=np.float32)
        self._check(self.lib.lb_download_level(self._ctx, level_id, code, N.ptr(out, C.c_float)),
                    "lb_download_level")
        return out

    def upload(self, level_id: int, field: str, data: np.ndarray) -> None:
        code = {"f": N.LB_FIELD_F, "rho": N.LB_FIELD_RHO, "vel": N.LB_FIELD_VEL, "g": N.LB_FIELD_G,
                "T": N.LB_FIELD_T}[field]
        a = np.ascontiguousarray(data, dtype=np.float32)
        self._check(self.lib.lb_upload_level(self._ctx, level_id, code, N.ptr(a, C.c_float)), "lb_upload_level")

    def profile(self, enable: bool = True) -> None:
        self._check(self.lib.lb_profile_enable(self._ctx, int(enable)), "lb_profile_enable")

    def profile_read(self, level_id: int, kernel: int = 0) -> tuple:
        """(summed kernel ms, launches) of a level since the last read;
        kernel 0 = stream-collide, 1 = thermal DDF."""
        ms = C.c_double()
        n = C.c_int64()
        self._check(self.lib.lb_profile_read_kernel(self._ctx, level_id, kernel, C.byref(ms), C.byref(n)),
                    "lb_profile_read_kernel")
        return ms.value, n.value

    def stats(self, level_id: int = 1) -> dict:
        out = np.zeros(8)
        self._check(self.lib.lb_stats(self._ctx, level_id, N.ptr(out, C.c_double)), "lb_stats")
        return {"rho_sum": out[0], "kinetic_energy": out[1], "rho_min": out[2], "rho_max": out[3],
                "nonfinite": int(out[4]), "fluid_cells": int(out[5]), "mass": out[6], "v_max": out[7]}

    def check_finite(self) -> None:
        """Divergence check (lb_check_finite): raises DivergenceError when a
        level holds non-finite density or velocity values."""
        n = C.c_int64()
        rc = self.lib.lb_check_finite(self._ctx, C.byref(n))
        if rc == N.LB_ENONFINITE:
            raise DivergenceError(f"case diverged: {self.lib.lb_last_error(self._ctx).decode(errors='replace')}",
                                  int(n.value))
        self._check(rc, "lb_check_finite")

    def mass(self) -> float:
        """FP64 sum of every population of every level, each level weighted by
        its cell volume relative to the coarsest (the global mass check)."""
        return sum(self.stats(g.level_id)["mass"] / 8.0 ** (g.level_id - 1) for g in self.grids)

    # -- main loop (main.jl:208-315) ----------------------------------------------
    def run(self, steps: int, diag_freq: int | None = None, on_diag=None, check_stability: bool = True) -> dict:
        """The main loop of main.jl:208-315: batches of gpu_async_depth coarse
        steps with the ramped inlet velocity; at every diagnostic step the
        divergence check (DivergenceError on non-finite values) and the
        forces, averaged over the samples after the ramp
        (accumulate_forces!, surface.jl:36-56)."""
        cfg = self.cfg
        batch = cfg.gpu_async_depth
        diag = diag_freq or cfg.diag_freq
        acc = ForceAccumulator()
        t = 1
        history = []
        while t <= steps:
            batch_end = min(t + batch - 1, steps)
            actual = batch_end - t + 1
            u_curr = ramp_velocity(cfg.u_lattice, batch_end, cfg.ramp_steps)
            self.execute_timestep_batch(t, actual, u_curr)
            if batch_end % diag < actual or batch_end == steps:
                diag_step = (batch_end // diag) * diag
                if t <= diag_step <= batch_end:
                    if check_stability:
                        self.check_finite()
                    rec = {"step": diag_step, "u_lat": u_curr}
                    if self.mesh is not None and cfg.forces_enabled:
                        f = self.compute_aerodynamics()
                        rec.update({k: f[k] for k in ("Cd", "Cl", "Cs")})
                        if diag_step > cfg.ramp_steps:
                            acc.add(f)
                    history.append(rec)
                    if on_diag:
                        on_diag(rec)
            t = batch_end + 1
        self.synchronize()
        res = {"history": history}
        res.update(averaged_coefficients(self.params, acc.sums, acc.n_samples))
        return res


def build_case(cfg: CaseConfig, mesh: SolverMesh, verbose: bool = False, gpu_device: int | None = None):
    """setup_multilevel_domain(stl_paths) (domain.jl:248-295) from a mesh
    already loaded: rotate, scale the domain, build every level (the
    voxelization, wall distances and Bouzidi Q-map on `gpu_device` if given)."""
    mesh = rotate_mesh(mesh, cfg.alpha_deg, cfg.beta_deg)
    params = compute_domain_from_mesh(cfg, mesh.min_bounds, mesh.max_bounds)
    grids = setup_multilevel_domain(cfg, mesh, params, verbose=verbose, gpu_device=gpu_device)
    return grids, params, mesh


def build_case_from_stl(cfg: CaseConfig, verbose: bool = False):
    meshes = [load_mesh(p, cfg.stl_scale, cfg.minimum_facet_size) for p in cfg.stl_files]
    return build_case(cfg, merge_meshes(meshes), verbose)


def mlups(grids: list, coarse_steps: int, seconds: float) -> tuple:
    """(MLUPS_true, MLUPS_ref): every substep counted / the reference's
    sub-cycling-blind definition (main.jl:175, 232, 377)."""
    lu_true = sum(g.n_cells * 2 ** (g.level_id - 1) for g in grids) * coarse_steps
    lu_ref = sum(g.n_cells for g in grids) * coarse_steps
    return lu_true / seconds / 1e6, lu_ref / seconds / 1e6
