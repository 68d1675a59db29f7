"""Domain scaling and block topology: the host side of OpenLUDWIG's setup
(src/solver/physics_scaling.jl, src/geometry/domain.jl,
src/geometry/domain_topology.jl) producing one ``BlockLevel`` of reference-
layout tables per level.  Set logic runs on dense boolean block grids; the
per-cell loops (voxelization, flood fill, sponge, wall distance, Bouzidi
Q-map) run in libludwig_setup.so.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .config import CaseConfig, f32
from .geometry import SolverMesh

BLOCK_SIZE = 8


@dataclass
class DomainParameters:
    """DomainParameters (physics_scaling.jl:7-53), the fields the path uses."""
    num_levels: int
    mesh_min: tuple
    mesh_max: tuple
    mesh_center: tuple
    mesh_extent: tuple
    reference_length: float
    reference_chord: float
    reference_area: float
    moment_center: tuple
    domain_size: tuple
    mesh_offset: tuple
    dx_fine: float
    dx_coarse: float
    dx_levels: list
    nx_coarse: int
    ny_coarse: int
    nz_coarse: int
    bx_max: int
    by_max: int
    bz_max: int
    nu_lattice: float
    tau_levels: list           # Float32 values as Python floats
    re_number: float
    u_physical: float
    rho_physical: float
    nu_physical: float
    length_scale: float
    time_scale: float
    velocity_scale: float
    force_scale: float
    tau_fine: float
    tau_g_fine: float
    tau_g_levels: list


def compute_tau_for_levels(Re: float, resolution: int, u_lattice: float) -> float:
    """compute_tau_for_levels (physics_scaling.jl:63-66)."""
    return 3.0 * (float(u_lattice) * resolution / Re) + 0.5


def compute_max_levels_for_domain(domain_size: float, dx_fine: float, block_size: int, min_blocks: int) -> int:
    ratio = domain_size / (dx_fine * min_blocks * block_size)
    return 1 if ratio < 1.0 else int(math.floor(1 + math.log2(ratio)))


def compute_domain_from_mesh(cfg: CaseConfig, mesh_min: tuple, mesh_max: tuple) -> DomainParameters:
    """compute_domain_from_mesh (physics_scaling.jl:83-191)."""
    mesh_center = tuple((mesh_min[i] + mesh_max[i]) / 2 for i in range(3))
    mesh_extent = tuple(mesh_max[i] - mesh_min[i] for i in range(3))
    if cfg.reference_length_for_meshing > 0:
        ref_length = cfg.reference_length_for_meshing
    else:
        ref_length = {"x": mesh_extent[0], "y": mesh_extent[1], "z": mesh_extent[2]}.get(
            cfg.reference_dimension, max(mesh_extent))
    ref_chord = cfg.reference_chord if cfg.reference_chord > 0 else mesh_extent[0]
    if cfg.reference_area_config > 0:
        ref_area = cfg.reference_area_config
    else:
        ref_area = mesh_extent[1] * mesh_extent[2] * 2 if cfg.symmetric_analysis else mesh_extent[1] * mesh_extent[2]
    mc_rel = tuple(float(v) for v in cfg.moment_center)
    u_phys, nu_phys, rho_phys = cfg.flow_velocity, cfg.kinematic_viscosity, cfg.fluid_density
    re_number = u_phys * ref_length / nu_phys
    tau_fine = max(compute_tau_for_levels(re_number, cfg.surface_resolution, cfg.u_lattice), cfg.tau_min)
    domain_x = ref_length * (cfg.domain_upstream + cfg.domain_downstream) + mesh_extent[0]
    domain_y = (mesh_max[1] + ref_length * cfg.domain_lateral) if cfg.symmetric_analysis else \
        (mesh_extent[1] + 2 * ref_length * cfg.domain_lateral)
    domain_z = mesh_extent[2] + 2 * ref_length * cfg.domain_height
    dx_fine = ref_length / cfg.surface_resolution
    min_domain = min(domain_x, domain_y, domain_z)
    max_levels_domain = compute_max_levels_for_domain(min_domain, dx_fine, cfg.block_size, cfg.min_coarse_blocks)
    if cfg.num_levels > 0:
        num_levels = min(cfg.num_levels, max_levels_domain)
    else:
        num_levels = min(max_levels_domain, cfg.max_levels) if cfg.auto_levels else min(8, max_levels_domain)
    dx_coarse = dx_fine * 2 ** (num_levels - 1)
    dx_levels = [dx_fine * 2 ** (num_levels - lvl) for lvl in range(1, num_levels + 1)]
    bs = cfg.block_size
    nx = max(bs, int(math.ceil(math.ceil(domain_x / dx_coarse) / bs) * bs))
    ny = max(bs, int(math.ceil(math.ceil(domain_y / dx_coarse) / bs) * bs))
    nz = max(bs, int(math.ceil(math.ceil(domain_z / dx_coarse) / bs) * bs))
    domain_x, domain_y, domain_z = nx * dx_coarse, ny * dx_coarse, nz * dx_coarse
    mesh_x = ref_length * cfg.domain_upstream
    mesh_y = 0.0 if cfg.symmetric_analysis else (domain_y / 2 - mesh_center[1])
    mesh_z = domain_z / 2 - mesh_center[2]
    mesh_offset = (mesh_x - mesh_min[0], mesh_y, mesh_z)
    length_scale = dx_fine
    velocity_scale = u_phys / float(cfg.u_lattice)
    time_scale = length_scale / velocity_scale
    nu_lattice_fine = nu_phys * time_scale / (length_scale ** 2)
    tau_levels = []
    for lvl in range(1, num_levels + 1):
        t = tau_fine if lvl == num_levels else 0.5 + (tau_fine - 0.5) * 2.0 ** (num_levels - lvl)
        tau_levels.append(f32(t))
    force_scale = rho_phys * length_scale ** 4 / time_scale ** 2
    moment_center = (mesh_min[0] + mesh_offset[0] + mc_rel[0] * ref_chord,
                     mesh_center[1] + mesh_offset[1] + mc_rel[1] * ref_chord,
                     mesh_center[2] + mesh_offset[2] + mc_rel[2] * ref_chord)
    tau_g_fine = 0.0
    tau_g_levels = []
    if cfg.ddf_enabled:
        kappa = nu_lattice_fine / cfg.prandtl
        tau_g_fine = 0.5 + 3.0 * kappa
        for lvl in range(1, num_levels + 1):
            t = tau_g_fine if lvl == num_levels else 0.5 + (tau_g_fine - 0.5) * 2.0 ** (num_levels - lvl)
            tau_g_levels.append(f32(t))
    return DomainParameters(
        num_levels=num_levels, mesh_min=tuple(mesh_min), mesh_max=tuple(mesh_max), mesh_center=mesh_center,
        mesh_extent=mesh_extent, reference_length=ref_length, reference_chord=ref_chord, reference_area=ref_area,
        moment_center=moment_center, domain_size=(domain_x, domain_y, domain_z), mesh_offset=mesh_offset,
        dx_fine=dx_fine, dx_coarse=dx_coarse, dx_levels=dx_levels, nx_coarse=nx, ny_coarse=ny, nz_coarse=nz,
        bx_max=nx // bs, by_max=ny // bs, bz_max=nz // bs, nu_lattice=nu_lattice_fine, tau_levels=tau_levels,
        re_number=re_number, u_physical=u_phys, rho_physical=rho_phys, nu_physical=nu_phys,
        length_scale=length_scale, time_scale=time_scale, velocity_scale=velocity_scale, force_scale=force_scale,
        tau_fine=tau_fine, tau_g_fine=tau_g_fine, tau_g_levels=tau_g_levels)


@dataclass
class BlockLevel:
    """Host tables of one level, in the reference's layout (blocks.jl:16-229)."""
    level_id: int
    dx: float                      # Float64(Float32(dx)) as the reference stores it
    tau: float
    tau_g: float
    coords: np.ndarray             # (nb,3) int32, 1-based, sorted (bx,by,bz)
    neighbor_table: np.ndarray     # (nb,27) int32, 1-based, 0 = none (Fortran order when flattened)
    block_pointer: np.ndarray      # (px,py,pz) int32 (max active coords), 1-based ids
    grid_dims: tuple               # level dims (bx_max*scale, ...)
    obstacle: np.ndarray           # (nb,512) uint8 == (8,8,8,nb) column-major
    sponge: np.ndarray             # (nb,512) float32
    wall_dist: np.ndarray          # (nb,512) float32 (100 where not near a wall)
    bouzidi_cells: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))  # (n,4) b,lx,ly,lz
    bouzidi_q: np.ndarray = field(default_factory=lambda: np.zeros((0, 27), np.uint16))     # FP16 bits
    bouzidi_tri: np.ndarray = field(default_factory=lambda: np.zeros((0, 27), np.int32))

    @property
    def n_blocks(self) -> int:
        return int(self.coords.shape[0])

    @property
    def n_boundary_cells(self) -> int:
        return int(self.bouzidi_cells.shape[0])

    @property
    def n_cells(self) -> int:
        return self.n_blocks * 512


def _dense_ptr(coords: np.ndarray, dims: tuple) -> np.ndarray:
    """Dense 1-based block pointer of the level dims, Fortran order (bx fastest)."""
    g = np.zeros(dims, dtype=np.int32, order="F")
    if coords.shape[0]:
        g[coords[:, 0] - 1, coords[:, 1] - 1, coords[:, 2] - 1] = np.arange(1, coords.shape[0] + 1, dtype=np.int32)
    return g


def get_active_blocks_for_level(mesh: SolverMesh, dx: float, offset, dims) -> np.ndarray:
    """get_active_blocks_for_level (domain_topology.jl:9-52) as a boolean grid."""
    grid = np.zeros(dims, dtype=bool)
    margin = dx * 0.01
    inv_bs_dx = 1.0 / (BLOCK_SIZE * dx)
    p = mesh.triangles + np.asarray(offset, dtype=np.float64)
    tmin = p.min(axis=1)
    tmax = p.max(axis=1)
    lo = np.floor((tmin - margin) * inv_bs_dx).astype(np.int64) + 1
    hi = np.floor((tmax + margin) * inv_bs_dx).astype(np.int64) + 1
    lo = np.maximum(lo, 1)
    hi = np.minimum(hi, np.asarray(dims))
    keep = (hi >= lo).all(axis=1)
    lo, hi = lo[keep] - 1, hi[keep] - 1          # 0-based, inclusive
    span = hi - lo
    # the union of the triangles' block boxes: boxes up to 4 blocks wide per
    # axis are set in 64 vectorised passes, wider ones one by one
    small = (span <= 3).all(axis=1)
    for ox in range(4):
        for oy in range(4):
            for oz in range(4):
                m = small & (span[:, 0] >= ox) & (span[:, 1] >= oy) & (span[:, 2] >= oz)
                grid[lo[m, 0] + ox, lo[m, 1] + oy, lo[m, 2] + oz] = True
    for t in np.nonzero(~small)[0]:
        grid[lo[t, 0]:hi[t, 0] + 1, lo[t, 1]:hi[t, 1] + 1, lo[t, 2]:hi[t, 2] + 1] = True
    return grid


def _shift_or(g: np.ndarray, dx: int, dy: int, dz: int) -> np.ndarray:
    """out[i] = g[i - d] (False outside)."""
    out = np.zeros_like(g)
    sx = slice(max(dx, 0), g.shape[0] + min(dx, 0)); tx = slice(max(-dx, 0), g.shape[0] + min(-dx, 0))
    sy = slice(max(dy, 0), g.shape[1] + min(dy, 0)); ty = slice(max(-dy, 0), g.shape[1] + min(-dy, 0))
    sz = slice(max(dz, 0), g.shape[2] + min(dz, 0)); tz = slice(max(-dz, 0), g.shape[2] + min(-dz, 0))
    out[sx, sy, sz] = g[tx, ty, tz]
    return out


def _sibling_expand(g: np.ndarray) -> np.ndarray:
    """All 8 children of every parent that has at least one set child, within bounds."""
    dims = g.shape
    pd = tuple((d + 1) // 2 for d in dims)
    pad = np.zeros(tuple(2 * p for p in pd), dtype=bool)
    pad[:dims[0], :dims[1], :dims[2]] = g
    par = np.zeros(pd, dtype=bool)
    for i in (0, 1):
        for j in (0, 1):
            for k in (0, 1):
                par |= pad[i::2, j::2, k::2]
    full = np.broadcast_to(par[:, None, :, None, :, None], (pd[0], 2, pd[1], 2, pd[2], 2))
    return full.reshape(2 * pd[0], 2 * pd[1], 2 * pd[2])[:dims[0], :dims[1], :dims[2]]


def add_halo_blocks_with_siblings(active: np.ndarray, layers: int) -> np.ndarray:
    """add_halo_blocks_with_siblings! (domain_topology.jl:54-99)."""
    for _ in range(layers):
        # 26-neighbourhood dilation, as three separable 1-D dilations
        dil = active
        for ax in range(3):
            d = [0, 0, 0]
            d[ax] = 1
            fwd = _shift_or(dil, *d)
            d[ax] = -1
            dil = dil | fwd | _shift_or(dil, *d)
        new = dil & ~active
        sib = _sibling_expand(new) & ~active & ~new
        active = active | new | sib
    return active


def ensure_complete_parent_coverage(active: np.ndarray, max_iterations: int = 10) -> np.ndarray:
    """ensure_complete_parent_coverage! (domain_topology.jl:101-133)."""
    for _ in range(max_iterations):
        nxt = active | _sibling_expand(active)
        if (nxt == active).all():
            break
        active = nxt
    return active


def build_neighbor_table(coords: np.ndarray, dims: tuple) -> np.ndarray:
    """build_neighbor_table (domain_topology.jl:135-160): (nb,27) int32, 1-based."""
    nb = coords.shape[0]
    ptr = _dense_ptr(coords, dims)
    table = np.zeros((nb, 27), dtype=np.int32)
    d = 0
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                n = coords + np.array([dx, dy, dz], dtype=np.int32)
                ok = (n[:, 0] >= 1) & (n[:, 0] <= dims[0]) & (n[:, 1] >= 1) & (n[:, 1] <= dims[1]) & \
                     (n[:, 2] >= 1) & (n[:, 2] <= dims[2])
                v = np.zeros(nb, dtype=np.int32)
                v[ok] = ptr[n[ok, 0] - 1, n[ok, 1] - 1, n[ok, 2] - 1]
                table[:, d] = v
                d += 1
    return table


def should_use_bouzidi(level_id: int, num_levels: int, method: str, bouzidi_levels: int) -> bool:
    """should_use_bouzidi (bouzidi_common.jl:28-34)."""
    return method == "bouzidi" and level_id > num_levels - bouzidi_levels


def _check_gpu(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc}): {N.hip_lib().lb_setup_gpu_last_error().decode()}")


def bouzidi_qmap(coords: np.ndarray, dense_c: np.ndarray, dims: tuple, tris: np.ndarray, dx: float, off: np.ndarray,
                 gpu_device: int | None = None) -> tuple:
    """compute_bouzidi_qmap_sparse (bouzidi_setup.jl:64-167): (cells (n,4),
    q bits (n,27), triangle ids (n,27)) in the canonical order, on the CPU
    (lbs_bouzidi_qmap) or on a HIP device (lb_bouzidi_qmap_gpu)."""
    nb = coords.shape[0]
    ci = N.ptr(coords, C.c_int32)
    gi = N.ptr(dense_c, C.c_int32)
    cells, q, tri = N.i32p(), N.u16p(), N.i32p()
    n = C.c_int64()
    args = (nb, ci, gi, dims[0], dims[1], dims[2], N.ptr(tris, C.c_double), tris.shape[0], dx,
            N.ptr(off, C.c_double), C.byref(cells), C.byref(q), C.byref(tri), C.byref(n))
    if gpu_device is None:
        lib = N.setup_lib()
        N.check_setup(lib.lbs_bouzidi_qmap(*args), "lbs_bouzidi_qmap")
        free = lib.lbs_free
    else:
        lib = N.hip_lib()
        rc = lib.lb_bouzidi_qmap_gpu(int(gpu_device), *args)
        if rc != 0:
            raise RuntimeError(f"lb_bouzidi_qmap_gpu failed ({rc}): {lib.lb_setup_gpu_last_error().decode()}")
        free = lib.lb_free
    try:
        m = n.value
        out = (np.ctypeslib.as_array(cells, shape=(max(m, 1) * 4,))[:m * 4].copy().reshape(m, 4),
               np.ctypeslib.as_array(q, shape=(max(m, 1) * 27,))[:m * 27].copy().reshape(m, 27),
               np.ctypeslib.as_array(tri, shape=(max(m, 1) * 27,))[:m * 27].copy().reshape(m, 27))
    finally:
        free(cells); free(q); free(tri)
    return out


def setup_multilevel_domain(cfg: CaseConfig, mesh: SolverMesh, params: DomainParameters,
                            verbose: bool = False, topology_only: bool = False, gpu_device: int | None = None) -> list:
    """setup_multilevel_domain(mesh, params) (domain.jl:1-246).  With
    topology_only, only the block sets are built and the block count of each
    level is returned (used to size serving requests).  With gpu_device, the
    voxelization, the wall distances and the Bouzidi Q-map run on that HIP
    device (lb_voxelize_gpu, lb_wall_distances_gpu, lb_bouzidi_qmap_gpu: the
    same bits as the CPU restatement)."""
    lib = N.setup_lib()
    off = np.ascontiguousarray(params.mesh_offset, dtype=np.float64)
    placed_min = np.array(params.mesh_min) + off
    placed_max = np.array(params.mesh_max) + off
    wake_start_x = placed_max[0] - params.reference_length * 0.1
    wake_end_x = placed_max[0] + params.reference_length * cfg.wake_length
    wcy = (placed_min[1] + placed_max[1]) / 2.0
    wcz = (placed_min[2] + placed_max[2]) / 2.0
    ww = (placed_max[1] - placed_min[1]) * cfg.wake_width_factor
    wh = (placed_max[2] - placed_min[2]) * cfg.wake_height_factor
    wmin_y, wmax_y = wcy - ww / 2.0, wcy + ww / 2.0
    wmin_z, wmax_z = wcz - wh / 2.0, wcz + wh / 2.0
    tris = np.ascontiguousarray(mesh.triangles, dtype=np.float64)
    grids = []
    prev_grid = None
    prev_coords_list = []
    for lvl in range(1, params.num_levels + 1):
        scale = 2 ** (lvl - 1)
        dx = params.dx_coarse / scale
        dims = (params.bx_max * scale, params.by_max * scale, params.bz_max * scale)
        if lvl == 1:
            active = np.ones(dims, dtype=bool)
        else:
            if cfg.refinement_strategy != "geometry_first":
                raise NotImplementedError("refinement strategy 'topology' is not supported")
            active = get_active_blocks_for_level(mesh, dx, off, dims)
            prev_coords = prev_coords_list[-1]
            if cfg.wake_enabled:
                prev_bs = BLOCK_SIZE * (params.dx_coarse / 2 ** (lvl - 2))
                c = prev_coords.astype(np.float64)
                bmin = (c - 1) * prev_bs
                bmax = c * prev_bs
                ov = (bmin[:, 0] <= wake_end_x) & (bmax[:, 0] >= wake_start_x) & \
                     (bmin[:, 1] <= wmax_y) & (bmax[:, 1] >= wmin_y) & \
                     (bmin[:, 2] <= wmax_z) & (bmax[:, 2] >= wmin_z)
                # the 8 children of every previous-level block the wake box touches
                kids = np.array([(dbx, dby, dbz) for dbz in (0, 1) for dby in (0, 1) for dbx in (0, 1)])
                f = (2 * prev_coords[ov].astype(np.int64) - 1)[:, None, :] + kids[None, :, :]
                f = f.reshape(-1, 3)
                f = f[((f >= 1) & (f <= np.asarray(dims))).all(axis=1)]
                active[f[:, 0] - 1, f[:, 1] - 1, f[:, 2] - 1] = True
            # orphan filter: parent (b+1)÷2 must be active on the previous level
            pexp = np.repeat(np.repeat(np.repeat(prev_grid, 2, 0), 2, 1), 2, 2)[:dims[0], :dims[1], :dims[2]]
            active &= pexp
        active = add_halo_blocks_with_siblings(active, cfg.refinement_margin)
        active = ensure_complete_parent_coverage(active)
        idx = np.argwhere(active)   # lexicographic (bx, by, bz) order, bx most significant
        coords = np.ascontiguousarray(idx.astype(np.int32) + 1)
        nb = coords.shape[0]
        prev_coords_list.append(coords)
        if topology_only:
            grids.append(nb)
            prev_grid = active
            continue
        nbt = build_neighbor_table(coords, dims)
        dense = np.asfortranarray(_dense_ptr(coords, dims))
        dense_c = np.ascontiguousarray(dense.ravel(order="F"))
        obst = np.zeros((nb, 512), dtype=np.uint8)
        sponge = np.zeros((nb, 512), dtype=np.float32)
        wall = np.full((nb, 512), 100.0, dtype=np.float32)
        ci = N.ptr(coords, C.c_int32)
        gi = N.ptr(dense_c, C.c_int32)
        if gpu_device is None:
            N.check_setup(lib.lbs_voxelize(nb, ci, gi, dims[0], dims[1], dims[2], N.ptr(tris, C.c_double),
                                           tris.shape[0], dx, N.ptr(off, C.c_double), N.ptr(obst, C.c_uint8)),
                          "lbs_voxelize")
        else:
            _check_gpu(N.hip_lib().lb_voxelize_gpu(gpu_device, nb, ci, gi, dims[0], dims[1], dims[2],
                                                   N.ptr(tris, C.c_double), tris.shape[0], dx,
                                                   N.ptr(off, C.c_double), N.ptr(obst, C.c_uint8)),
                       "lb_voxelize_gpu")
        filled = lib.lbs_flood_fill(nb, ci, gi, dims[0], dims[1], dims[2], N.ptr(obst, C.c_uint8))
        N.check_setup(lib.lbs_sponge(nb, ci, dx, params.domain_size[0], params.domain_size[1],
                                     params.domain_size[2], float(cfg.sponge_thickness),
                                     int(cfg.symmetric_analysis), N.ptr(sponge, C.c_float)), "lbs_sponge")
        if cfg.wall_model_enabled:
            if gpu_device is None:
                lib.lbs_wall_distances(nb, ci, gi, dims[0], dims[1], dims[2], N.ptr(obst, C.c_uint8), dx,
                                       N.ptr(wall, C.c_float))
            else:
                _check_gpu(N.hip_lib().lb_wall_distances_gpu(gpu_device, nb, ci, gi, dims[0], dims[1], dims[2],
                                                             N.ptr(obst, C.c_uint8), dx, N.ptr(wall, C.c_float),
                                                             None), "lb_wall_distances_gpu")
        lev = BlockLevel(level_id=lvl, dx=f32(dx), tau=params.tau_levels[lvl - 1],
                         tau_g=params.tau_g_levels[lvl - 1] if params.tau_g_levels else 0.0,
                         coords=coords, neighbor_table=nbt,
                         block_pointer=_dense_ptr(coords, tuple(int(v) for v in coords.max(axis=0))) if nb else
                         np.zeros((0, 0, 0), np.int32),
                         grid_dims=dims, obstacle=obst, sponge=sponge, wall_dist=wall)
        if should_use_bouzidi(lvl, params.num_levels, cfg.boundary_method, cfg.bouzidi_levels):
            lev.bouzidi_cells, lev.bouzidi_q, lev.bouzidi_tri = bouzidi_qmap(
                coords, dense_c, dims, tris, dx, off, gpu_device)
        if verbose:
            print(f"[setup] level {lvl}: {nb} blocks ({nb * 512 / 1e6:.2f} M cells), filled {filled}, "
                  f"bouzidi cells {lev.n_boundary_cells}")
        grids.append(lev)
        prev_grid = active
    return grids
