"""Setup on the GPU against the CPU restatement, bit for bit, on a sphere,
the subdivided cube and the rotated NACA wing: the Bouzidi Q-map of
lb_bouzidi_qmap_gpu (the boundary-cell list in canonical order, every q as
FP16 bits and every triangle id), the SAT voxelization of lb_voxelize_gpu
(every obstacle flag of every level) and the wall distances of
lb_wall_distances_gpu (every FP32 value)."""
import ctypes as C

import numpy as np
import pytest

from openludwig_amd import _native as N
from openludwig_amd import cases
from openludwig_amd.domain import bouzidi_qmap, compute_domain_from_mesh, setup_multilevel_domain
from openludwig_amd.domain import _dense_ptr
from openludwig_amd.geometry import rotate_mesh


def _finest(cfg, mesh):
    mesh = rotate_mesh(mesh, cfg.alpha_deg, cfg.beta_deg)
    params = compute_domain_from_mesh(cfg, mesh.min_bounds, mesh.max_bounds)
    grids = setup_multilevel_domain(cfg, mesh, params)
    g = grids[-1]
    dense = np.ascontiguousarray(np.asfortranarray(_dense_ptr(g.coords, g.grid_dims)).ravel(order="F"))
    tris = np.ascontiguousarray(mesh.triangles, dtype=np.float64)
    off = np.ascontiguousarray(params.mesh_offset, dtype=np.float64)
    return g, dense, tris, off


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sphere", "cube", "wing"])
def test_qmap_gpu_equals_cpu(hip_device, case):
    if case == "sphere":
        cfg, mesh = cases.sphere_re1e3(16, 3)
    elif case == "cube":
        cfg, mesh = cases.cube_re1e5(48)
    else:
        cfg, mesh = cases.naca_wing_thermal(48, 3, n_chord=60, n_span=30)
    g, dense, tris, off = _finest(cfg, mesh)
    cpu = bouzidi_qmap(g.coords, dense, g.grid_dims, tris, float(g.dx), off, None)
    gpu = bouzidi_qmap(g.coords, dense, g.grid_dims, tris, float(g.dx), off, hip_device)
    assert cpu[0].shape[0] > 100
    for a, b, what in zip(cpu, gpu, ("cells", "q", "tri")):
        assert a.shape == b.shape, f"{case} {what}: {a.shape} vs {b.shape}"
        assert np.array_equal(a, b), f"{case} {what}: {np.count_nonzero(a != b)} entries differ"


def _levels(cfg, mesh):
    mesh = rotate_mesh(mesh, cfg.alpha_deg, cfg.beta_deg)
    params = compute_domain_from_mesh(cfg, mesh.min_bounds, mesh.max_bounds)
    return mesh, params, setup_multilevel_domain(cfg, mesh, params)


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["sphere", "cube", "wing"])
def test_voxelize_and_wall_distance_gpu_equal_cpu(hip_device, case):
    """Every level: the SAT shell of the GPU equals the CPU's, and the whole
    set-up (shell, flood fill, wall distances) with gpu_device gives the same
    obstacle and wall-distance arrays as the CPU set-up."""
    if case == "sphere":
        cfg, mesh = cases.sphere_re1e5(16, 3)
    elif case == "cube":
        cfg, mesh = cases.cube_re1e5(48)
    else:
        cfg, mesh = cases.naca_wing_thermal(48, 3, n_chord=60, n_span=30)
    assert cfg.wall_model_enabled
    mesh_r, params, cpu_levels = _levels(cfg, mesh)
    gpu_levels = setup_multilevel_domain(cfg, mesh_r, params, gpu_device=hip_device)
    tris = np.ascontiguousarray(mesh_r.triangles, dtype=np.float64)
    off = np.ascontiguousarray(params.mesh_offset, dtype=np.float64)
    lib, hip = N.setup_lib(), N.hip_lib()
    n_wall = 0
    for a, b in zip(cpu_levels, gpu_levels):
        assert np.array_equal(a.obstacle, b.obstacle), f"{case} level {a.level_id}: obstacle differs"
        assert np.array_equal(a.wall_dist, b.wall_dist), f"{case} level {a.level_id}: wall distance differs"
        n_wall += int(np.count_nonzero(a.wall_dist < 100.0))
        # the SAT shell alone (before the flood fill)
        dense = np.ascontiguousarray(np.asfortranarray(_dense_ptr(a.coords, a.grid_dims)).ravel(order="F"))
        s_cpu = np.zeros((a.n_blocks, 512), np.uint8)
        s_gpu = np.zeros((a.n_blocks, 512), np.uint8)
        args = (a.n_blocks, N.ptr(a.coords, C.c_int32), N.ptr(dense, C.c_int32), *a.grid_dims,
                N.ptr(tris, C.c_double), tris.shape[0], float(params.dx_coarse / 2 ** (a.level_id - 1)),
                N.ptr(off, C.c_double))
        N.check_setup(lib.lbs_voxelize(*args, N.ptr(s_cpu, C.c_uint8)), "lbs_voxelize")
        assert hip.lb_voxelize_gpu(hip_device, *args, N.ptr(s_gpu, C.c_uint8)) == 0
        assert np.array_equal(s_cpu, s_gpu), f"{case} level {a.level_id}: {np.count_nonzero(s_cpu != s_gpu)} cells"
    assert n_wall > 100
