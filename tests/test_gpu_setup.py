"""Setup on the GPU: the Bouzidi Q-map of lb_bouzidi_qmap_gpu equals the CPU
restatement (lbs_bouzidi_qmap) bit for bit -- the boundary-cell list in
canonical order, every q (FP16 bits) and every triangle id -- on a sphere,
the subdivided cube and the rotated NACA wing."""
import numpy as np
import pytest

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
