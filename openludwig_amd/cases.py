"""Named workloads.

sphere_re1e3     D = 1 m icosphere (subdiv 4, 5120 triangles), Re 1e3, coarse
                 64x32x32, 3 levels (2 nested 2:1 refinements, sphere spans 16
                 finest cells), cumulant + positivity limiter, Bouzidi on the
                 finest level, no wall model, u_lat 0.05, ramp 500.
cube_re1e5       1 m cube (12-facet STL refined with minimum facet 0.025 to
                 49,152 triangles), Re 1e5, cumulant + WALE + WMLES, 4 levels,
                 ~50 M active cells in 8^3 blocks, u_lat 0.05.
sphere_re1e5     D = 1 m icosphere (subdiv 6, 81,920 triangles) at Re 1e5 with
                 WMLES, the second body of the precision form.
tiny_sphere      a 2-level sphere case small enough for the CPU oracle.
naca_wing_thermal  extruded NACA 0012 (chord 1 m, span 2 m, 40,400
                 triangles) at alpha 2 deg, compressible form: thermal DDF +
                 O(u^4) equilibrium, u_lat 0.15 (lattice Mach 0.26) with
                 U = 0.7 * 343 m/s; numerics, thermal and domain settings of
                 CASES/Wing_0_deg/config.yaml (positivity limiter, WMLES,
                 adiabatic walls, Pr 0.71), full span (no symmetry plane).
"""
from __future__ import annotations

from .config import CaseConfig, f32
from .geometry import cube, icosphere, mesh_from_triangles, naca0012_wing, subdivide_mesh


def sphere_re1e3(resolution: int = 16, num_levels: int = 3):
    cfg = CaseConfig(
        surface_resolution=resolution, num_levels=num_levels, reference_area_full_model=0.7854,
        reference_chord=1.0, reference_length_for_meshing=1.0, reference_dimension="x",
        fluid_density=1.225, kinematic_viscosity=1.5e-5, flow_velocity=0.015,
        steps=2000, ramp_steps=500, output_freq=10000,
        u_lattice=f32(0.05), c_wale=f32(0.5), tau_min=f32(0.501), inlet_turbulence_intensity=f32(0.0),
        nu_sgs_background=f32(0.0), limiter="positivity", lambda_param=f32(0.1667),
        wall_model_enabled=False, domain_upstream=5.0, domain_downstream=10.0, domain_lateral=3.5,
        domain_height=3.5, sponge_thickness=f32(0.10), refinement_margin=2, wake_enabled=True,
        wake_length=0.25, wake_width_factor=0.1, wake_height_factor=0.1, boundary_method="bouzidi",
        bouzidi_levels=1, q_min_threshold=f32(0.001), forces_enabled=True, moment_center=(0.25, 0.0, 0.0),
        diag_freq=100,
    ).validate()
    return cfg, mesh_from_triangles(icosphere(4, 0.5))


def cube_re1e5(resolution: int = 96, num_levels: int = 4, facet: float = 0.025):
    cfg = CaseConfig(
        surface_resolution=resolution, num_levels=num_levels, reference_area_full_model=1.0,
        reference_chord=1.0, reference_length_for_meshing=1.0, reference_dimension="x",
        minimum_facet_size=facet, fluid_density=1.225, kinematic_viscosity=1.5e-5, flow_velocity=1.5,
        steps=2500, ramp_steps=2000, output_freq=10000,
        u_lattice=f32(0.05), c_wale=f32(0.5), tau_min=f32(0.501), inlet_turbulence_intensity=f32(0.0),
        nu_sgs_background=f32(0.0), limiter="positivity", lambda_param=f32(0.1667),
        wall_model_enabled=True, domain_upstream=5.0, domain_downstream=5.0, domain_lateral=4.0,
        domain_height=4.0, sponge_thickness=f32(0.10), refinement_margin=2, wake_enabled=True,
        wake_length=2.0, wake_width_factor=1.0, wake_height_factor=1.0, boundary_method="bouzidi",
        bouzidi_levels=1, q_min_threshold=f32(0.001), forces_enabled=True, moment_center=(0.25, 0.0, 0.0),
        diag_freq=200,
    ).validate()
    mesh = mesh_from_triangles(cube(1.0))
    if facet > 0:
        mesh = subdivide_mesh(mesh, facet)
    return cfg, mesh


def sphere_re1e5(resolution: int = 96, num_levels: int = 4, subdiv: int = 6):
    cfg, _ = sphere_re1e3(resolution, num_levels)
    cfg = cfg.replace(flow_velocity=1.5, wall_model_enabled=True, ramp_steps=2000, steps=2500, diag_freq=200)
    return cfg, mesh_from_triangles(icosphere(subdiv, 0.5))


def tiny_sphere(resolution: int = 8, num_levels: int = 2, wall_model: bool = False, limiter: str = "positivity",
                ddf: bool = False):
    cfg, mesh = sphere_re1e3(resolution, num_levels)
    cfg = cfg.replace(domain_upstream=1.5, domain_downstream=2.0, domain_lateral=1.0, domain_height=1.0,
                      wall_model_enabled=wall_model, limiter=limiter, compressibility_correction=ddf,
                      flow_velocity=0.015 if not wall_model else 1.5, ramp_steps=20, min_coarse_blocks=1)
    return cfg, mesh


def naca_wing_thermal(resolution: int = 160, num_levels: int = 4, alpha_deg: float = 2.0,
                      n_chord: int = 100, n_span: int = 100):
    cfg = CaseConfig(
        surface_resolution=resolution, num_levels=num_levels, reference_area_full_model=2.0,
        reference_chord=1.0, reference_length_for_meshing=1.0, reference_dimension="x",
        alpha_deg=alpha_deg, fluid_density=1.225, kinematic_viscosity=1.5e-5, flow_velocity=0.7 * 343.0,
        steps=12000, ramp_steps=2000, output_freq=10 ** 9,
        u_lattice=f32(0.15), c_wale=f32(0.5), tau_min=f32(0.501), inlet_turbulence_intensity=f32(0.0),
        nu_sgs_background=f32(0.0), limiter="positivity", lambda_param=f32(0.1667),
        wall_model_enabled=True, compressibility_correction=True, prandtl=0.71, t_inlet=1.0, t_wall=1.0,
        t_initial=1.0, wall_bc="adiabatic",
        domain_upstream=0.75, domain_downstream=1.5, domain_lateral=0.75, domain_height=0.75,
        sponge_thickness=f32(0.10), refinement_margin=2, wake_enabled=False, boundary_method="bouzidi",
        bouzidi_levels=1, q_min_threshold=f32(0.001), forces_enabled=True, moment_center=(0.25, 0.0, 0.0),
        diag_freq=100,
    ).validate()
    return cfg, mesh_from_triangles(naca0012_wing(1.0, 2.0, n_chord, n_span))
