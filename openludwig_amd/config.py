"""Case configuration: the globals that OpenLUDWIG's load_case_configuration
sets from a case's config.yaml (src/io/config_loader.jl:5-330), with the same
keys, defaults and validation errors.  Only what the time step, its setup and
the force reduction read is kept.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

VALID_COLLISION_OPERATORS = {"cumulant"}          # regularized_bgk is outside this library
VALID_CUMULANT_LIMITERS = {"none", "factored", "positivity"}
VALID_WALL_BC = {"adiabatic", "isothermal"}


def f32(x: float) -> float:
    """The Float64 value of Float32(x), as Julia stores a Float32 global."""
    return float(np.float32(x))


@dataclass
class CaseConfig:
    # basic
    stl_files: list = field(default_factory=list)
    stl_scale: float = 1.0
    surface_resolution: int = 200
    num_levels: int = 0
    minimum_facet_size: float = 0.0
    symmetric_analysis: bool = False
    alpha_deg: float = 0.0
    beta_deg: float = 0.0
    reference_area_full_model: float = 0.0
    reference_chord: float = 0.0
    reference_length_for_meshing: float = 0.0
    reference_dimension: str = "x"
    fluid_density: float = 1.225
    kinematic_viscosity: float = 1.5e-5
    flow_velocity: float = 10.0
    steps: int = 1000
    ramp_steps: int = 1500
    output_freq: int = 100
    domain_length_covered: float = 0.0
    # numerics
    collision_operator: str = "cumulant"
    u_lattice: float = f32(0.08)
    c_wale: float = f32(0.20)
    tau_min: float = f32(0.501)
    inlet_turbulence_intensity: float = f32(0.01)
    nu_sgs_background: float = f32(0.0)
    sponge_blend_distributions: bool = False
    temporal_interpolation: bool = True
    omega_bulk: float = f32(1.0)
    omega_3: float = f32(1.0)
    omega_4: float = f32(1.0)
    adaptive_omega_4: bool = True
    lambda_param: float = f32(1.0 / 6.0)
    limiter: str = "factored"
    compressibility_correction: bool = False
    # thermal (DDF follows compressibility_correction, config_loader.jl:225)
    prandtl: float = 0.71
    t_inlet: float = 1.0
    t_wall: float = 1.0
    t_initial: float = 1.0
    wall_bc: str = "adiabatic"
    # high-Re
    auto_levels: bool = False
    max_levels: int = 12
    min_coarse_blocks: int = 4
    wall_model_enabled: bool = False
    # domain
    domain_upstream: float = 0.75
    domain_downstream: float = 1.5
    domain_lateral: float = 0.75
    domain_height: float = 0.75
    sponge_thickness: float = f32(0.10)
    # refinement
    block_size: int = 8
    refinement_margin: int = 2
    refinement_strategy: str = "geometry_first"
    wake_enabled: bool = False
    wake_length: float = 0.25
    wake_width_factor: float = 0.1
    wake_height_factor: float = 0.1
    # boundary
    boundary_method: str = "bouzidi"
    bouzidi_levels: int = 1
    q_min_threshold: float = f32(0.001)
    # forces
    forces_enabled: bool = True
    moment_center: tuple = (0.25, 0.0, 0.0)
    # diagnostics / gpu
    diag_freq: int = 500
    gpu_async_depth: int = 8
    # case folder and outputs (read by the case runner and the result writers)
    case_dir: str = ""
    output_dir: str = "output"
    output_density: bool = True
    output_velocity: bool = True
    output_velocity_magnitude: bool = True
    output_vorticity: bool = True
    output_obstacle: bool = True
    output_level: bool = True
    output_bouzidi: bool = True
    # read and kept for the summaries; they do not change the time step
    tau_safety_factor: float = f32(1.0)
    wall_model_type: str = "equilibrium"
    wall_model_y_plus_target: float = 30.0
    force_output_freq: int = 0
    net_force_export: bool = False
    stability_check: bool = True
    print_tau_warning: bool = True
    use_streams: bool = True
    prefetch_neighbors: bool = True

    @property
    def ddf_enabled(self) -> bool:
        return bool(self.compressibility_correction)

    @property
    def reference_area_config(self) -> float:
        a = self.reference_area_full_model
        return a / 2.0 if self.symmetric_analysis else a

    @property
    def limiter_code(self) -> int:
        return {"none": 0, "factored": 1, "positivity": 2}[self.limiter]

    def validate(self) -> "CaseConfig":
        if self.collision_operator not in VALID_COLLISION_OPERATORS:
            raise ValueError(f'Unknown collision_operator: "{self.collision_operator}". '
                             f"Valid options: {', '.join(sorted(VALID_COLLISION_OPERATORS))}")
        if self.limiter not in VALID_CUMULANT_LIMITERS:
            raise ValueError(f'Unknown cumulant limiter: "{self.limiter}". '
                             f"Valid: {', '.join(sorted(VALID_CUMULANT_LIMITERS))}")
        if self.ddf_enabled and self.wall_bc not in VALID_WALL_BC:
            raise ValueError(f'Unknown thermal wall_bc: "{self.wall_bc}". Valid: adiabatic, isothermal')
        if self.block_size != 8:
            raise ValueError("block_size must be 8 (BLOCK_SIZE, src/geometry/blocks.jl:14)")
        if self.symmetric_analysis and self.beta_deg != 0.0:
            self.beta_deg = 0.0
        return self

    def replace(self, **kw) -> "CaseConfig":
        return dataclasses.replace(self, **kw).validate()


def _get(d, *keys, default=None, required=False):
    cur = d
    for i, k in enumerate(keys):
        if not isinstance(cur, dict) or k not in cur:
            if required or default is None and required:
                raise KeyError("Missing config key: " + " → ".join(keys[: i + 1]))
            return default
        cur = cur[k]
    return default if cur is None and default is not None else cur


def load_case_configuration(config_path: str | Path) -> CaseConfig:
    """load_case_configuration (config_loader.jl:133-330) for one config.yaml."""
    import yaml

    path = Path(config_path)
    if not path.is_file():
        raise FileNotFoundError(f"config.yaml not found: {path}")
    cfg = yaml.safe_load(path.read_text())
    case_dir = path.parent
    raw = _get(cfg, "basic", "stl_file", required=True)
    if isinstance(raw, str):
        names = [raw]
    elif isinstance(raw, list):
        names = [str(s) for s in raw]
    else:
        raise TypeError(f"stl_file must be a string or list of strings, got: {type(raw)}")
    c = CaseConfig()
    c.stl_files = [str(case_dir / n) for n in names]
    c.stl_scale = float(_get(cfg, "basic", "stl_scale", required=True))
    c.surface_resolution = int(_get(cfg, "basic", "surface_resolution", required=True))
    c.num_levels = int(_get(cfg, "basic", "num_levels", required=True))
    c.minimum_facet_size = float(_get(cfg, "basic", "minimum_facet_size", default=0.0))
    c.symmetric_analysis = bool(_get(cfg, "advanced", "refinement", "symmetric_analysis", default=False))
    c.alpha_deg = float(_get(cfg, "basic", "alpha", default=0.0))
    c.beta_deg = float(_get(cfg, "basic", "beta", default=0.0))
    c.reference_area_full_model = float(_get(cfg, "basic", "reference_area_of_full_model", default=0.0))
    c.reference_chord = float(_get(cfg, "basic", "reference_chord", default=0.0))
    c.reference_length_for_meshing = float(_get(cfg, "basic", "reference_length_for_meshing", default=0.0))
    c.reference_dimension = str(_get(cfg, "basic", "reference_dimension", default="x"))
    c.fluid_density = float(_get(cfg, "basic", "fluid", "density", default=1.225))
    c.kinematic_viscosity = float(_get(cfg, "basic", "fluid", "kinematic_viscosity", default=1.5e-5))
    c.flow_velocity = float(_get(cfg, "basic", "flow", "velocity", default=10.0))
    c.domain_length_covered = float(_get(cfg, "basic", "simulation", "domain_length_covered", default=0.0))
    c.steps = int(_get(cfg, "basic", "simulation", "steps", default=0))
    c.ramp_steps = int(_get(cfg, "basic", "simulation", "ramp_steps", required=True))
    c.output_freq = int(_get(cfg, "basic", "simulation", "output_freq", required=True))
    c.case_dir = str(case_dir)
    c.output_dir = str(_get(cfg, "basic", "simulation", "output_dir", required=True))
    of = ("basic", "simulation", "output_fields")
    c.output_density = bool(_get(cfg, *of, "density", default=True))
    c.output_velocity = bool(_get(cfg, *of, "velocity", default=True))
    c.output_velocity_magnitude = bool(_get(cfg, *of, "velocity_magnitude", default=True))
    c.output_vorticity = bool(_get(cfg, *of, "vorticity", default=True))
    c.output_obstacle = bool(_get(cfg, *of, "obstacle", default=True))
    c.output_level = bool(_get(cfg, *of, "level", default=True))
    c.output_bouzidi = bool(_get(cfg, *of, "bouzidi", default=True))
    num = ("advanced", "numerics")
    c.collision_operator = str(_get(cfg, *num, "collision_operator", default="cumulant"))
    c.u_lattice = f32(_get(cfg, *num, "u_lattice", default=0.08))
    c.c_wale = f32(_get(cfg, *num, "c_wale", default=0.20))
    c.tau_min = f32(_get(cfg, *num, "tau_min", default=0.501))
    c.tau_safety_factor = f32(_get(cfg, *num, "tau_safety_factor", default=1.0))
    c.inlet_turbulence_intensity = f32(_get(cfg, *num, "inlet_turbulence_intensity", default=0.01))
    c.nu_sgs_background = f32(_get(cfg, *num, "nu_sgs_background", default=0.0))
    c.sponge_blend_distributions = bool(_get(cfg, *num, "sponge_blend_distributions", default=False))
    c.temporal_interpolation = bool(_get(cfg, *num, "temporal_interpolation", default=True))
    cum = (*num, "cumulant")
    c.omega_bulk = f32(_get(cfg, *cum, "omega_bulk", default=1.0))
    c.omega_3 = f32(_get(cfg, *cum, "omega_3", default=1.0))
    c.omega_4 = f32(_get(cfg, *cum, "omega_4", default=1.0))
    c.adaptive_omega_4 = bool(_get(cfg, *cum, "adaptive_omega_4", default=True))
    c.lambda_param = f32(_get(cfg, *cum, "lambda_param", default=1.0 / 6.0))
    c.limiter = str(_get(cfg, *cum, "limiter", default="factored"))
    c.compressibility_correction = bool(_get(cfg, *num, "compressibility_correction", default=False))
    if c.compressibility_correction:
        th = ("advanced", "thermal")
        c.prandtl = float(_get(cfg, *th, "prandtl", default=0.71))
        c.t_inlet = float(_get(cfg, *th, "t_inlet", default=1.0))
        c.t_wall = float(_get(cfg, *th, "t_wall", default=1.0))
        c.t_initial = float(_get(cfg, *th, "t_initial", default=1.0))
        c.wall_bc = str(_get(cfg, *th, "wall_bc", default="adiabatic"))
    hr = ("advanced", "high_re")
    c.auto_levels = bool(_get(cfg, *hr, "auto_levels", default=False))
    c.max_levels = int(_get(cfg, *hr, "max_levels", default=12))
    c.min_coarse_blocks = int(_get(cfg, *hr, "min_coarse_blocks", default=4))
    c.wall_model_enabled = bool(_get(cfg, *hr, "wall_model", "enabled", default=False))
    c.wall_model_type = str(_get(cfg, *hr, "wall_model", "type", default="equilibrium"))
    c.wall_model_y_plus_target = float(_get(cfg, *hr, "wall_model", "y_plus_target", default=30.0))
    dm = ("advanced", "domain")
    c.domain_upstream = float(_get(cfg, *dm, "upstream", default=0.75))
    c.domain_downstream = float(_get(cfg, *dm, "downstream", default=1.5))
    c.domain_lateral = float(_get(cfg, *dm, "lateral", default=0.75))
    c.domain_height = float(_get(cfg, *dm, "height", default=0.75))
    c.sponge_thickness = f32(_get(cfg, *dm, "sponge_thickness", default=0.10))
    rf = ("advanced", "refinement")
    c.block_size = int(_get(cfg, *rf, "block_size", default=8))
    c.refinement_margin = int(_get(cfg, *rf, "margin", default=2))
    c.refinement_strategy = str(_get(cfg, *rf, "strategy", default="geometry_first"))
    c.wake_enabled = bool(_get(cfg, *rf, "wake_enabled", default=False))
    c.wake_length = float(_get(cfg, *rf, "wake_length", default=0.25))
    c.wake_width_factor = float(_get(cfg, *rf, "wake_width_factor", default=0.1))
    c.wake_height_factor = float(_get(cfg, *rf, "wake_height_factor", default=0.1))
    bd = ("advanced", "boundary")
    c.boundary_method = str(_get(cfg, *bd, "method", default="bouzidi"))
    c.bouzidi_levels = int(_get(cfg, *bd, "bouzidi_levels", default=1))
    c.q_min_threshold = f32(_get(cfg, *bd, "q_min_threshold", default=0.001))
    fo = ("advanced", "forces")
    c.forces_enabled = bool(_get(cfg, *fo, "enabled", default=True))
    mc = _get(cfg, *fo, "moment_center", default=[0.25, 0.0, 0.0])
    c.moment_center = tuple(float(v) for v in mc)
    c.force_output_freq = int(_get(cfg, *fo, "output_freq", default=0))
    c.net_force_export = bool(_get(cfg, *fo, "net_force_export", default=False))
    dg = ("advanced", "diagnostics")
    c.diag_freq = int(_get(cfg, *dg, "freq", default=500))
    c.stability_check = bool(_get(cfg, *dg, "stability_check", default=True))
    c.print_tau_warning = bool(_get(cfg, *dg, "print_tau_warning", default=True))
    gp = ("advanced", "gpu")
    c.gpu_async_depth = int(_get(cfg, *gp, "async_depth", default=8))
    c.use_streams = bool(_get(cfg, *gp, "use_streams", default=True))
    c.prefetch_neighbors = bool(_get(cfg, *gp, "prefetch_neighbors", default=True))
    return c.validate()
