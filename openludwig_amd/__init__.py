"""openludwig_amd -- the OpenLUDWIG lattice-Boltzmann time step on MI355X (gfx950).

The product is libludwig_hip.so (include/ludwig_hip.h): hand-written HIP
kernels behind a C ABI.  This package is its Python host: case configuration,
geometry and block-table setup, and the driver loop that calls the C ABI.
"""
from .config import CaseConfig, load_case_configuration  # noqa: F401

__all__ = ["CaseConfig", "load_case_configuration"]
