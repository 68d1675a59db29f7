"""The C-ABI libraries load and export every function their headers declare
(no device calls: runs on the CPU)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from openludwig_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def declared(header: Path) -> list:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lb[s]?_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader", [("ludwig_hip.h", N.hip_lib), ("ludwig_setup.h", N.setup_lib)])
def test_exports_match_header(header, loader):
    names = declared(ROOT / "include" / header)
    assert len(names) >= 8
    lib = loader()
    for n in names:
        assert hasattr(lib, n), f"{n} declared in {header} but not exported"
    so = Path(lib._name)
    out = subprocess.run(["nm", "-D", "--defined-only", str(so)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (lb[s]?_\w+)", out))
    assert set(names) <= exported


def test_bindings_cover_header():
    names = declared(ROOT / "include" / "ludwig_hip.h")
    assert set(names) <= set(N.HIP_EXPORTS), set(names) - set(N.HIP_EXPORTS)
    names = declared(ROOT / "include" / "ludwig_setup.h")
    assert set(names) <= set(N.SETUP_EXPORTS), set(names) - set(N.SETUP_EXPORTS)


def test_error_paths_without_device():
    lib = N.hip_lib()
    assert lib.lb_version().decode().startswith("ludwig_hip")
    # null config / context are rejected before any device call
    ctx = C.c_void_p()
    assert lib.lb_create(C.byref(ctx), 0, None) == -1
    cfg = N.LbConfig()
    cfg.n_levels = 0
    assert lib.lb_create(C.byref(ctx), 0, C.byref(cfg)) == -1
    assert lib.lb_step_batch(None, 1, 1, 0.0, None) == -1
    assert lib.lb_synchronize(None) == -1
    assert lib.lb_device_bytes(None) == 0


def test_footprint_plans_without_gpu():
    """lb_footprint runs the upload code in counting mode: no device needed,
    FP16 storage needs less than FP32, and the count grows with the mesh.
    Attaching to an engine adds the ramp table (4 bytes per batch) and the
    80-byte force-sample accumulator."""
    from openludwig_amd import _native as N
    from openludwig_amd import cases
    from openludwig_amd.solver import build_case, footprint
    cfg, mesh = cases.tiny_sphere(8, 2)
    grids, params, mesh = build_case(cfg, mesh)
    b32 = footprint(cfg, grids, params, mesh, N.LB_STORAGE_FP32)
    b16 = footprint(cfg, grids, params, mesh, N.LB_STORAGE_FP16)
    n = sum(g.n_cells for g in grids)
    assert 216 * n < b32 < 400 * n
    assert 108 * n < b16 < b32
    assert footprint(cfg, grids, params, mesh, N.LB_STORAGE_FP16, 10) == b16 + 40 + 80
