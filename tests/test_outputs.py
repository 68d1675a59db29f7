"""Result files and the case loader (CPU).

The CSV headers and row formats are compared with the strings in the
reference's own source (src/main.jl, src/forces/io.jl) when the reference
checkout is present, so a renamed column or a changed printf format fails
here; the .lbmp layout is checked byte for byte against its documented
format (src/forces/io.jl:230-250) on a hand-built file; the VTU writers are
read back; every CASES/*/config.yaml of the reference loads, with its values
taken from the YAML itself.
"""
from __future__ import annotations

import re
import struct
from pathlib import Path

import numpy as np
import pytest
import yaml

from openludwig_amd import outputs as O
from openludwig_amd.config import CaseConfig, load_case_configuration
from openludwig_amd.runner import steps_for_case

REF = Path("/root/reference")
needs_ref = pytest.mark.skipif(not (REF / "src").is_dir(), reason="reference checkout not present")


def _julia_strings(path: Path) -> list:
    return re.findall(r'"((?:[^"\\]|\\.)*)"', path.read_text())


@needs_ref
def test_csv_headers_match_reference_source():
    main = _julia_strings(REF / "src/main.jl")
    io = _julia_strings(REF / "src/forces/io.jl")
    assert O.CONVERGENCE_HEADER in main
    assert O.FORCES_HEADER in io


@needs_ref
def test_csv_row_formats_match_reference_source():
    main = (REF / "src/main.jl").read_text()
    io = (REF / "src/forces/io.jl").read_text()
    conv_fmt = re.search(r'@printf\(io, "(%d,%.6e,[^"]*)\\n"', main).group(1)
    force_fmt = re.search(r'@printf\(io, "(%d,%.6e,%.6f,[^"]*)\\n"', io).group(1)
    row = O.convergence_row(100, 0.123456789, 12.3456, np.float32(0.05), 0.998765, 1234.56, "0.5000", "N/A", "N/A")
    assert row == (conv_fmt % (100, 0.123456789, 12.3456, float(np.float32(0.05)), 0.998765, 1234.56,
                               "0.5000", "N/A", "N/A")) + "\n"
    fd = {k: float(i) * 1.1 for i, k in enumerate(
        ("Fx", "Fy", "Fz", "Fx_pressure", "Fy_pressure", "Fz_pressure", "Fx_viscous", "Fy_viscous", "Fz_viscous",
         "Mx", "My", "Mz", "Cd", "Cl", "Cs", "Cmx", "Cmy", "Cmz"))}
    row = O.force_row(200, 1.5, fd, np.float32(0.08))
    vals = [fd[k] for k in ("Fx", "Fy", "Fz", "Fx_pressure", "Fy_pressure", "Fz_pressure", "Fx_viscous",
                            "Fy_viscous", "Fz_viscous", "Mx", "My", "Mz", "Cd", "Cl", "Cs", "Cmx", "Cmy", "Cmz")]
    assert row == (force_fmt % (200, 1.5, float(np.float32(0.08)), *vals)) + "\n"


def test_csv_files_round_trip(tmp_path):
    p = tmp_path / "convergence.csv"
    O.write_convergence_header(p)
    O.append_convergence(p, 500, 2.5e-3, 1.0, 0.05, 0.99, 100.0, "1.2345", "0.0001", "-0.0002")
    head, rows = O.read_csv(p)
    assert head == O.CONVERGENCE_HEADER.split(",")
    assert rows == [["500", "2.500000e-03", "1.000000e+00", "0.0500", "0.990000", "100.0", "1.2345", "0.0001",
                     "-0.0002"]]


def _lbmp_by_hand(tris, normals, p, sx, sy, sz, rho, u, names, ids):
    """The documented layout, assembled value by value with struct."""
    out = bytearray(b"LBMP") + struct.pack("<III", 1, len(p), len(names))
    for n in names:
        b = n.encode()
        out += struct.pack("<I", len(b)) + b
    for t in tris:
        for v in t:
            out += struct.pack("<3f", *(float(np.float32(c)) for c in v))
    for nrm in normals:
        out += struct.pack("<3f", *(float(np.float32(c)) for c in nrm))
    q = np.float32(0.5 * rho * u * u)
    mag = [np.sqrt(np.float32(a) * np.float32(a) + np.float32(b) * np.float32(b) + np.float32(c) * np.float32(c))
           for a, b, c in zip(sx, sy, sz)]
    out += struct.pack(f"<{len(p)}f", *(float(np.float32(v)) for v in p))
    out += struct.pack(f"<{len(p)}f", *(float(v) for v in mag))
    out += struct.pack(f"<{len(p)}f", *(float(np.float32(v) / q) for v in p))
    out += struct.pack(f"<{len(p)}i", *ids)
    return bytes(out)


def test_lbmp_bytes_match_documented_layout(tmp_path):
    rng = np.random.default_rng(7)
    n = 6
    tris = rng.normal(size=(n, 3, 3))
    normals = rng.normal(size=(n, 3))
    p = rng.normal(size=n).astype(np.float32) * 50
    sx, sy, sz = (rng.normal(size=n).astype(np.float32) for _ in range(3))
    names = ["wing", "flap", "wing", "wing", "pylon", "flap"]
    path = tmp_path / "s.lbmp"
    O.export_surface_pressure_bin(path, tris, normals, p, (sx, sy, sz), 1.225, 240.1, object_names=names)
    want = _lbmp_by_hand(tris, normals, p, sx, sy, sz, 1.225, 240.1, ["wing", "flap", "pylon"], [1, 2, 1, 1, 3, 2])
    assert path.read_bytes() == want
    r = O.read_surface_pressure_bin(path)
    assert r["names"] == ["wing", "flap", "pylon"] and list(r["object_id"]) == [1, 2, 1, 1, 3, 2]
    # single object: the name table holds "default"
    O.export_surface_pressure_bin(path, tris, normals, p, (sx, sy, sz), 1.225, 240.1)
    assert O.read_surface_pressure_bin(path)["names"] == ["default"]


def test_surface_vtu_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n = 9
    tris = rng.normal(size=(n, 3, 3))
    normals = rng.normal(size=(n, 3))
    areas = rng.random(n)
    p = rng.normal(size=n).astype(np.float32)
    sh = tuple(rng.normal(size=n).astype(np.float32) for _ in range(3))
    f = O.save_surface_vtk(tmp_path / "surface_000010", tris, normals, areas, p, sh, offset=(1.0, 2.0, 3.0),
                           object_names=["a"] * 4 + ["b"] * 5)
    v = O.read_vtu(f)
    assert v["n_points"] == 3 * n and v["n_cells"] == n
    assert np.array_equal(v["points"], (tris + [1.0, 2.0, 3.0]).reshape(-1, 3))
    assert np.array_equal(v["connectivity"], np.arange(3 * n)) and np.all(v["types"] == O.VTK_TRIANGLE)
    cd = v["cell_data"]
    assert np.array_equal(cd["Pressure"], p) and np.array_equal(cd["ShearStress"], np.stack(sh, 1))
    assert np.array_equal(cd["ShearMag"], np.sqrt(sh[0] ** 2 + sh[1] ** 2 + sh[2] ** 2))
    assert list(cd["ObjectID"]) == [1] * 4 + [2] * 5
    assert np.array_equal(cd["Area"], areas.astype(np.float32))


def test_flow_vtu_block_selection_and_voxels(tmp_path):
    """A coarse block whose 8 children exist is left out; every voxel is an
    axis-aligned cube of side dx at its cell's position."""
    rng = np.random.default_rng(1)
    c1 = np.array([[1, 1, 1], [2, 1, 1], [1, 2, 1]])
    c2 = np.array([[x, y, z] for z in (1, 2) for y in (1, 2) for x in (1, 2)] + [[3, 1, 1]])
    lv = []
    for c, dx in ((c1, 0.5), (c2, 0.25)):
        n = len(c) * 512
        lv.append({"coords": c, "dx": dx, "rho": rng.random(n).astype(np.float32),
                   "vel": rng.random((n, 3)).astype(np.float32), "obstacle": (rng.random(n) < 0.1).astype(np.uint8)})
    lv[0]["rho"][600] = np.nan
    f = O.export_flow_vtu(tmp_path / "flow_000100", lv, CaseConfig(), compress=False)
    v = O.read_vtu(f)
    assert v["n_cells"] == (2 + 9) * 512                     # coarse (1,1,1) is fully refined
    cd = v["cell_data"]
    assert list(np.unique(cd["Level"])) == [1, 2] and (cd["Level"] == 1).sum() == 2 * 512
    # coarse block 2 (index 1) comes first; its cell 100 = (4, 4, 1)
    i = 100
    assert cd["Density"][i] == lv[0]["rho"][512 + i]
    assert np.array_equal(cd["Velocity"][i], lv[0]["vel"][512 + i])
    assert cd["Obstacle"][i] == lv[0]["obstacle"][512 + i]
    assert cd["Density"][600 - 512] == 0.0                   # the NaN is written as 0
    pts = v["points"][v["connectivity"].reshape(-1, 8)]        # (ncells, 8, 3)
    lo = pts.min(axis=1)
    side = pts.max(axis=1) - lo
    dx = np.where(cd["Level"] == 1, 0.5, 0.25)
    assert np.allclose(side, dx[:, None])
    # cell (x, y, z) of coarse block (2,1,1) starts at ((8 + x) dx, y dx, z dx)
    assert np.allclose(lo[i], [(8 + 4) * 0.5, 4 * 0.5, 1 * 0.5])
    # VTK_VOXEL corner order: x fastest, then y, then z
    d = pts[i] - lo[i]
    assert np.allclose(d / 0.5, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1], [1, 1, 1]])


@needs_ref
def test_every_reference_case_config_loads():
    cfgs = sorted((REF / "CASES").glob("*/config.yaml"))
    assert len(cfgs) >= 10
    for p in cfgs:
        raw = yaml.safe_load(p.read_text())
        c = load_case_configuration(p)
        b, adv = raw["basic"], raw.get("advanced", {})
        assert c.surface_resolution == b["surface_resolution"]
        assert c.num_levels == b["num_levels"]
        assert c.output_dir == b["simulation"]["output_dir"]
        assert c.ramp_steps == b["simulation"]["ramp_steps"]
        stl = b["stl_file"] if isinstance(b["stl_file"], list) else [b["stl_file"]]
        assert [Path(s).name for s in c.stl_files] == [Path(s).name for s in stl]
        num = adv.get("numerics", {})
        if "u_lattice" in num:
            assert c.u_lattice == float(np.float32(num["u_lattice"]))
        lim = num.get("cumulant", {}).get("limiter")
        if lim is not None:
            assert c.limiter == lim
        assert c.compressibility_correction == bool(num.get("compressibility_correction", False))
        assert c.ddf_enabled == c.compressibility_correction


def test_steps_from_domain_length_covered():
    """STEPS = ceil(L_cov * L_x / U / dt) rounded up to a multiple of the
    diagnostic interval (main.jl:84-98)."""
    class P:
        domain_size = (12.0, 4.0, 4.0)
        u_physical = 0.015
        time_scale = 0.01
    cfg = CaseConfig(domain_length_covered=3.0, diag_freq=500, steps=0)
    # 3 * 12 / 0.015 = 2400 s -> 240000 steps, already a multiple of 500
    assert steps_for_case(cfg, P) == 240000
    cfg = CaseConfig(domain_length_covered=1.0, diag_freq=700, steps=0)
    assert steps_for_case(cfg, P) == 80500                 # 80000 -> next multiple of 700
    assert steps_for_case(CaseConfig(steps=123), P) == 123
    with pytest.raises(ValueError):
        steps_for_case(CaseConfig(steps=0), P)
