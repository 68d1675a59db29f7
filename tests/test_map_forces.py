"""Net-force exports and the FE load mapper (CPU).

Pins: a force at the centre of an equilateral node triangle splits into
thirds (symmetry + minimum norm); on random clouds the nodal loads keep the
total force and the total moment of the triangle forces; the pseudo-inverse
solution is the minimum-norm least-squares one (numpy lstsq); the output
formats equal the printf strings in the reference source when the checkout
is present; the inverse rotation of the net-force CSV undoes R = Rz Ry.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

from openludwig_amd import map_forces as MF
from openludwig_amd import outputs as O

REF = Path("/root/reference")
needs_ref = pytest.mark.skipif(not (REF / "src").is_dir(), reason="reference checkout not present")


def test_equilateral_split_in_thirds():
    c = np.array([0.3, -0.2, 1.0])
    ang = np.deg2rad([90.0, 210.0, 330.0])
    nodes = c + 0.5 * np.stack([np.cos(ang), np.sin(ang), np.zeros(3)], 1)
    F = np.array([1.0, -2.0, 0.5])
    f = MF.distribute_force(F, c, nodes)
    assert np.allclose(f, F / 3.0, atol=1e-14)


def test_min_norm_solution_and_equilibrium():
    rng = np.random.default_rng(0)
    for _ in range(20):
        c = rng.normal(size=3)
        nodes = c + rng.normal(size=(3, 3))
        F = rng.normal(size=3)
        f = MF.distribute_force(F, c, nodes)
        A = np.zeros((6, 9))
        for i in range(3):
            A[0:3, 3 * i:3 * i + 3] = np.eye(3)
            A[3:6, 3 * i:3 * i + 3] = MF.skew(nodes[i] - c)
        b = np.concatenate([F, np.zeros(3)])
        ref = np.linalg.lstsq(A, b, rcond=None)[0]
        assert np.allclose(f.ravel(), ref, atol=1e-12)
        assert np.allclose(f.sum(0), F, atol=1e-12)
        assert np.allclose(sum(np.cross(nodes[i] - c, f[i]) for i in range(3)), 0.0, atol=1e-12)
    assert np.array_equal(MF.distribute_force(F, c, nodes[:1]), F.reshape(1, 3))


def _write_inputs(tmp_path, rng, n_nodes=40, n_tri=200):
    objs = ["wing", "flap"]
    nodes = [(i + 1, objs[i % 2], *rng.normal(size=3)) for i in range(n_nodes)]
    forces = [(i + 1, objs[i % 2], *rng.normal(size=3), *rng.normal(size=3)) for i in range(n_tri)]
    (tmp_path / "nodes.csv").write_text("node_id,object_name,cx,cy,cz\n" +
                                        "".join("%d,%s,%.17g,%.17g,%.17g\n" % n for n in nodes))
    (tmp_path / "net.csv").write_text("triangle_id,object_name,cx,cy,cz,Fx,Fy,Fz\n" +
                                      "".join("%d,%s,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n" % f for f in forces))
    return nodes, forces


def test_mapping_keeps_force_and_moment(tmp_path):
    rng = np.random.default_rng(5)
    nodes, forces = _write_inputs(tmp_path, rng)
    res = MF.run(tmp_path / "nodes.csv", tmp_path / "net.csv", tmp_path / "out", log=lambda *a: None)
    for obj, r in res["objects"].items():
        scale = max(1.0, np.abs(r["F_tri"]).max(), np.abs(r["M_tri"]).max())
        assert np.allclose(r["F_tri"], r["F_nod"], atol=1e-9 * scale), obj
        assert np.allclose(r["M_tri"], r["M_nod"], atol=1e-9 * scale), obj
    head = (tmp_path / "out" / "nodal_loads.csv").read_text().splitlines()
    assert head[0] == "node_id,object_name,cx,cy,cz,Fx,Fy,Fz" and len(head) == len(nodes) + 1
    assert "**PASS**" in (tmp_path / "out" / "Nodal_loads_summary.md").read_text()


@needs_ref
def test_mapper_formats_match_reference_source():
    src = (REF / "src/PREPOST/map_forces_to_nodes.jl").read_text()
    assert '"node_id,object_name,cx,cy,cz,Fx,Fy,Fz"' in src
    assert '"%d,%s,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e\\n"' in src
    assert '"| Fx [N]   | %+.6e | %+.6e | %+.6e |\\n"' in src
    io = (REF / "src/forces/io.jl").read_text()
    assert '"triangle_id,object_name,cx,cy,cz,Fx,Fy,Fz"' in io
    assert '"%d,%s,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e\\n"' in io
    assert '"          %.8e %.8e %.8e\\n"' in io and '"          %.6e %.6e %.6e\\n"' in io


def test_summary_table_rows():
    r = {"F_tri": np.array([1.0, 2.0, 3.0]), "F_nod": np.array([1.0, 2.0, 3.0 - 1e-9]),
         "M_tri": np.zeros(3), "M_nod": np.zeros(3)}
    rows = MF._table(r)
    assert rows[2] == "| Fx [N]   | +1.000000e+00 | +1.000000e+00 | +0.000000e+00 |"
    assert rows[4] == "| Fz [N]   | +3.000000e+00 | +3.000000e+00 | +1.000000e-09 |"


def test_net_force_csv_inverse_rotation(tmp_path):
    """Forces given in the rotated frame come back in the STL frame: R^T R = I
    up to Float32 rounding of the matrix and the products."""
    rng = np.random.default_rng(2)
    n = 50
    F = rng.normal(size=(n, 3)).astype(np.float32)
    alpha, beta = 7.0, -12.0
    a, b = np.deg2rad(alpha), np.deg2rad(beta)
    Ry = np.array([[np.cos(a), 0.0, -np.sin(a)], [0.0, 1.0, 0.0], [np.sin(a), 0.0, np.cos(a)]])
    Rz = np.array([[np.cos(b), np.sin(b), 0.0], [-np.sin(b), np.cos(b), 0.0], [0.0, 0.0, 1.0]])
    R = Rz @ Ry
    Fr = (F.astype(np.float64) @ R.T).astype(np.float32)          # rotated-frame forces
    centers = rng.normal(size=(n, 3))
    names = ["body"] * n
    O.export_net_triangle_forces_csv(tmp_path / "net.csv", tuple(Fr.T), centers, alpha_deg=alpha, beta_deg=beta,
                                     object_names=names)
    lines = (tmp_path / "net.csv").read_text().splitlines()
    assert lines[0] == "triangle_id,object_name,cx,cy,cz,Fx,Fy,Fz" and len(lines) == n + 1
    back = np.array([[float(v) for v in ln.split(",")[5:8]] for ln in lines[1:]])
    assert np.allclose(back, F, atol=2e-6 * np.abs(F).max())
    c = np.array([[float(v) for v in ln.split(",")[2:5]] for ln in lines[1:]])
    assert np.array_equal(c, np.array([[float("%.6e" % v) for v in row] for row in centers.astype(np.float32)]))
    assert lines[1].split(",")[:2] == ["1", "body"]


def test_net_force_vtp(tmp_path):
    rng = np.random.default_rng(4)
    n = 7
    net = tuple(rng.normal(size=n).astype(np.float32) for _ in range(3))
    f = O.export_net_triangle_forces_vtk(tmp_path / "net_forces_000100", net, rng.normal(size=(n, 3)),
                                         rng.normal(size=(n, 3)), rng.random(n), offset=(1.0, 2.0, 3.0))
    txt = f.read_text()
    assert f.suffix == ".vtp" and f'NumberOfPoints="{n}" NumberOfVerts="{n}"' in txt
    body = re.findall(r'Name="NetForceMagnitude" format="ascii">\n(.*?)\n        </DataArray>', txt, re.S)[0]
    mag = np.array([float(v) for v in body.split()])
    ref = np.sqrt(net[0] ** 2 + net[1] ** 2 + net[2] ** 2)
    assert np.array_equal(mag, np.array([float("%.6e" % v) for v in ref]))
