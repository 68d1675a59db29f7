"""rotate_mesh against the reference's own rotation literals (CPU).

The reference builds Ry(alpha) and Rz(beta) as SMatrix{3,3,Float64}(...)
literals and applies R = Rz * Ry about the centre of the bounding box
(src/geometry/geometry.jl:267-344).  Julia fills an SMatrix literal column
by column.  The test reads those literals from the reference source when the
checkout is present, evaluates them for sample angles, and checks that
rotate_mesh moves vertices, normals and bounds exactly as that R does.  A
transposed matrix, a flipped sign or the product order Ry * Rz fails it.
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

from openludwig_amd.geometry import SolverMesh, rotate_mesh

SRC = Path("/root/reference/src/geometry/geometry.jl")
needs_ref = pytest.mark.skipif(not SRC.is_file(), reason="reference checkout not present")


def _literal(text: str, name: str) -> list[str]:
    m = re.search(name + r"\s*=\s*SMatrix\{3,\s*3,\s*Float64\}\(([^)]*)\)", text)
    assert m, f"{name} literal not found"
    items = [t.strip() for t in m.group(1).replace("\n", " ").split(",")]
    assert len(items) == 9, items
    return items


def _evaluate(items: list[str], env: dict) -> np.ndarray:
    vals = []
    for t in items:
        neg = t.startswith("-")
        t = t.lstrip("-").strip()
        v = env[t] if t in env else float(t)
        vals.append(-v if neg else v)
    return np.array(vals, dtype=np.float64).reshape(3, 3, order="F")  # column-major fill


@needs_ref
@pytest.mark.parametrize("alpha,beta", [(7.0, 0.0), (0.0, -12.0), (7.0, -12.0), (90.0, 30.0)])
def test_rotate_mesh_matches_reference_literals(alpha, beta):
    text = SRC.read_text()
    body = text[text.index("function rotate_mesh"):]
    body = body[:body.index("\nend\n")]
    assert re.search(r"\bR\s*=\s*Rz\s*\*\s*Ry\b", body), "product order changed in the reference"
    a, b = np.deg2rad(alpha), np.deg2rad(beta)
    env = {"ca": np.cos(a), "sa": np.sin(a), "cb": np.cos(b), "sb": np.sin(b)}
    R = _evaluate(_literal(body, "Rz"), env) @ _evaluate(_literal(body, "Ry"), env)

    rng = np.random.default_rng(11)
    tris = rng.normal(size=(40, 3, 3)) + np.array([3.0, -1.0, 0.5])
    p = tris.reshape(-1, 3)
    lo, hi = p.min(axis=0), p.max(axis=0)
    normals = rng.normal(size=(40, 3))
    normals /= np.linalg.norm(normals, axis=1, keepdims=True)
    mesh = SolverMesh(tris, tuple(lo), tuple(hi), normals, np.zeros(40), tris.mean(axis=1))

    out = rotate_mesh(mesh, alpha, beta)
    origin = (lo + hi) / 2.0
    want = (p - origin) @ R.T + origin
    assert np.allclose(out.triangles.reshape(-1, 3), want, rtol=0.0, atol=1e-12)
    assert np.allclose(out.normals, normals @ R.T, rtol=0.0, atol=1e-12)
    assert np.allclose(out.min_bounds, want.min(axis=0), rtol=0.0, atol=1e-12)
    assert np.allclose(out.max_bounds, want.max(axis=0), rtol=0.0, atol=1e-12)
    assert np.allclose(out.centers, want.reshape(-1, 3, 3).mean(axis=1), rtol=0.0, atol=1e-12)


def test_rotate_mesh_zero_angles_is_identity():
    tris = np.arange(27, dtype=np.float64).reshape(3, 3, 3)
    mesh = SolverMesh(tris, (0.0, 1.0, 2.0), (24.0, 25.0, 26.0), np.eye(3), np.ones(3), tris.mean(axis=1))
    assert rotate_mesh(mesh, 0.0, 0.0) is mesh
