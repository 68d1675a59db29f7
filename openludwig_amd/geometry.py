"""Surface meshes: the reference's SolverMesh operations (src/geometry/geometry.jl)
and the analytic bodies used by the benchmark cases (icosphere, cube,
ellipsoid, extruded NACA 0012).

Triangles are float64 arrays of shape (n, 3, 3) (triangle, vertex, xyz).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _native as N


@dataclass
class SolverMesh:
    """SolverMesh (geometry.jl:16-24)."""
    triangles: np.ndarray   # (n,3,3) float64
    min_bounds: tuple
    max_bounds: tuple
    normals: np.ndarray     # (n,3) float64
    areas: np.ndarray       # (n,) float64
    centers: np.ndarray     # (n,3) float64

    @property
    def n(self) -> int:
        return int(self.triangles.shape[0])


def _bounds(tris: np.ndarray):
    p = tris.reshape(-1, 3)
    return tuple(float(v) for v in p.min(axis=0)), tuple(float(v) for v in p.max(axis=0))


def geometry_properties(tris: np.ndarray):
    """compute_geometry_properties (geometry.jl:75-103)."""
    tris = np.ascontiguousarray(tris, dtype=np.float64)
    n = tris.shape[0]
    normals = np.zeros((n, 3)); areas = np.zeros(n); centers = np.zeros((n, 3))
    lib = N.setup_lib()
    N.check_setup(lib.lbs_geometry_properties(N.ptr(tris, C.c_double), n, N.ptr(normals, C.c_double),
                                              N.ptr(areas, C.c_double), N.ptr(centers, C.c_double)),
                  "lbs_geometry_properties")
    return normals, areas, centers


def mesh_from_triangles(tris: np.ndarray) -> SolverMesh:
    tris = np.ascontiguousarray(tris, dtype=np.float64).reshape(-1, 3, 3)
    lo, hi = _bounds(tris)
    normals, areas, centers = geometry_properties(tris)
    return SolverMesh(tris, lo, hi, normals, areas, centers)


def read_stl(path: str | Path, scale: float = 1.0) -> np.ndarray:
    lib = N.setup_lib()
    out = N.f64p()
    n = C.c_int64()
    N.check_setup(lib.lbs_stl_read(str(path).encode(), float(scale), C.byref(out), C.byref(n)), "lbs_stl_read")
    try:
        tris = np.ctypeslib.as_array(out, shape=(n.value * 9,)).copy().reshape(-1, 3, 3)
    finally:
        lib.lbs_free(out)
    return tris


def subdivide_mesh(mesh: SolverMesh, min_size: float) -> SolverMesh:
    """subdivide_mesh (geometry.jl:174-252): children inherit the parent normal."""
    if min_size <= 0.0:
        return mesh
    lib = N.setup_lib()
    tris = np.ascontiguousarray(mesh.triangles)
    out = N.f64p()
    par = N.i64p()
    n = C.c_int64()
    N.check_setup(lib.lbs_subdivide(N.ptr(tris, C.c_double), tris.shape[0], float(min_size), C.byref(out),
                                    C.byref(par), C.byref(n)), "lbs_subdivide")
    try:
        new = np.ctypeslib.as_array(out, shape=(n.value * 9,)).copy().reshape(-1, 3, 3)
        parent = np.ctypeslib.as_array(par, shape=(n.value,)).copy()
    finally:
        lib.lbs_free(out)
        lib.lbs_free(par)
    _, areas, centers = geometry_properties(new)
    lo, hi = _bounds(new)
    return SolverMesh(new, lo, hi, mesh.normals[parent].copy(), areas, centers)


def load_mesh(path: str | Path, scale: float = 1.0, min_facet_size: float = 0.0) -> SolverMesh:
    """load_mesh (geometry.jl:458-520), without writing the subdivided STL."""
    mesh = mesh_from_triangles(read_stl(path, scale))
    if min_facet_size > 0.0:
        mesh = subdivide_mesh(mesh, float(min_facet_size))
    return mesh


def merge_meshes(meshes: list) -> SolverMesh:
    """merge_meshes (geometry.jl:533-576)."""
    if len(meshes) == 1:
        return meshes[0]
    tris = np.concatenate([m.triangles for m in meshes])
    lo, hi = _bounds(tris)
    return SolverMesh(tris, lo, hi, np.concatenate([m.normals for m in meshes]),
                      np.concatenate([m.areas for m in meshes]), np.concatenate([m.centers for m in meshes]))


def rotate_mesh(mesh: SolverMesh, alpha_deg: float, beta_deg: float) -> SolverMesh:
    """rotate_mesh (geometry.jl:267-344): R = Rz(beta) * Ry(alpha) about the
    bounding-box centre; normals rotated, areas/centres recomputed."""
    if alpha_deg == 0.0 and beta_deg == 0.0:
        return mesh
    a, b = np.deg2rad(alpha_deg), np.deg2rad(beta_deg)
    ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
    # SMatrix{3,3}(...) is filled column-major
    Ry = np.array([[ca, 0.0, -sa], [0.0, 1.0, 0.0], [sa, 0.0, ca]])
    Rz = np.array([[cb, sb, 0.0], [-sb, cb, 0.0], [0.0, 0.0, 1.0]])
    R = Rz @ Ry
    origin = (np.array(mesh.min_bounds) + np.array(mesh.max_bounds)) / 2.0
    p = mesh.triangles.reshape(-1, 3) - origin
    rot = np.stack([R[i, 0] * p[:, 0] + R[i, 1] * p[:, 1] + R[i, 2] * p[:, 2] for i in range(3)], axis=1) + origin
    tris = rot.reshape(-1, 3, 3)
    nrm = mesh.normals
    normals = np.stack([R[i, 0] * nrm[:, 0] + R[i, 1] * nrm[:, 1] + R[i, 2] * nrm[:, 2] for i in range(3)], axis=1)
    _, areas, centers = geometry_properties(tris)
    lo, hi = _bounds(tris)
    return SolverMesh(np.ascontiguousarray(tris), lo, hi, normals, areas, centers)


# ---------------------------------------------------------------------------
# Analytic bodies

def icosphere(subdiv: int, radius: float = 0.5, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Icosphere: icosahedron, each face split into 4 per level, new vertices
    projected on the sphere.  20*4^subdiv triangles, outward winding."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t],
                  [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
                  [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    tris = v[f]
    for _ in range(subdiv):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]
        ab = (a + b); ab /= np.linalg.norm(ab, axis=1, keepdims=True)
        bc = (b + c); bc /= np.linalg.norm(bc, axis=1, keepdims=True)
        ca = (c + a); ca /= np.linalg.norm(ca, axis=1, keepdims=True)
        tris = np.stack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1), np.stack([ca, bc, c], 1),
                         np.stack([ab, bc, ca], 1)], axis=1).reshape(-1, 3, 3)
    return np.ascontiguousarray(tris * radius + np.asarray(center, dtype=np.float64))


# cube1m.stl vertex pattern (12 facets, unit cube centred at the origin):
# per facet three vertices, each coordinate a sign of 0.5.
_CUBE_SIGNS = np.array([
    [[1, 1, 1], [-1, 1, 1], [-1, -1, 1]], [[1, 1, 1], [-1, -1, 1], [1, -1, 1]],
    [[1, -1, -1], [1, -1, 1], [-1, -1, 1]], [[1, -1, -1], [-1, -1, 1], [-1, -1, -1]],
    [[-1, -1, -1], [-1, -1, 1], [-1, 1, 1]], [[-1, -1, -1], [-1, 1, 1], [-1, 1, -1]],
    [[-1, 1, -1], [1, 1, -1], [1, -1, -1]], [[-1, 1, -1], [1, -1, -1], [-1, -1, -1]],
    [[1, 1, -1], [1, 1, 1], [1, -1, 1]], [[1, 1, -1], [1, -1, 1], [1, -1, -1]],
    [[-1, 1, -1], [-1, 1, 1], [1, 1, 1]], [[-1, 1, -1], [1, 1, 1], [1, 1, -1]],
], dtype=np.float64)


def cube(side: float = 1.0) -> np.ndarray:
    """The 12-facet cube of CASES/*/cube1m.stl, scaled to `side`."""
    return np.ascontiguousarray(_CUBE_SIGNS * (0.5 * side))


def ellipsoid(subdiv: int, axes=(1.0, 1.0, 1.0)) -> np.ndarray:
    """Icosphere of diameter 1 stretched by `axes` (full lengths)."""
    return np.ascontiguousarray(icosphere(subdiv, 0.5) * np.asarray(axes, dtype=np.float64))


def naca0012_wing(chord: float = 1.0, span: float = 2.0, n_chord: int = 100, n_span: int = 100) -> np.ndarray:
    """Extruded NACA 0012 (closed trailing edge), span along y, chord along x,
    flat end caps.  2*n_chord*n_span*2 side triangles + caps."""
    beta = np.linspace(0.0, np.pi, n_chord + 1)
    xc = 0.5 * (1.0 - np.cos(beta))
    yt = 0.6 * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 + 0.2843 * xc ** 3 - 0.1036 * xc ** 4)
    # closed loop: upper surface TE->LE, lower LE->TE (TE point shared)
    xu, zu = xc[::-1], yt[::-1]
    xl, zl = xc[1:], -yt[1:]
    px = np.concatenate([xu, xl])[:-1] * chord
    pz = np.concatenate([zu, zl])[:-1] * chord
    m = px.size
    ys = np.linspace(0.0, span, n_span + 1)
    tris = []
    for j in range(n_span):
        y0, y1 = ys[j], ys[j + 1]
        for i in range(m):
            i2 = (i + 1) % m
            a = (px[i], y0, pz[i]); b = (px[i2], y0, pz[i2])
            c = (px[i2], y1, pz[i2]); d = (px[i], y1, pz[i])
            tris.append((a, c, b))
            tris.append((a, d, c))
    # end caps: fan from the mid-chord centre point
    for y, flip in ((0.0, False), (span, True)):
        cen = (0.5 * chord, y, 0.0)
        for i in range(m):
            i2 = (i + 1) % m
            p, q = (px[i], y, pz[i]), (px[i2], y, pz[i2])
            tris.append((cen, q, p) if not flip else (cen, p, q))
    return np.ascontiguousarray(np.array(tris, dtype=np.float64))
