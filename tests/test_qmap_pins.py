"""Pins of the Bouzidi Q-map restatement (lbs_bouzidi_qmap; the GPU Q-map
equals it bit for bit) against geometry it must reproduce
(compute_bouzidi_qmap_sparse, bouzidi_setup.jl:64-167; compute_q_for_cell,
bouzidi_math.jl:53-102):

  * an icosphere (subdivision 4) of radius R: every stored q, scaled to a
    distance t = q dx |c|, is the exact ray-sphere distance from the cell
    centre along c_k to within the icosphere's chord error over the
    incidence cosine plus Float16 rounding; every link whose exact distance
    lies well inside (0, dx |c|] is found; the stored triangle is hit by
    that ray at that distance;
  * ties and order: a triangle listed twice is reported with its first
    (lowest) index -- the first minimum wins (strict <) -- while a nearer
    triangle wins whatever its index; hits beyond one link length (q > 1)
    are dropped.
"""
import ctypes as C

import numpy as np

from openludwig_amd import _native as N
from openludwig_amd.geometry import icosphere

K = np.arange(27)
CVEC = np.stack([K % 3 - 1, (K // 3) % 3 - 1, K // 9 - 1], 1).astype(np.float64)


def _box(nbx, nby, nbz):
    coords = np.array([[x, y, z] for x in range(1, nbx + 1) for y in range(1, nby + 1) for z in range(1, nbz + 1)],
                      dtype=np.int32)
    order = np.lexsort((coords[:, 2], coords[:, 1], coords[:, 0]))
    coords = np.ascontiguousarray(coords[order])
    grid = np.zeros((nbx, nby, nbz), dtype=np.int32, order="F")
    grid[coords[:, 0] - 1, coords[:, 1] - 1, coords[:, 2] - 1] = np.arange(1, len(coords) + 1)
    return coords, np.ascontiguousarray(grid.ravel(order="F")), (nbx, nby, nbz)


def qmap(coords, grid, dims, tris, dx, off):
    lib = N.setup_lib()
    tris = np.ascontiguousarray(tris, dtype=np.float64)
    off = np.ascontiguousarray(off, dtype=np.float64)
    cells, q, tri = N.i32p(), N.u16p(), N.i32p()
    n = C.c_int64()
    rc = lib.lbs_bouzidi_qmap(len(coords), N.ptr(coords, C.c_int32), N.ptr(grid, C.c_int32), *dims,
                              N.ptr(tris, C.c_double), tris.shape[0], dx, N.ptr(off, C.c_double), C.byref(cells),
                              C.byref(q), C.byref(tri), C.byref(n))
    assert rc == 0
    m = n.value
    try:
        cl = np.ctypeslib.as_array(cells, shape=(max(m, 1) * 4,))[:m * 4].copy().reshape(m, 4)
        qb = np.ctypeslib.as_array(q, shape=(max(m, 1) * 27,))[:m * 27].copy().reshape(m, 27)
        tb = np.ctypeslib.as_array(tri, shape=(max(m, 1) * 27,))[:m * 27].copy().reshape(m, 27)
    finally:
        lib.lbs_free(cells); lib.lbs_free(q); lib.lbs_free(tri)
    return cl, qb.view(np.float16).astype(np.float64), tb


def _ray_sphere(o, d, c, r):
    """Smallest t > 1e-9 with |o + t d - c| = r (d unit), or inf."""
    oc = o - c
    b = oc @ d
    disc = b * b - (oc @ oc - r * r)
    if disc < 0:
        return np.inf
    s = np.sqrt(disc)
    for t in (-b - s, -b + s):
        if t > 1e-9:
            return t
    return np.inf


def _ray_tri(o, d, tri):
    e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
    h = np.cross(d, e2)
    a = e1 @ h
    if abs(a) < 1e-12:
        return np.inf
    s = o - tri[0]
    u = (s @ h) / a
    qv = np.cross(s, e1)
    v = (d @ qv) / a
    if u < -1e-9 or v < -1e-9 or u + v > 1 + 1e-9:
        return np.inf
    return (e2 @ qv) / a


def test_icosphere_q_against_exact_sphere():
    R, dx = 0.5, 1.0 / 16
    center = np.array([1.0, 1.0, 1.0])
    tris = icosphere(4, R)                         # centred at the origin; offset places it
    coords, grid, dims = _box(4, 4, 4)             # 32^3 cells of dx: a 2 m box
    cl, q, tb = qmap(coords, grid, dims, tris, dx, center)
    assert len(cl) > 1000
    # chord error: the largest distance of a face from the sphere (sagitta of its circumcircle)
    edge = max(np.linalg.norm(tris[:, 1] - tris[:, 0], axis=1).max(),
               np.linalg.norm(tris[:, 2] - tris[:, 1], axis=1).max())
    sag = R - np.sqrt(R * R - (edge / np.sqrt(3.0)) ** 2)
    checked = 0
    listed = {tuple(c) for c in cl.tolist()}
    for i, (b, lx, ly, lz) in enumerate(cl):
        bx, by, bz = coords[b - 1]
        o = (np.array([(bx - 1) * 8 + lx, (by - 1) * 8 + ly, (bz - 1) * 8 + lz]) - 0.5) * dx
        for k in range(27):
            if k == 13:
                assert q[i, k] == 0
                continue
            cn = np.linalg.norm(CVEC[k])
            d = CVEC[k] / cn
            te = _ray_sphere(o, d, center, R)
            if q[i, k] > 0:
                t = q[i, k] * dx * cn
                hit = o + te * d
                cos_inc = abs((hit - center) @ d) / R if np.isfinite(te) else 0.0
                if cos_inc > 0.5:
                    checked += 1
                    assert abs(t - te) <= sag / cos_inc + 1e-3 * t + 1e-9, (i, k, t, te)
                # the stored triangle is hit by this ray at the stored distance
                tri = tris[tb[i, k] - 1] + center
                assert abs(_ray_tri(o, d, tri) - t) <= 1e-3 * t + 1e-9
            elif np.isfinite(te):
                hit = o + te * d
                cos_inc = abs((hit - center) @ d) / R
                if cos_inc > 0.5 and 0.05 * dx * cn < te < 0.95 * dx * cn - sag / cos_inc:
                    raise AssertionError(f"link missed: cell {i} k {k} exact t {te}")
    # and no cell with such a link is missing from the list
    for b, (bx, by, bz) in enumerate(coords, start=1):
        for lz in range(1, 9):
            for ly in range(1, 9):
                for lx in range(1, 9):
                    if (b, lx, ly, lz) in listed:
                        continue
                    o = (np.array([(bx - 1) * 8 + lx, (by - 1) * 8 + ly, (bz - 1) * 8 + lz]) - 0.5) * dx
                    r0 = np.linalg.norm(o - center)
                    if abs(r0 - R) > 2 * dx:
                        continue
                    for k in range(27):
                        if k == 13:
                            continue
                        cn = np.linalg.norm(CVEC[k])
                        d = CVEC[k] / cn
                        te = _ray_sphere(o, d, center, R)
                        if np.isfinite(te):
                            cos_inc = abs((o + te * d - center) @ d) / R
                            if cos_inc > 0.5 and 0.05 * dx * cn < te < 0.95 * dx * cn - sag / cos_inc:
                                raise AssertionError(f"cell {(b, lx, ly, lz)} missing, k {k}")
    assert checked > 2000


def test_first_minimum_in_index_order_and_q_limit():
    """A 1-block box of dx = 0.1 and a plane x = 0.43 made of two triangles.
    Listed twice (triangles 1-2, then the same two as 3-4), a ray from the
    centre of cell x = 4 (x = 0.35) along +x ties between a triangle and its
    copy at t = 0.08 (q = 0.8): the first of the tie (1 or 2) is kept.  With
    a nearer plane x = 0.41 listed after it (triangles 3-4), the nearer hit
    (q = 0.6) wins whatever its index."""
    dx = 0.1
    coords, grid, dims = _box(1, 1, 1)
    big = 5.0

    def plane(x0):
        return [[[x0, -big, -big], [x0, big, -big], [x0, big, big]], [[x0, -big, -big], [x0, big, big],
                                                                        [x0, -big, big]]]

    a = plane(0.43)
    tris_ab = np.array(a + a)                       # indices 1-2 and 3-4: the same two triangles twice
    cl, q, tb = qmap(coords, grid, dims, tris_ab, dx, np.zeros(3))
    x4 = cl[:, 1] == 4
    kx = 14                                         # c = (+1, 0, 0)
    assert x4.any()
    assert np.all(q[x4, kx] == np.float64(np.float16(0.8)))
    assert np.all((tb[x4, kx] == 1) | (tb[x4, kx] == 2))   # one of the first pair, never the later copy
    tris = np.array(a + plane(0.41))                # the nearer plane listed after the farther one
    cl, q, tb = qmap(coords, grid, dims, tris, dx, np.zeros(3))
    x4 = cl[:, 1] == 4
    assert np.all(q[x4, kx] == np.float64(np.float16(0.6)))
    assert np.all((tb[x4, kx] == 3) | (tb[x4, kx] == 4))
    # cells farther than one link from every plane keep q = 0 in that direction
    x2 = cl[:, 1] == 2 if (cl[:, 1] == 2).any() else None
    assert x2 is None or np.all(q[x2, kx] == 0)
