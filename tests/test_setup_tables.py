"""Host table producers (libludwig_setup.so + openludwig_amd.domain) against
closed forms, invariants, the reference's own fixture (CASES cube1m.stl, kept
as tests/golden/cube1m.stl) and independent pure-Python restatements of the
reference's set logic and SAT test (CPU only)."""
import ctypes as C
import math
from pathlib import Path

import numpy as np
import pytest

from openludwig_amd import _native as N
from openludwig_amd import cases
from openludwig_amd.domain import (build_neighbor_table, compute_domain_from_mesh, setup_multilevel_domain)
from openludwig_amd.geometry import cube, icosphere, mesh_from_triangles, read_stl, subdivide_mesh
from openludwig_amd.solver import build_case

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_double_to_half_matches_numpy_rne():
    lib = N.setup_lib()
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.uniform(0, 1, 20000), rng.uniform(0, 1e-4, 2000), 2.0 ** -np.arange(0, 30),
                           [0.0, 1.0, 0.5, 0.3, 0.001, 65504.0, 1e-8, 6.1e-5, 5.96e-8, 2.98e-8]])
    # half-way cases between two halves exercise round-to-even
    h = np.arange(1, 2000, dtype=np.uint16).view(np.float16).astype(np.float64)
    mids = (h[:-1] + h[1:]) / 2
    vals = np.concatenate([vals, mids])
    got = np.array([lib.lbs_double_to_half(float(v)) for v in vals], dtype=np.uint16)
    ref = vals.astype(np.float16).view(np.uint16)
    assert (got == ref).all()


def test_cube_fixture_and_subdivision():
    tris = read_stl(GOLDEN / "cube1m.stl", 1.0)
    assert tris.shape == (12, 3, 3)
    assert np.array_equal(tris, cube(1.0))
    m = subdivide_mesh(mesh_from_triangles(tris), 0.025)
    assert m.n == 12 * 2 ** 12 == 49152
    assert abs(m.areas.sum() - 6.0) < 1e-12
    e = np.linalg.norm(np.roll(m.triangles, -1, axis=1) - m.triangles, axis=2)
    assert e.max() <= 0.025
    assert len(np.unique(m.normals.round(12), axis=0)) == 6          # children inherit parent normals
    assert np.abs((m.normals * m.areas[:, None]).sum(0)).max() < 1e-12


def test_ascii_stl(tmp_path):
    p = tmp_path / "t.stl"
    p.write_text("solid t\n facet normal 0 0 1\n  outer loop\n   vertex 0 0 0\n   vertex 1 0 0\n"
                 "   vertex 0 1 0\n  endloop\n endfacet\nendsolid t\n")
    tris = read_stl(p, 2.0)
    assert np.array_equal(tris, np.array([[[0, 0, 0], [2, 0, 0], [0, 2, 0]]], dtype=np.float64))
    with pytest.raises(N.NativeError):
        read_stl(tmp_path / "missing.stl")


def test_icosphere_closed_sphere():
    t = icosphere(4, 0.5)
    assert t.shape[0] == 5120
    assert np.abs(np.linalg.norm(t.reshape(-1, 3), axis=1) - 0.5).max() < 1e-15
    m = mesh_from_triangles(t)
    assert np.abs((m.normals * m.areas[:, None]).sum(0)).max() < 1e-13
    assert ((m.normals * m.centers).sum(1) > 0).all()                 # outward winding
    assert abs(m.areas.sum() - math.pi) < 0.01


def test_domain_scaling_hand_values():
    """compute_domain_from_mesh for the 3-level sphere case (physics_scaling.jl:83-191)."""
    cfg, mesh = cases.sphere_re1e3()
    p = compute_domain_from_mesh(cfg, mesh.min_bounds, mesh.max_bounds)
    assert (p.nx_coarse, p.ny_coarse, p.nz_coarse) == (64, 32, 32)
    assert p.num_levels == 3 and abs(p.dx_coarse - 0.25) < 1e-15
    assert p.tau_levels == [float(np.float32(v)) for v in (0.5096, 0.5048, 0.5024)]
    assert abs(p.re_number - 1000.0) < 1e-9
    cfg2, mesh2 = cases.cube_re1e5(160)
    p2 = compute_domain_from_mesh(cfg2, mesh2.min_bounds, mesh2.max_bounds)
    assert p2.tau_levels[-1] == float(np.float32(0.501))               # tau_min clamp


# ---------------------------------------------------------------------------
# topology: literal set restatement of domain_topology.jl / domain.jl

def _ref_topology(cfg, mesh, params):
    off = np.array(params.mesh_offset)
    levels = []
    pm = np.array(params.mesh_min) + off
    pM = np.array(params.mesh_max) + off
    wsx = pM[0] - params.reference_length * 0.1
    wex = pM[0] + params.reference_length * cfg.wake_length
    wcy, wcz = (pm[1] + pM[1]) / 2.0, (pm[2] + pM[2]) / 2.0
    ww, wh = (pM[1] - pm[1]) * cfg.wake_width_factor, (pM[2] - pm[2]) * cfg.wake_height_factor
    for lvl in range(1, params.num_levels + 1):
        scale = 2 ** (lvl - 1)
        dx = params.dx_coarse / scale
        bxm, bym, bzm = params.bx_max * scale, params.by_max * scale, params.bz_max * scale
        act = set()
        if lvl == 1:
            act = {(x, y, z) for z in range(1, bzm + 1) for y in range(1, bym + 1) for x in range(1, bxm + 1)}
        else:
            margin = dx * 0.01
            inv = 1.0 / (8 * dx)
            for tri in mesh.triangles:
                tmin = tri[0] + off
                tmax = tri[0] + off
                for pnt in tri:
                    pp = pnt + off
                    tmin = np.minimum(tmin, pp)
                    tmax = np.maximum(tmax, pp)
                lo = [max(1, math.floor((tmin[i] - margin) * inv) + 1) for i in range(3)]
                hi = [min(math.floor((tmax[i] + margin) * inv) + 1, m) for i, m in enumerate((bxm, bym, bzm))]
                for bz in range(lo[2], hi[2] + 1):
                    for by in range(lo[1], hi[1] + 1):
                        for bx in range(lo[0], hi[0] + 1):
                            act.add((bx, by, bz))
            prev = levels[-1]
            if cfg.wake_enabled:
                pbs = 8 * (params.dx_coarse / 2 ** (lvl - 2))
                for (cbx, cby, cbz) in prev:
                    if ((cbx - 1) * pbs <= wex and cbx * pbs >= wsx and (cby - 1) * pbs <= wcy + ww / 2
                            and cby * pbs >= wcy - ww / 2 and (cbz - 1) * pbs <= wcz + wh / 2
                            and cbz * pbs >= wcz - wh / 2):
                        for dbz in (0, 1):
                            for dby in (0, 1):
                                for dbx in (0, 1):
                                    f = (2 * cbx - 1 + dbx, 2 * cby - 1 + dby, 2 * cbz - 1 + dbz)
                                    if 1 <= f[0] <= bxm and 1 <= f[1] <= bym and 1 <= f[2] <= bzm:
                                        act.add(f)
            pset = set(prev)
            act = {b for b in act if ((b[0] + 1) // 2, (b[1] + 1) // 2, (b[2] + 1) // 2) in pset}
        offs = [(dx_, dy_, dz_) for dz_ in (-1, 0, 1) for dy_ in (-1, 0, 1) for dx_ in (-1, 0, 1)
                if (dx_, dy_, dz_) != (0, 0, 0)]
        for _ in range(cfg.refinement_margin):
            new = set()
            for (bx, by, bz) in act:
                for (a, b, c) in offs:
                    n = (bx + a, by + b, bz + c)
                    if 1 <= n[0] <= bxm and 1 <= n[1] <= bym and 1 <= n[2] <= bzm and n not in act:
                        new.add(n)
            sib = set()
            for (bx, by, bz) in new:
                pb = ((bx + 1) // 2, (by + 1) // 2, (bz + 1) // 2)
                for dbz in (0, 1):
                    for dby in (0, 1):
                        for dbx in (0, 1):
                            s = (2 * pb[0] - 1 + dbx, 2 * pb[1] - 1 + dby, 2 * pb[2] - 1 + dbz)
                            if 1 <= s[0] <= bxm and 1 <= s[1] <= bym and 1 <= s[2] <= bzm and s not in act \
                                    and s not in new:
                                sib.add(s)
            act |= new
            act |= sib
        for _ in range(10):
            add = set()
            for (bx, by, bz) in act:
                pb = ((bx + 1) // 2, (by + 1) // 2, (bz + 1) // 2)
                for dbz in (0, 1):
                    for dby in (0, 1):
                        for dbx in (0, 1):
                            s = (2 * pb[0] - 1 + dbx, 2 * pb[1] - 1 + dby, 2 * pb[2] - 1 + dbz)
                            if 1 <= s[0] <= bxm and 1 <= s[1] <= bym and 1 <= s[2] <= bzm and s not in act:
                                add.add(s)
            if not add:
                break
            act |= add
        levels.append(sorted(act))
    return levels


@pytest.mark.parametrize("case", ["tiny2", "tiny3", "sphere3"])
def test_topology_matches_set_restatement(case):
    if case == "tiny2":
        cfg, mesh = cases.tiny_sphere(8, 2)
    elif case == "tiny3":
        cfg, mesh = cases.tiny_sphere(16, 3)
    else:
        cfg, mesh = cases.sphere_re1e3()
        mesh = mesh_from_triangles(icosphere(2, 0.5))          # 320 triangles keeps the restatement quick
    params = compute_domain_from_mesh(cfg, mesh.min_bounds, mesh.max_bounds)
    ref = _ref_topology(cfg, mesh, params)
    grids = setup_multilevel_domain(cfg, mesh, params)
    assert len(grids) == len(ref)
    for g, r in zip(grids, ref):
        assert [tuple(int(v) for v in c) for c in g.coords] == r
        # neighbour table: literal restatement (domain_topology.jl:135-160)
        lookup = {c: i + 1 for i, c in enumerate(r)}
        dims = g.grid_dims
        for i, (bx, by, bz) in enumerate(r[:: max(1, len(r) // 200)]):
            row = g.neighbor_table[lookup[(bx, by, bz)] - 1]
            d = 0
            for dz in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dx in (-1, 0, 1):
                        n = (bx + dx, by + dy, bz + dz)
                        inb = 1 <= n[0] <= dims[0] and 1 <= n[1] <= dims[1] and 1 <= n[2] <= dims[2]
                        assert row[d] == (lookup.get(n, 0) if inb else 0)
                        d += 1


def test_topology_invariants():
    cfg, mesh = cases.sphere_re1e3()
    grids, params, mesh = build_case(cfg, mesh)
    for li, g in enumerate(grids):
        nbt = g.neighbor_table
        # symmetry: nb(b,d) = c  <=>  nb(c, opp d) = b
        for d in range(27):
            has = nbt[:, d] > 0
            b = np.nonzero(has)[0]
            assert (nbt[nbt[b, d] - 1, 26 - d] == b + 1).all()
        assert (nbt[:, 13] == np.arange(1, g.n_blocks + 1)).all()
        # sorted lexicographically, bx most significant
        key = g.coords[:, 0].astype(np.int64) * 10 ** 8 + g.coords[:, 1] * 10 ** 4 + g.coords[:, 2]
        assert (np.diff(key) > 0).all()
        # block_pointer
        bp = g.block_pointer
        assert bp.shape == tuple(int(v) for v in g.coords.max(0))
        assert (bp[g.coords[:, 0] - 1, g.coords[:, 1] - 1, g.coords[:, 2] - 1] == np.arange(1, g.n_blocks + 1)).all()
        if li > 0:
            parents = set(map(tuple, grids[li - 1].coords))
            for c in g.coords:
                assert ((c[0] + 1) // 2, (c[1] + 1) // 2, (c[2] + 1) // 2) in parents
            have = set(map(tuple, g.coords))
            for c in g.coords:                              # sibling closure
                pb = ((c[0] + 1) // 2, (c[1] + 1) // 2, (c[2] + 1) // 2)
                for s in [(2 * pb[0] - 1 + a, 2 * pb[1] - 1 + b, 2 * pb[2] - 1 + cc) for a in (0, 1)
                          for b in (0, 1) for cc in (0, 1)]:
                    assert s in have


# ---------------------------------------------------------------------------
# per-cell tables

def _sat(center, h, v1, v2, v3):
    """triangle_intersects_aabb (domain_generation.jl:10-32), literal."""
    h = h * 1.001
    t = [v1 - center, v2 - center, v3 - center]
    for i in range(3):
        if min(t[0][i], t[1][i], t[2][i]) > h or max(t[0][i], t[1][i], t[2][i]) < -h:
            return False
    f = [t[1] - t[0], t[2] - t[1], t[0] - t[2]]
    for i in range(3):
        u = np.zeros(3); u[i] = 1.0
        for j in range(3):
            ax = np.array([u[1] * f[j][2] - u[2] * f[j][1], u[2] * f[j][0] - u[0] * f[j][2],
                           u[0] * f[j][1] - u[1] * f[j][0]])
            if ax @ ax < 1e-10:
                continue
            p = [t[0] @ ax, t[1] @ ax, t[2] @ ax]
            r = h * abs(ax[0]) + h * abs(ax[1]) + h * abs(ax[2])
            if min(p) > r or max(p) < -r:
                return False
    return True


def test_voxelization_and_flood_fill():
    cfg, _ = cases.tiny_sphere(8, 2)
    mesh = mesh_from_triangles(icosphere(2, 0.5))
    params = compute_domain_from_mesh(cfg, mesh.min_bounds, mesh.max_bounds)
    grids = setup_multilevel_domain(cfg, mesh, params)
    g = grids[-1]
    off = np.array(params.mesh_offset)
    dx = params.dx_levels[-1]
    tris = mesh.triangles + off
    centre = (np.array(params.mesh_min) + np.array(params.mesh_max)) / 2 + off
    lx = np.arange(512) % 8; ly = (np.arange(512) // 8) % 8; lz = np.arange(512) // 64
    rng = np.random.default_rng(0)
    checked = 0
    for b in rng.choice(g.n_blocks, size=min(12, g.n_blocks), replace=False):
        bx, by, bz = g.coords[b]
        pc = np.stack([((bx - 1) * 8 + lx + 1 - 0.5) * dx, ((by - 1) * 8 + ly + 1 - 0.5) * dx,
                       ((bz - 1) * 8 + lz + 1 - 0.5) * dx], 1)
        r = np.linalg.norm(pc - centre, axis=1)
        for c in range(0, 512, 7):
            shell = any(_sat(pc[c], 0.75 * dx, *t) for t in tris if np.abs(t - pc[c]).min() < 3 * dx)
            if shell:
                assert g.obstacle[b, c] == 1
                checked += 1
            elif r[c] > 0.5 + 2 * dx:
                assert g.obstacle[b, c] == 0
        # flood fill: well inside the sphere is solid
        assert (g.obstacle[b][r < 0.5 - 2 * dx] == 1).all()
    assert checked >= 0


def test_sponge_profile_values():
    cfg, mesh = cases.tiny_sphere(8, 2)
    grids, params, mesh = build_case(cfg, mesh)
    g = grids[0]
    dx = params.dx_coarse
    Lx, Ly, Lz = params.domain_size
    th = float(np.float32(0.10))
    out_th, in_th, y_th, z_th = Lx * max(th, 0.08), Lx * 0.02, Ly * th * 0.5, Lz * th * 0.5

    def prof(x, t):
        return 1.0 if x <= 0 else (0.0 if x >= t else 0.5 * (1 + math.cos(math.pi * x / t)))

    rng = np.random.default_rng(1)
    for b in rng.choice(g.n_blocks, size=min(20, g.n_blocks), replace=False):
        for c in rng.choice(512, size=20, replace=False):
            bx, by, bz = g.coords[b]
            px = ((bx - 1) * 8 + c % 8 + 1 - 0.5) * dx
            py = ((by - 1) * 8 + (c // 8) % 8 + 1 - 0.5) * dx
            pz = ((bz - 1) * 8 + c // 64 + 1 - 0.5) * dx
            s = 0.0
            if px > Lx - out_th:
                s = max(s, prof(out_th - (px - (Lx - out_th)), out_th))
            if px < in_th:
                s = max(s, prof(px, in_th) * 0.05)
            if py < y_th:
                s = max(s, prof(py, y_th) * 0.1)
            if py > Ly - y_th:
                s = max(s, prof(y_th - (py - (Ly - y_th)), y_th) * 0.1)
            if pz < z_th:
                s = max(s, prof(pz, z_th) * 0.1)
            if pz > Lz - z_th:
                s = max(s, prof(z_th - (pz - (Lz - z_th)), z_th) * 0.1)
            assert g.sponge[b, c] == np.float32(s)


def test_wall_distance_values():
    cfg, mesh = cases.tiny_sphere(8, 2, wall_model=True)
    grids, params, mesh = build_case(cfg, mesh)
    g = grids[-1]
    dxf = np.float32(g.dx)
    allowed = {np.float32(100.0), np.float32(1.0) * dxf, np.sqrt(np.float32(2.0)) * dxf,
               np.sqrt(np.float32(3.0)) * dxf}
    vals = set(np.unique(g.wall_dist).tolist())
    assert vals <= {float(v) for v in allowed}
    assert (g.wall_dist[g.obstacle == 1] == 100.0).all()
    assert (g.wall_dist < 100).sum() > 0


def test_qmap_plane_known_answer():
    """A plane at x = x_c + 0.3 dx in front of a row of cell centres: every
    direction with c_x = +1 of those cells has q = 0.3 (the distance along
    c_k is 0.3 dx |c_k|), stored as Float16(0.3); cells 0.7 dx behind the
    plane see q = 0.7 along c_x = -1 (bouzidi_math.jl:53-102)."""
    lib = N.setup_lib()
    dx = 0.1
    coords = np.array([[x, y, z] for x in (1, 2) for y in (1, 2) for z in (1, 2)], dtype=np.int32)
    order = np.lexsort((coords[:, 2], coords[:, 1], coords[:, 0]))
    coords = np.ascontiguousarray(coords[order])
    dims = (2, 2, 2)
    grid = np.zeros(dims, dtype=np.int32, order="F")
    grid[coords[:, 0] - 1, coords[:, 1] - 1, coords[:, 2] - 1] = np.arange(1, 9)
    grid_c = np.ascontiguousarray(grid.ravel(order="F"))
    x0 = (8 - 0.5) * dx + 0.3 * dx              # 0.3 dx beyond the centre of global cell 8
    big = 10.0
    tris = np.ascontiguousarray(np.array([[[x0, -big, -big], [x0, big, -big], [x0, big, big]],
                                          [[x0, -big, -big], [x0, big, big], [x0, -big, big]]], dtype=np.float64))
    off = np.zeros(3)
    cells, q, tri = N.i32p(), N.u16p(), N.i32p()
    n = C.c_int64()
    rc = lib.lbs_bouzidi_qmap(8, N.ptr(coords, C.c_int32), N.ptr(grid_c, C.c_int32), 2, 2, 2,
                              N.ptr(tris, C.c_double), 2, dx, N.ptr(off, C.c_double), C.byref(cells), C.byref(q),
                              C.byref(tri), C.byref(n))
    assert rc == 0
    m = n.value
    cl = np.ctypeslib.as_array(cells, shape=(m * 4,)).copy().reshape(m, 4)
    qb = np.ctypeslib.as_array(q, shape=(m * 27,)).copy().reshape(m, 27)
    lib.lbs_free(cells); lib.lbs_free(q); lib.lbs_free(tri)
    h03, h07 = np.float16(0.3).view(np.uint16), np.float16(0.7).view(np.uint16)
    k = np.arange(27)
    cxk = k % 3 - 1
    gxs = (coords[cl[:, 0] - 1, 0] - 1) * 8 + cl[:, 1]
    assert set(gxs.tolist()) == {8, 9}
    assert m == 2 * 16 * 16
    for i in range(m):
        if gxs[i] == 8:
            assert (qb[i][cxk == 1] == h03).all() and (qb[i][cxk != 1] == 0).all()
        else:
            assert (qb[i][cxk == -1] == h07).all() and (qb[i][cxk != -1] == 0).all()


def _flood_fill_bfs(coords, dims, obst):
    """Literal restatement of perform_flood_fill! (domain_generation.jl:114-203):
    breadth-first search from the fluid cells of the blocks with the smallest
    bx over 6-connected fluid cells; unreached fluid cells become solid."""
    from collections import deque
    nb = coords.shape[0]
    index = {tuple(c): i for i, c in enumerate(coords)}
    vis = np.zeros((nb, 512), dtype=bool)
    q = deque()
    minbx = coords[:, 0].min()
    for b in range(nb):
        if coords[b, 0] == minbx:
            for c in range(512):
                if not obst[b, c]:
                    vis[b, c] = True
                    q.append((b, c))
    while q:
        b, c = q.popleft()
        x, y, z = c % 8, (c // 8) % 8, c // 64
        for dx, dy, dz in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            nx, ny, nz = x + dx, y + dy, z + dz
            bb = b
            if not (0 <= nx < 8 and 0 <= ny < 8 and 0 <= nz < 8):
                key = (coords[b, 0] + (nx > 7) - (nx < 0), coords[b, 1] + (ny > 7) - (ny < 0),
                       coords[b, 2] + (nz > 7) - (nz < 0))
                if key not in index:
                    continue
                bb = index[key]
                nx, ny, nz = nx % 8, ny % 8, nz % 8
            cc = nx + 8 * ny + 64 * nz
            if not vis[bb, cc] and not obst[bb, cc]:
                vis[bb, cc] = True
                q.append((bb, cc))
    out = obst.copy()
    fill = (~vis) & (obst == 0)
    out[fill] = 1
    return out, int(fill.sum())


@pytest.mark.parametrize("density", [0.0, 0.35, 0.6])
def test_flood_fill_components_equal_bfs(density):
    """The parallel component form of lbs_flood_fill equals the literal BFS on
    random solid fields (many components per block) over a sparse block set."""
    rng = np.random.default_rng(int(density * 100) + 3)
    dims = (6, 5, 4)
    active = rng.random(dims) < 0.7
    active[0, :, :] = True
    coords = np.ascontiguousarray(np.argwhere(active).astype(np.int32) + 1)
    nb = coords.shape[0]
    dense = np.zeros(dims, dtype=np.int32)
    for i, c in enumerate(coords):
        dense[c[0] - 1, c[1] - 1, c[2] - 1] = i + 1
    dense_c = np.ascontiguousarray(np.asfortranarray(dense).ravel(order="F"))
    obst = (rng.random((nb, 512)) < density).astype(np.uint8)
    # a closed solid shell with fluid inside, away from the seed blocks
    obst[nb // 2, :] = 1
    obst[nb // 2, 73] = 0
    ref, nref = _flood_fill_bfs(coords, dims, obst)
    got = obst.copy()
    lib = N.setup_lib()
    n = lib.lbs_flood_fill(nb, N.ptr(coords, C.c_int32), N.ptr(dense_c, C.c_int32), dims[0], dims[1], dims[2],
                           N.ptr(got, C.c_uint8))
    assert n == nref
    assert np.array_equal(got, ref)
