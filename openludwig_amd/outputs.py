"""Result files of a case run, in the reference's formats:

convergence.csv   main.jl:56-57 (header), 245-257 (rows)
forces.csv        src/forces/io.jl:5-22
surface_NNNNNN.lbmp   export_surface_pressure_bin (src/forces/io.jl:230-335)
surface_NNNNNN.vtu    save_surface_vtk (src/forces/io.jl:75-137)
flow_NNNNNN.vtu       export_merged_mesh_sync (src/io/io_vtk.jl:6-129)
net_forces_NNNNNN.csv / .vtp   export_net_triangle_forces_csv / _vtk
                      (src/forces/io.jl:139-228, 337-444)

The CSV rows and the net-force VTP (an ASCII file) use the reference's
printf formats, so files written by either program diff line by line.  The
.lbmp writer produces the reference's byte layout (little-endian header,
name table, contiguous Float32/Int32 arrays) with the same Float32 rounding
of every value.  The VTU files are VTK XML
UnstructuredGrid files holding the reference's arrays under its names; they
are written with inline base64 binary data (WriteVTK's compressed form is
one of several encodings VTK readers accept), and `read_vtu` reads them back.
"""
from __future__ import annotations

import base64
import struct
import zlib
from pathlib import Path

import numpy as np

CONVERGENCE_HEADER = "Step,Time_phys_s,Domain_travel_pct,U_inlet_lat,Rho_min,MLUPS,Cd,Cl,Cs"
FORCES_HEADER = ("Step,Time_phys_s,U_inlet_lat,Fx_N,Fy_N,Fz_N,Fx_p,Fy_p,Fz_p,Fx_v,Fy_v,Fz_v,"
                 "Mx,My,Mz,Cd,Cl,Cs,Cmx,Cmy,Cmz")
BLOCK = 8


# ---------------------------------------------------------------------------
# CSV files

def write_convergence_header(path: str | Path) -> None:
    Path(path).write_text(CONVERGENCE_HEADER + "\n")


def convergence_row(step: int, time_phys: float, travel_pct: float, u_inlet: float, rho_min: float,
                    mlups: float, cd: str, cl: str, cs: str) -> str:
    """One row as main.jl:254-256 prints it; cd/cl/cs are the "%.4f" strings
    (or "N/A" without forces)."""
    return "%d,%.6e,%.6e,%.4f,%.6f,%.1f,%s,%s,%s\n" % (step, time_phys, travel_pct, u_inlet, rho_min, mlups,
                                                      cd, cl, cs)


def append_convergence(path: str | Path, *row) -> None:
    with open(path, "a") as fh:
        fh.write(convergence_row(*row))


def write_force_csv_header(path: str | Path) -> None:
    """write_force_csv_header (src/forces/io.jl:5-9)."""
    Path(path).write_text(FORCES_HEADER + "\n")


def force_row(step: int, time_phys: float, fd: dict, u_inlet: float) -> str:
    """append_force_csv (src/forces/io.jl:11-22): fd holds the force-reduction
    keys (Fx, Fx_pressure, ..., Cmz)."""
    return ("%d,%.6e,%.6f,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e,"
            "%.6f,%.6f,%.6f,%.6f,%.6f,%.6f\n") % (
        step, time_phys, float(np.float32(u_inlet)),
        fd["Fx"], fd["Fy"], fd["Fz"],
        fd["Fx_pressure"], fd["Fy_pressure"], fd["Fz_pressure"],
        fd["Fx_viscous"], fd["Fy_viscous"], fd["Fz_viscous"],
        fd["Mx"], fd["My"], fd["Mz"],
        fd["Cd"], fd["Cl"], fd["Cs"], fd["Cmx"], fd["Cmy"], fd["Cmz"])


def append_force_csv(path: str | Path, step: int, time_phys: float, fd: dict, u_inlet: float) -> None:
    with open(path, "a") as fh:
        fh.write(force_row(step, time_phys, fd, u_inlet))


def read_csv(path: str | Path) -> tuple:
    """(header list, rows as lists of strings)."""
    lines = Path(path).read_text().splitlines()
    return lines[0].split(","), [ln.split(",") for ln in lines[1:] if ln]


# ---------------------------------------------------------------------------
# Surface pressure binary (.lbmp)

def _object_table(n_tri: int, object_names):
    if object_names is None:
        names = ["default"]
        ids = np.ones(n_tri, np.int32)
    else:
        names = list(dict.fromkeys(object_names))   # first-appearance order, as Julia's unique
        idx = {n: i + 1 for i, n in enumerate(names)}
        ids = np.array([idx[n] for n in object_names], np.int32)
    return names, ids


def export_surface_pressure_bin(path: str | Path, triangles: np.ndarray, normals: np.ndarray,
                                pressure: np.ndarray, shear: tuple, rho_ref: float, u_ref: float,
                                object_names=None, offset=(0.0, 0.0, 0.0)) -> None:
    """export_surface_pressure_bin (src/forces/io.jl:230-335).

    triangles (n,3,3) and normals (n,3) are the mesh the maps belong to (the
    reference passes the unrotated mesh with no offset); pressure and shear
    are the Float32 per-triangle maps of the force reduction."""
    tris = np.asarray(triangles, np.float64).reshape(-1, 3, 3)
    n = tris.shape[0]
    p = np.asarray(pressure, np.float32)
    sx, sy, sz = (np.asarray(a, np.float32) for a in shear)
    shear_mag = np.sqrt(sx * sx + sy * sy + sz * sz)             # Float32 throughout
    q_inf = np.float32(0.5 * rho_ref * u_ref ** 2)
    cp = p / q_inf if q_inf > np.float32(1e-10) else np.zeros(n, np.float32)
    names, ids = _object_table(n, object_names)
    off = np.array([np.float32(v) for v in offset], np.float32)
    verts = tris.astype(np.float32) + off                          # Float32(v) + offset
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    with open(path, "wb") as fh:
        fh.write(b"LBMP")
        fh.write(struct.pack("<III", 1, n, len(names)))
        for name in names:
            b = name.encode("utf-8")
            fh.write(struct.pack("<I", len(b)))
            fh.write(b)
        fh.write(np.ascontiguousarray(verts, "<f4").tobytes())
        fh.write(np.ascontiguousarray(np.asarray(normals, np.float64).astype(np.float32), "<f4").tobytes())
        for arr in (p, shear_mag, cp.astype(np.float32)):
            fh.write(np.ascontiguousarray(arr, "<f4").tobytes())
        fh.write(np.ascontiguousarray(ids, "<i4").tobytes())


def read_surface_pressure_bin(path: str | Path) -> dict:
    """The arrays of an .lbmp file (either program's)."""
    raw = Path(path).read_bytes()
    if raw[:4] != b"LBMP":
        raise ValueError(f"{path}: not an LBMP file")
    version, n, n_obj = struct.unpack_from("<III", raw, 4)
    pos = 16
    names = []
    for _ in range(n_obj):
        (ln,) = struct.unpack_from("<I", raw, pos)
        names.append(raw[pos + 4:pos + 4 + ln].decode("utf-8"))
        pos += 4 + ln

    def take(count, dt):
        nonlocal pos
        a = np.frombuffer(raw, dt, count, pos)
        pos += count * 4
        return a

    out = {"version": version, "names": names,
           "vertices": take(9 * n, "<f4").reshape(n, 3, 3), "normals": take(3 * n, "<f4").reshape(n, 3),
           "pressure": take(n, "<f4"), "shear_mag": take(n, "<f4"), "cp": take(n, "<f4"),
           "object_id": take(n, "<i4")}
    if pos != len(raw):
        raise ValueError(f"{path}: {len(raw) - pos} trailing bytes")
    return out


# ---------------------------------------------------------------------------
# VTK XML unstructured grids

VTK_TRIANGLE = 5
VTK_VOXEL = 11
_VTK_TYPES = {np.dtype("float32"): "Float32", np.dtype("float64"): "Float64", np.dtype("int32"): "Int32",
              np.dtype("int64"): "Int64", np.dtype("uint8"): "UInt8"}


def _b64(arr: np.ndarray, compress: bool) -> str:
    data = np.ascontiguousarray(arr).tobytes()
    if compress:
        body = zlib.compress(data, 6)
        # one block: [nblocks, block size, last block size, compressed sizes], UInt64
        head = struct.pack("<QQQQ", 1, len(data), len(data), len(body))
        return base64.b64encode(head).decode() + base64.b64encode(body).decode()
    return base64.b64encode(struct.pack("<Q", len(data)) + data).decode()


def write_vtu(path: str | Path, points: np.ndarray, connectivity: np.ndarray, offsets: np.ndarray,
              types: np.ndarray, cell_data: dict, compress: bool = True) -> None:
    """One-piece UnstructuredGrid; points (n,3), cell arrays (ncells,) or
    (ncells, ncomp)."""
    pts = np.ascontiguousarray(points)
    ncell = int(types.shape[0])
    enc = ' compressor="vtkZLibDataCompressor"' if compress else ""
    out = ['<?xml version="1.0"?>',
           f'<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64"{enc}>',
           "  <UnstructuredGrid>",
           f'    <Piece NumberOfPoints="{pts.shape[0]}" NumberOfCells="{ncell}">',
           "      <Points>",
           f'        <DataArray type="{_VTK_TYPES[pts.dtype]}" NumberOfComponents="3" format="binary">'
           f"{_b64(pts, compress)}</DataArray>",
           "      </Points>",
           "      <Cells>",
           f'        <DataArray type="Int64" Name="connectivity" format="binary">'
           f"{_b64(connectivity.astype(np.int64), compress)}</DataArray>",
           f'        <DataArray type="Int64" Name="offsets" format="binary">'
           f"{_b64(offsets.astype(np.int64), compress)}</DataArray>",
           f'        <DataArray type="UInt8" Name="types" format="binary">'
           f"{_b64(types.astype(np.uint8), compress)}</DataArray>",
           "      </Cells>",
           "      <CellData>"]
    for name, arr in cell_data.items():
        a = np.ascontiguousarray(arr)
        nc = 1 if a.ndim == 1 else a.shape[1]
        out.append(f'        <DataArray type="{_VTK_TYPES[a.dtype]}" Name="{name}" NumberOfComponents="{nc}" '
                   f'format="binary">{_b64(a, compress)}</DataArray>')
    out += ["      </CellData>", "    </Piece>", "  </UnstructuredGrid>", "</VTKFile>", ""]
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    Path(path).write_text("\n".join(out))


def _decode(text: str, dtype: str, compressed: bool) -> np.ndarray:
    dt = {"Float32": "<f4", "Float64": "<f8", "Int32": "<i4", "Int64": "<i8", "UInt8": "u1"}[dtype]
    if compressed:
        head = base64.b64decode(text[:44])
        nblocks, _, _, csize = struct.unpack("<QQQQ", head[:32])
        if nblocks != 1:
            raise ValueError("only single-block compressed arrays are read")
        body = base64.b64decode(text[44:])
        return np.frombuffer(zlib.decompress(body[:csize]), dt)
    raw = base64.b64decode(text)
    (nbytes,) = struct.unpack("<Q", raw[:8])
    return np.frombuffer(raw[8:8 + nbytes], dt)


def read_vtu(path: str | Path) -> dict:
    """Points, cells and cell data of a file written by write_vtu."""
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    compressed = root.get("compressor") is not None
    piece = root.find("UnstructuredGrid/Piece")
    res = {"n_points": int(piece.get("NumberOfPoints")), "n_cells": int(piece.get("NumberOfCells")), "cell_data": {}}
    pa = piece.find("Points/DataArray")
    res["points"] = _decode(pa.text, pa.get("type"), compressed).reshape(-1, 3)
    for da in piece.find("Cells"):
        res[da.get("Name")] = _decode(da.text, da.get("type"), compressed)
    for da in piece.find("CellData"):
        a = _decode(da.text, da.get("type"), compressed)
        nc = int(da.get("NumberOfComponents", "1"))
        res["cell_data"][da.get("Name")] = a.reshape(-1, nc) if nc > 1 else a
    return res


def save_surface_vtk(path: str | Path, triangles: np.ndarray, normals: np.ndarray, areas: np.ndarray,
                     pressure: np.ndarray, shear: tuple, offset=(0.0, 0.0, 0.0), object_names=None) -> Path:
    """save_surface_vtk (src/forces/io.jl:75-137): three points per triangle
    (offset applied), cell arrays Pressure, ShearStress, ShearMag, Normal,
    Area and, for several objects, ObjectID.  Returns the .vtu path."""
    tris = np.asarray(triangles, np.float64).reshape(-1, 3, 3)
    n = tris.shape[0]
    pts = (tris + np.asarray(offset, np.float64)).reshape(-1, 3)
    sx, sy, sz = (np.asarray(a, np.float32) for a in shear)
    data = {"Pressure": np.asarray(pressure, np.float32),
            "ShearStress": np.stack([sx, sy, sz], axis=1),
            "ShearMag": np.sqrt(sx * sx + sy * sy + sz * sz),
            "Normal": np.asarray(normals, np.float64).reshape(n, 3),
            "Area": np.asarray(areas, np.float64).astype(np.float32)}
    if object_names is not None:
        data["ObjectID"] = _object_table(n, object_names)[1]
    out = Path(str(path) + ".vtu")
    write_vtu(out, pts, np.arange(3 * n), 3 * np.arange(1, n + 1), np.full(n, VTK_TRIANGLE), data)
    return out


def export_flow_vtu(path: str | Path, levels: list, cfg, compress: bool = True) -> Path:
    """export_merged_mesh_sync (src/io/io_vtk.jl:6-129): every block of every
    level except those whose 8 children all exist on the next level, as 8^3
    voxels of (8+1)^3 points each.  `levels` holds, per level in order, a dict
    with "coords" ((nb,3) 1-based block coordinates), "dx", and the cell
    arrays "rho" (nb*512), "vel" ((nb*512, 3)), "obstacle" (nb*512, bool) and
    optionally "T", in block-major, x-fastest cell order.  Non-finite values
    are written as 0 and the Float32 points are (block origin + p) * dx."""
    sets = [set(map(tuple, np.asarray(L["coords"]).tolist())) for L in levels]
    chosen = []
    for li, L in enumerate(levels):
        nxt = sets[li + 1] if li + 1 < len(levels) else None
        for b, (bx, by, bz) in enumerate(np.asarray(L["coords"]).tolist()):
            if nxt is not None and all((2 * bx - 1 + dx, 2 * by - 1 + dy, 2 * bz - 1 + dz) in nxt
                                       for dz in (0, 1) for dy in (0, 1) for dx in (0, 1)):
                continue
            chosen.append((li, b, bx, by, bz))
    n = len(chosen)
    npt, ncl = (BLOCK + 1) ** 3, BLOCK ** 3
    r = np.arange(BLOCK + 1, dtype=np.float32)
    pz, py, px = np.meshgrid(r, r, r, indexing="ij")
    local = np.stack([px.ravel(), py.ravel(), pz.ravel()], axis=1)            # x fastest
    c = np.arange(BLOCK)
    cz, cy, cx = np.meshgrid(c, c, c, indexing="ij")
    base = (cx + cy * (BLOCK + 1) + cz * (BLOCK + 1) ** 2).ravel()
    s1, s2 = BLOCK + 1, (BLOCK + 1) ** 2
    vox = np.stack([base, base + 1, base + s1, base + s1 + 1, base + s2, base + s2 + 1, base + s2 + s1,
                    base + s2 + s1 + 1], axis=1)                                   # VTK_VOXEL order
    points = np.empty((n * npt, 3), np.float32)
    conn = np.empty((n * ncl, 8), np.int64)
    rho = np.empty(n * ncl, np.float32)
    vel = np.empty((n * ncl, 3), np.float32)
    obst = np.empty(n * ncl, np.uint8)
    lvl = np.empty(n * ncl, np.int32)
    has_t = all("T" in L and L["T"] is not None for L in levels)
    temp = np.empty(n * ncl, np.float32) if has_t else None
    for i, (li, b, bx, by, bz) in enumerate(chosen):
        L = levels[li]
        dx = np.float32(L["dx"])
        org = np.array([(bx - 1) * BLOCK, (by - 1) * BLOCK, (bz - 1) * BLOCK], np.float32)
        points[i * npt:(i + 1) * npt] = (org + local) * dx
        conn[i * ncl:(i + 1) * ncl] = vox + i * npt
        sl = slice(b * ncl, (b + 1) * ncl)
        o = slice(i * ncl, (i + 1) * ncl)
        rho[o] = L["rho"][sl]
        vel[o] = np.asarray(L["vel"]).reshape(-1, 3)[sl]
        obst[o] = np.asarray(L["obstacle"])[sl].astype(bool)
        lvl[o] = li + 1
        if has_t:
            temp[o] = L["T"][sl]
    for a in (rho, vel, temp):
        if a is not None:
            a[~np.isfinite(a)] = 0.0
    data = {}
    if cfg.output_density:
        data["Density"] = rho
    if cfg.output_velocity:
        data["Velocity"] = vel
    if cfg.output_velocity_magnitude:
        data["VelocityMagnitude"] = np.sqrt(vel[:, 0] ** 2 + vel[:, 1] ** 2 + vel[:, 2] ** 2)
    if cfg.output_obstacle:
        data["Obstacle"] = obst
    if cfg.output_level:
        data["Level"] = lvl
    if has_t:
        data["Temperature"] = temp
    out = Path(str(path) + ".vtu")
    write_vtu(out, points, conn.ravel(), 8 * np.arange(1, n * ncl + 1), np.full(n * ncl, VTK_VOXEL), data,
              compress)
    return out


# ---------------------------------------------------------------------------
# Net per-triangle forces

def export_net_triangle_forces_csv(path: str | Path, net: tuple, centers: np.ndarray, offset=(0.0, 0.0, 0.0),
                                   alpha_deg: float = 0.0, beta_deg: float = 0.0, object_names=None,
                                   unrotated: bool = True) -> None:
    """export_net_triangle_forces_csv (src/forces/io.jl:139-228).

    With `unrotated` (what main.jl passes), `centers` are the unrotated
    triangle centres, no offset is added, and the Float32 force vectors are
    turned back into the STL frame with R^T = (Rz(beta) Ry(alpha))^T, its
    entries rounded to Float32 and applied in Float32.  Otherwise `centers`
    are the solver mesh's Float32 centres and `offset` is added."""
    f32 = np.float32
    fx, fy, fz = (np.asarray(a, f32) for a in net)
    c = np.asarray(centers, np.float64).astype(f32)
    if unrotated:
        off = (f32(0.0), f32(0.0), f32(0.0))
    else:
        off = tuple(f32(v) for v in offset)
    if unrotated and (alpha_deg != 0.0 or beta_deg != 0.0):
        a, b = np.deg2rad(alpha_deg), np.deg2rad(beta_deg)
        ca, sa, cb, sb = np.cos(a), np.sin(a), np.cos(b), np.sin(b)
        r = [[f32(ca * cb), f32(-ca * sb), f32(sa)],
             [f32(sb), f32(cb), f32(0.0)],
             [f32(-sa * cb), f32(sa * sb), f32(ca)]]
        fx, fy, fz = (r[i][0] * fx + r[i][1] * fy + r[i][2] * fz for i in range(3))
    names = object_names if object_names is not None else ["default"] * len(fx)
    lines = ["triangle_id,object_name,cx,cy,cz,Fx,Fy,Fz"]
    cx, cy, cz = c[:, 0] + off[0], c[:, 1] + off[1], c[:, 2] + off[2]
    for i in range(len(fx)):
        lines.append("%d,%s,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e" % (i + 1, names[i], cx[i], cy[i], cz[i],
                                                             fx[i], fy[i], fz[i]))
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    Path(path).write_text("\n".join(lines) + "\n")


def export_net_triangle_forces_vtk(path: str | Path, net: tuple, centers: np.ndarray, normals: np.ndarray,
                                   areas: np.ndarray, offset=(0.0, 0.0, 0.0)) -> Path:
    """export_net_triangle_forces_vtk (src/forces/io.jl:337-444): an ASCII
    VTK PolyData of one vertex per triangle centre (Float32 centre + Float64
    offset) with NetForce, NetForceMagnitude, Normal and Area."""
    f32 = np.float32
    fx, fy, fz = (np.asarray(a, f32) for a in net)
    c = np.asarray(centers, np.float64).astype(f32).astype(np.float64)
    nrm = np.asarray(normals, np.float64).astype(f32)
    ar = np.asarray(areas, np.float64).astype(f32)
    n = len(fx)
    mag = np.sqrt(fx * fx + fy * fy + fz * fz)
    o = [float(v) for v in offset]
    L = ['<?xml version="1.0"?>',
         '<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian">',
         "  <PolyData>",
         f'    <Piece NumberOfPoints="{n}" NumberOfVerts="{n}" NumberOfLines="0" NumberOfStrips="0" '
         f'NumberOfPolys="0">',
         "      <Points>",
         '        <DataArray type="Float64" NumberOfComponents="3" format="ascii">']
    L += ["          %.8e %.8e %.8e" % (c[i, 0] + o[0], c[i, 1] + o[1], c[i, 2] + o[2]) for i in range(n)]
    L += ["        </DataArray>", "      </Points>", "      <Verts>",
          '        <DataArray type="Int64" Name="connectivity" format="ascii">']
    L += ["          %d" % i for i in range(n)]
    L += ["        </DataArray>", '        <DataArray type="Int64" Name="offsets" format="ascii">']
    L += ["          %d" % i for i in range(1, n + 1)]
    L += ["        </DataArray>", "      </Verts>", '      <PointData Vectors="NetForce">',
          '        <DataArray type="Float32" Name="NetForce" NumberOfComponents="3" format="ascii">']
    L += ["          %.6e %.6e %.6e" % (fx[i], fy[i], fz[i]) for i in range(n)]
    L += ["        </DataArray>", '        <DataArray type="Float32" Name="NetForceMagnitude" format="ascii">']
    L += ["          %.6e" % mag[i] for i in range(n)]
    L += ["        </DataArray>",
          '        <DataArray type="Float32" Name="Normal" NumberOfComponents="3" format="ascii">']
    L += ["          %.6e %.6e %.6e" % (nrm[i, 0], nrm[i, 1], nrm[i, 2]) for i in range(n)]
    L += ["        </DataArray>", '        <DataArray type="Float32" Name="Area" format="ascii">']
    L += ["          %.6e" % ar[i] for i in range(n)]
    L += ["        </DataArray>", "      </PointData>", "    </Piece>", "  </PolyData>", "</VTKFile>"]
    out = Path(str(path) + ".vtp")
    out.parent.mkdir(parents=True, exist_ok=True)
    out.write_text("\n".join(L) + "\n")
    return out
