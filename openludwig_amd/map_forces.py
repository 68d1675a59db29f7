"""Net triangle forces -> structural FE node loads, preserving force and
moment equilibrium (src/PREPOST/map_forces_to_nodes.jl).

For every triangle force (net_forces_NNNNNN.csv) the 3 nearest structural
nodes of the same object share it: the nodal forces f (3k values) solve
    sum_i f_i = F            (force equilibrium)
    sum_i r_i x f_i = 0      (moment equilibrium about the triangle centre)
with the minimum-norm solution f = pinv(A) b, A the 6 x 3k constraint matrix
(map_forces_to_nodes.jl:164-220).  One node takes the whole force.  The
pseudo-inverse cut-off is Julia's default, eps * min(size(A)) relative to the
largest singular value.

Command line:
    python -m openludwig_amd.map_forces nodes.csv net_forces.csv [output_dir]
writes nodal_loads.csv and Nodal_loads_summary.md as the reference does.
"""
from __future__ import annotations

import sys
import time
from pathlib import Path

import numpy as np


def parse_nodes_csv(path) -> list:
    """parse_nodes_csv (map_forces_to_nodes.jl:40-72): node_id, object_name,
    cx, cy, cz; malformed lines are skipped."""
    out = []
    lines = Path(path).read_text().splitlines()
    for ln in lines[1:]:
        s = ln.strip()
        if not s:
            continue
        p = s.split(",")
        if len(p) < 5:
            continue
        out.append((int(p[0].strip()), p[1].strip(), float(p[2]), float(p[3]), float(p[4])))
    return out


def parse_forces_csv(path) -> list:
    """parse_forces_csv (map_forces_to_nodes.jl:74-108): triangle_id,
    object_name, cx, cy, cz, Fx, Fy, Fz."""
    out = []
    lines = Path(path).read_text().splitlines()
    for ln in lines[1:]:
        s = ln.strip()
        if not s:
            continue
        p = s.split(",")
        if len(p) < 8:
            continue
        out.append((int(p[0].strip()), p[1].strip(), *(float(v) for v in p[2:8])))
    return out


def group_by_object(items) -> dict:
    groups: dict = {}
    for it in items:
        groups.setdefault(it[1], []).append(it)
    return groups


def find_k_nearest(c, node_coords: np.ndarray, k: int) -> np.ndarray:
    """Indices of the k nearest nodes, nearest first (squared distances)."""
    k = min(k, node_coords.shape[0])
    d = node_coords - np.asarray(c, np.float64)
    dist = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    return np.argsort(dist, kind="stable")[:k]


def skew(r) -> np.ndarray:
    return np.array([[0.0, -r[2], r[1]], [r[2], 0.0, -r[0]], [-r[1], r[0], 0.0]])


def distribute_force(F, centroid, node_positions: np.ndarray) -> np.ndarray:
    """distribute_force (map_forces_to_nodes.jl:164-220): (k, 3) nodal forces."""
    F = np.asarray(F, np.float64)
    k = node_positions.shape[0]
    if k == 1:
        return F.reshape(1, 3).copy()
    A = np.zeros((6, 3 * k))
    b = np.zeros(6)
    b[:3] = F
    for i in range(k):
        A[0, 3 * i] = A[1, 3 * i + 1] = A[2, 3 * i + 2] = 1.0
        A[3:6, 3 * i:3 * i + 3] = skew(node_positions[i] - np.asarray(centroid, np.float64))
    rcond = np.finfo(np.float64).eps * min(A.shape)
    return (np.linalg.pinv(A, rcond=rcond) @ b).reshape(k, 3)


def compute_resultant(positions: np.ndarray, forces: np.ndarray) -> tuple:
    """Total force and moment about the origin, summed in point order."""
    F = np.zeros(3)
    M = np.zeros(3)
    for (px, py, pz), (fx, fy, fz) in zip(positions, forces):
        F += (fx, fy, fz)
        M[0] += py * fz - pz * fy
        M[1] += pz * fx - px * fz
        M[2] += px * fy - py * fx
    return F, M


def map_forces(nodes: list, forces: list, log=print) -> tuple:
    """The mapping loop of main() (map_forces_to_nodes.jl:420-520): returns
    (nodal_forces {global node index: [Fx,Fy,Fz]}, per-object results)."""
    node_groups, force_groups = group_by_object(nodes), group_by_object(forces)
    nodal: dict = {}
    results = {}
    index = {(n[0], n[1]): i for i, n in reversed(list(enumerate(nodes)))}   # first match wins
    for obj in sorted(set(node_groups) | set(force_groups)):
        of, on = force_groups.get(obj), node_groups.get(obj)
        if not of:
            log(f"  [SKIP] {obj}: no triangle forces")
            continue
        if not on:
            log(f"  [WARN] {obj}: {len(of)} forces but no nodes; skipped")
            continue
        coords = np.array([[n[2], n[3], n[4]] for n in on])
        gidx = [index[(n[0], n[1])] for n in on]
        k = min(3, len(on))
        for tf in of:
            c = tf[2:5]
            near = find_k_nearest(c, coords, k)
            nf = distribute_force(tf[5:8], c, coords[near])
            for j, li in enumerate(near):
                acc = nodal.setdefault(gidx[li], np.zeros(3))
                acc += nf[j]
        F_tri, M_tri = compute_resultant(np.array([t[2:5] for t in of]), np.array([t[5:8] for t in of]))
        F_nod, M_nod = compute_resultant(coords, np.array([nodal.get(g, np.zeros(3)) for g in gidx]))
        results[obj] = {"n_nodes": len(on), "n_forces": len(of), "F_tri": F_tri, "M_tri": M_tri,
                        "F_nod": F_nod, "M_nod": M_nod}
    return nodal, results


def write_nodal_loads_csv(path, nodes: list, nodal: dict) -> None:
    """write_nodal_loads_csv (map_forces_to_nodes.jl:262-276)."""
    lines = ["node_id,object_name,cx,cy,cz,Fx,Fy,Fz"]
    for i, n in enumerate(nodes):
        F = nodal.get(i, np.zeros(3))
        lines.append("%d,%s,%.6e,%.6e,%.6e,%.6e,%.6e,%.6e" % (n[0], n[1], n[2], n[3], n[4], F[0], F[1], F[2]))
    Path(path).write_text("\n".join(lines) + "\n")


def _table(r) -> list:
    rows = ["| Quantity | Triangle Forces | Nodal Loads | Difference |",
            "|----------|---------------:|------------:|-----------:|"]
    for lab, key, i in (("Fx [N]  ", "F", 0), ("Fy [N]  ", "F", 1), ("Fz [N]  ", "F", 2),
                        ("Mx [N.m]", "M", 0), ("My [N.m]", "M", 1), ("Mz [N.m]", "M", 2)):
        a, b = r[f"{key}_tri"][i], r[f"{key}_nod"][i]
        rows.append("| %s | %+.6e | %+.6e | %+.6e |" % (lab, a, b, a - b))
    return rows


def write_summary_md(path, results: dict, glob: dict, n_nodes: int, n_forces: int, elapsed: float) -> None:
    """write_summary_md (map_forces_to_nodes.jl:278-360)."""
    L = ["# Nodal Loads Mapping Summary", "",
         "**Generated:** %s" % time.strftime("%Y-%m-%d %H:%M:%S"),
         "**Elapsed time:** %.3f s" % elapsed, "**Total nodes:** %d" % n_nodes,
         "**Total triangle forces:** %d" % n_forces, "", "## Per-Object Validation", ""]
    for obj in sorted(results):
        r = results[obj]
        L += [f"### Object: `{obj}`", "", "- **Nodes:** %d" % r["n_nodes"],
              "- **Triangle forces:** %d" % r["n_forces"], ""] + _table(r) + [""]
    L += ["## Global Validation (All Objects)", ""] + _table(glob) + [""]
    dF = np.abs(glob["F_tri"] - glob["F_nod"]).max()
    dM = np.abs(glob["M_tri"] - glob["M_nod"]).max()
    L += ["**Max force difference:**  %.2e N" % dF, "**Max moment difference:** %.2e N.m" % dM, ""]
    L.append("> **PASS** - Force and moment equilibrium preserved within machine precision."
             if dF < 1e-6 and dM < 1e-6 else
             "> **WARNING** - Differences exceed expected machine precision. Check input data.")
    Path(path).write_text("\n".join(L) + "\n")


def run(nodes_path, forces_path, output_dir=None, log=print) -> dict:
    t0 = time.time()
    for p, what in ((nodes_path, "Nodes"), (forces_path, "Forces")):
        if not Path(p).is_file():
            raise FileNotFoundError(f"[MapForces] {what} file not found: {p}")
    nodes, forces = parse_nodes_csv(nodes_path), parse_forces_csv(forces_path)
    if not nodes or not forces:
        raise ValueError("[MapForces] Empty input data. Cannot proceed.")
    out = Path(output_dir) if output_dir else (Path(forces_path).parent if str(Path(forces_path).parent) else Path("."))
    nodal, results = map_forces(nodes, forces, log)
    F_tri, M_tri = compute_resultant(np.array([f[2:5] for f in forces]), np.array([f[5:8] for f in forces]))
    keys = list(nodal)
    F_nod, M_nod = compute_resultant(np.array([nodes[g][2:5] for g in keys]).reshape(-1, 3),
                                     np.array([nodal[g] for g in keys]).reshape(-1, 3))
    glob = {"F_tri": F_tri, "M_tri": M_tri, "F_nod": F_nod, "M_nod": M_nod}
    out.mkdir(parents=True, exist_ok=True)
    write_nodal_loads_csv(out / "nodal_loads.csv", nodes, nodal)
    write_summary_md(out / "Nodal_loads_summary.md", results, glob, len(nodes), len(forces), time.time() - t0)
    return {"nodal": nodal, "objects": results, "global": glob}


def main(argv=None) -> int:
    a = sys.argv[1:] if argv is None else argv
    if len(a) < 2:
        print(__doc__)
        return 2
    run(a[0], a[1], a[2] if len(a) > 2 else None)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
