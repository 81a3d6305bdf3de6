"""DB-SPH wall elements (Domain%tipo = "bsph"): the host-side pre-processing
the device's DB-SPH terms read.

A DB-SPH frontier is a surface mesh read from ``.ply`` files
(input/07_DBSPH/surface_mesh_list.txt names them).  Each triangle of the mesh
is a wall element ("semi-particle") with a position (its centroid), an
inward normal, an area ("weight"), a semi-particle volume and a mass.  This
module restates the reference's units for that:

* :func:`tank_mesh` writes the mesh of an open-top rectangular tank (floor
  and four walls, quadrilaterals of the fluid spacing, normals into the
  tank): the automatic tank-mesh generator, for decks that have no mesh;
* :func:`read_ply` / :func:`write_ply`: the ASCII ``.ply`` subset the
  reference reads (Import_ply_surface_meshes.f90:93-111: a header of eleven
  lines, vertex positions, faces as vertex counts and 0-based indices);
* :func:`wall_elements`: Import_ply_surface_meshes.f90:161-345 (faces split
  into triangles, area and normal by area_triangle.f90), the 3D branch of
  DBSPH_IC_surface_elements.f90:89-120 (centroid, cell, normal, weight),
  DBSPH_find_close_faces.f90:53-106 (edge-adjacent triangles found through
  the grid cells) and semi_particle_volumes.f90:56-134 with
  adjacent_faces_isolated_points.f90 (depth coefficient k_d, volume, mass).

3D only; the wall elements do not move (no DBSPH kinematics records), and
there are no DB-SPH inlet or outlet sections.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .case import PIGRECO, Grid, cell_number


@dataclass
class DBSPHParams:
    """The DBSPH deck section (ReadDBSPH.f90:35-99) for one surface mesh,
    without monitors, inlets or outlets, with slip_ID = 0.  kinematics: the
    mesh's records (t, vx, vy, vz, wx, wy, wz), the first at t <= 0, with
    its rotation centre (DBSPH_kinematics.f90); empty for static walls."""
    dx_dxw: float = 1.0                 # fluid spacing / wall-element spacing
    k_w: float = 0.5                    # semi-particle coefficient
    gamma_limiter: bool = False         # Gamma <= 1
    negative_wall_p_allowed: bool = False
    fs_allowed: bool = True             # free-surface detection via sigma
    muscl_boundary: bool = False        # wall terms in the MUSCL gradients
    kinematics: tuple = ()              # ((t, vx, vy, vz, wx, wy, wz), ...)
    rotation_centre: tuple = (0.0, 0.0, 0.0)

    def deck_lines(self, nfiles=1):
        b = lambda v: ".true." if v else ".false."  # noqa: E731
        return [f"{float(self.dx_dxw)!r} {b(self.muscl_boundary)} {float(self.k_w)!r} 0 {b(self.gamma_limiter)}",
                f"{b(self.negative_wall_p_allowed)} {b(self.fs_allowed)}", "0 0",
                f"{int(nfiles)} .false.",
                " ".join([str(len(self.kinematics))] + [repr(float(v)) for v in self.rotation_centre])] + \
            [" ".join(repr(float(v)) for v in r) for r in self.kinematics] + ["0 0 4"]


def tank_mesh(tank, dx):
    """Open-top box [0,X] x [0,Y] x [0,Z]: floor, then the walls y = 0, y = Y,
    x = 0 and x = X, each as a lattice of quadrilaterals of side about dx,
    each quadrilateral's vertex order giving a normal into the tank.
    Returns (vertices (nv, 3), quads (nq, 4) 0-based)."""
    X, Y, Z = (float(v) for v in tank)
    nx, ny, nz = (max(1, int(round(v / dx))) for v in (X, Y, Z))
    verts, quads = [], []

    def patch(p0, u, v, nu, nv, inward):
        base = len(verts)
        for a in range(nu + 1):
            for b in range(nv + 1):
                verts.append(p0 + u * a / nu + v * b / nv)
        for a in range(nu):
            for b in range(nv):
                q = [base + a * (nv + 1) + b, base + (a + 1) * (nv + 1) + b,
                     base + (a + 1) * (nv + 1) + b + 1, base + a * (nv + 1) + b + 1]
                P = [verts[k] for k in q[:3]]
                if np.dot(np.cross(P[1] - P[0], P[2] - P[0]), inward) < 0:
                    q = q[::-1]
                quads.append(q)

    o = np.zeros(3)
    ex, ey, ez = np.array([X, 0.0, 0.0]), np.array([0.0, Y, 0.0]), np.array([0.0, 0.0, Z])
    patch(o, ex, ey, nx, ny, np.array([0.0, 0.0, 1.0]))
    patch(o, ex, ez, nx, nz, np.array([0.0, 1.0, 0.0]))
    patch(ey, ex, ez, nx, nz, np.array([0.0, -1.0, 0.0]))
    patch(o, ey, ez, ny, nz, np.array([1.0, 0.0, 0.0]))
    patch(ex, ey, ez, ny, nz, np.array([-1.0, 0.0, 0.0]))
    return np.array(verts, dtype=np.float64), np.array(quads, dtype=np.int64)


def write_ply(path, verts, faces):
    """ASCII .ply with the eleven header lines the reader skips."""
    hdr = ["ply", "format ascii 1.0", "comment open-top tank", "obj_info tank",
           f"element vertex {len(verts)}", "property float x", "property float y",
           "property float z", f"element face {len(faces)}",
           "property list uchar int vertex_indices", "end_header"]
    body = [f"{float(v[0])!r} {float(v[1])!r} {float(v[2])!r}" for v in verts]
    body += [f"{len(f)} " + " ".join(str(int(k)) for k in f) for f in faces]
    Path(path).write_text("\n".join(hdr + body) + "\n")


def read_ply(path):
    """The vertex and face counts from the header (ply_headings.f90), then,
    after the eleven header lines, the vertices and the faces."""
    lines = Path(path).read_text().splitlines()
    nv = nf = None
    for ln in lines[:11]:
        t = ln.split()
        if len(t) == 3 and t[0] == "element" and t[1] == "vertex":
            nv = int(t[2])
        if len(t) == 3 and t[0] == "element" and t[1] == "face":
            nf = int(t[2])
    if nv is None or nf is None:
        raise ValueError("ply header without vertex/face counts")
    verts = np.array([[float(v) for v in lines[11 + k].split()[:3]] for k in range(nv)])
    faces = []
    for k in range(nf):
        t = [int(v) for v in lines[11 + nv + k].split()]
        faces.append(t[1:1 + t[0]])
    return verts, faces


def _triangles(faces):
    """Import_ply_surface_meshes.f90:161-290 (3D): a triangle stays, a
    quadrilateral gives (1,2,3) and (1,3,4), a pentagon (1,2,5), (2,3,5),
    (3,4,5), a hexagon (1,2,6), (2,5,6), (2,3,5), (3,4,5); 0-based here."""
    tri = []
    for f in faces:
        n = len(f)
        if n == 3:
            tri.append((f[0], f[1], f[2]))
        elif n == 4:
            tri += [(f[0], f[1], f[2]), (f[0], f[2], f[3])]
        elif n == 5:
            tri += [(f[0], f[1], f[4]), (f[1], f[2], f[4]), (f[2], f[3], f[4])]
        elif n == 6:
            tri += [(f[0], f[1], f[5]), (f[1], f[4], f[5]), (f[1], f[2], f[4]), (f[2], f[3], f[4])]
        else:
            raise ValueError(f"a face of {n} vertices")
    return np.array(tri, dtype=np.int64).reshape(-1, 3)


def _area_normal(P1, P2, P3):
    """area_triangle.f90:38-49 with Vector_Product: cross product of the two
    edges from P1, area = |cross|/2, normal = cross / (2 area)."""
    a = P2 - P1
    b = P3 - P1
    n = np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])
    area = 0.5 * math.sqrt(float(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]))
    return area, n / (2.0 * area)


@dataclass
class WallElements:
    """pg_w(1:n_w) of the reference after the pre-processing."""
    coord: np.ndarray        # (n_w, 3) centroids
    normal: np.ndarray       # (n_w, 3) unit normals into the fluid
    weight: np.ndarray       # areas
    k_d: np.ndarray          # depth coefficients
    volume: np.ndarray       # semi-particle volumes
    mass: np.ndarray
    adjacent: np.ndarray     # (n_w, 3) 1-based adjacent elements, 0 = none
    cell: np.ndarray         # 1-based grid cell, 0 outside
    vel: np.ndarray          # (n_w, 3) zero: static walls

    def __len__(self):
        return len(self.weight)


def _isolated(face1, face2):
    """adjacent_faces_isolated_points.f90:45-95 (3D): the 1-based index of the
    vertex of each triangle that the other does not share (from the sum of
    the shared vertices' indices), 0 when they do not share an edge."""
    t1 = t2 = 0
    for i in range(3):
        for j in range(3):
            if face1[i][0] == face2[j][0] and face1[i][1] == face2[j][1] and face1[i][2] == face2[j][2]:
                t1 += i + 1
                t2 += j + 1
                break
    iso = {5: 1, 4: 2, 3: 3}
    return iso.get(t1, 0), iso.get(t2, 0)


def wall_elements(verts, faces, grid: Grid, dx, den0, params: DBSPHParams) -> WallElements:
    """The wall elements of a mesh (see the module docstring for the units)."""
    verts = np.asarray(verts, dtype=np.float64)
    tri = _triangles(faces)
    nw = len(tri)
    coord = np.zeros((nw, 3))
    normal = np.zeros((nw, 3))
    weight = np.zeros(nw)
    for i, (a, b, c) in enumerate(tri):
        weight[i], normal[i] = _area_normal(verts[a], verts[b], verts[c])
        coord[i] = (verts[a] + verts[b] + verts[c]) / 3.0
    cell = cell_number(grid, coord)
    # wall elements per cell in ascending id (OrdGrid1.f90:108-137)
    order = np.lexsort((np.arange(nw), cell))
    cstart = {}
    for k, i in enumerate(order):
        cstart.setdefault(int(cell[i]), []).append(int(i))
    ni, nj, nk = (int(v) for v in grid.ncd)
    adjacent = np.zeros((nw, 3), np.int64)
    P = verts[tri]  # (nw, 3, 3)
    Pt = [tuple(tuple(float(c) for c in v) for v in t) for t in P]  # exact coordinate tuples
    for i in range(nw):
        if cell[i] == 0:
            continue
        c0 = int(cell[i]) - 1
        ic, jc, kc = c0 % ni + 1, (c0 // ni) % nj + 1, c0 // (ni * nj) + 1
        nadj = 0
        done = False
        for jr in range(jc - 1, jc + 2):
            for ir in range(ic - 1, ic + 2):
                for kr in range(kc - 1, kc + 2):
                    if not (1 <= ir <= ni and 1 <= jr <= nj and 1 <= kr <= nk):
                        continue
                    cid = ((kr - 1) * nj + (jr - 1)) * ni + ir
                    for j in cstart.get(cid, ()):
                        if j == i or done:
                            continue
                        common = 0
                        for u in range(3):
                            for v in range(3):
                                if Pt[i][u] == Pt[j][v]:
                                    common += 1
                        if common >= 2:
                            adjacent[i, nadj] = j + 1
                            nadj += 1
                            if nadj == 3:
                                done = True
        # semi_particle_volumes.f90:56-134
    k_d = np.zeros(nw)
    for i in range(nw):
        alfa_sum = 0.0
        nadj = 0
        for j1 in adjacent[i]:
            if j1 == 0:
                continue
            j = int(j1) - 1
            nadj += 1
            i0, ib = _isolated(P[i], P[j])
            if i0 == 0 or ib == 0:
                raise ValueError(f"wall elements {i + 1} and {j + 1} are not adjacent")
            s = (normal[i][0] * normal[j][0] + normal[i][1] * normal[j][1]) + normal[i][2] * normal[j][2]
            d = P[j, ib - 1] - P[i, i0 - 1]
            s2 = (normal[i][0] * d[0] + normal[i][1] * d[1]) + normal[i][2] * d[2]
            if s2 >= 0.000001:
                alfa = PIGRECO + math.acos(s)
            elif s2 <= -0.000001:
                alfa = PIGRECO - math.acos(s)
            else:
                alfa = PIGRECO
            alfa_sum = alfa_sum + alfa
        if alfa_sum > (nadj * PIGRECO * 0.5):
            k_d[i] = alfa_sum / (PIGRECO * 0.5 * nadj) - 1.0
    volume = k_d * params.k_w * weight * dx / params.dx_dxw
    mass = volume * den0
    return WallElements(coord, normal, weight, k_d, volume, mass, adjacent, cell.astype(np.int32),
                        np.zeros((nw, 3)))
