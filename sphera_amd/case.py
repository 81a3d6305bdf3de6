"""Host-side set-up of a 3D dam-break solve: geometry, grid, initial particles,
SA-SPH frontier tables.

This is the part of the reference's pre-processing the device path needs in
order to run without the Fortran front end.  Every function restates one
reference unit (cited per function) so that a case built here is the case the
reference builds from the equivalent v10 deck (``oracle/decks.py`` writes that
deck from the same :class:`DamBreak3D`).  The tables are plain numpy arrays in
the reference's own layout and units; they are handed to the C ABI unchanged.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# Static_allocation_module.f90 constants
PIGRECO = 3.14159265358979
KERNELCONST3D = 1.0 / PIGRECO
FI_STEPS = 8
TETA_STEPS = 4 * FI_STEPS
BITROWS = FI_STEPS * TETA_STEPS
INT_KERNELTABLE = 40
MAX_POSITIVE = 3.4e38

# boundary type codes carried to the device (Tratto%tipo)
BT_FIXE, BT_TAPI, BT_OPEN, BT_SOUR, BT_VELO, BT_FLOW, BT_PERI = 0, 1, 2, 3, 4, 5, 6
_BT_NAMES = {"fixe": BT_FIXE, "tapi": BT_TAPI, "open": BT_OPEN, "sour": BT_SOUR,
             "velo": BT_VELO, "flow": BT_FLOW, "peri": BT_PERI}


@dataclass
class DamBreak3D:
    """A rectangular tank (open top, SA-SPH "fixed" walls) with a water column
    in the corner x=y=z=0, hydrostatic at rest.  Units SI."""
    tank: tuple = (1.6, 0.6, 0.6)
    column: tuple = (0.4, 0.6, 0.3)
    dx: float = 0.005
    trunc: float = 1.3            # h/dx
    rho0: float = 1000.0
    bulk: float | None = None     # K; None -> rho0*(10*sqrt(g*H))^2 (linear EoS)
    alfa_mon: float = 0.02
    mu: float = 1.0e-3            # dynamic viscosity as the deck gives it
    grav: tuple = (0.0, 0.0, -9.81)
    p_ref: float = 0.0
    cfl: float = 0.05
    vsc_coeff: float = 0.05
    pesodt: float = 0.0
    teta_p: float = 0.2
    teta_v: float = 0.2
    coefnmaxpartj: float = 2.0
    maxcloseboundfaces: int = 100
    slip_coeff: float = 0.0       # slip_coefficient_mode=1 input value
    slip_mode: int = 1            # 1: slip coefficient from input; 2: SA-SPH wall function
    roughness: tuple = (0.0, 0.0)  # slip_coefficient_mode=2 inputs (eps_sigma_z, H_sgr)
    density_thresholds: int = 0   # 1: Leapfrog_continuity's 0.9/1.1 rho0 clamp
    time_split: int = 1           # 1: leapfrog; 0: explicit RK (RKscheme 1, Euler)
    # (x0, x1, z_top): a block on the floor across the tank's width, x0 < x < x1,
    # whose upstream, top and downstream faces join the frontier zone (two
    # convex edges, FindBoundaryConvexEdges3D.f90); None: no block
    obstacle: tuple | None = None
    # True: the tank's downstream wall (x = tank length) is an open face
    # (zone 3, type "open"): particles leaving through it are removed
    outlet: bool = False
    # DB-SPH walls instead of SA-SPH ones (Domain%tipo = "bsph"): a
    # sphera_amd.dbsph.DBSPHParams; the tank's surface mesh is
    # sphera_amd.dbsph.tank_mesh (explicit RK1 only)
    dbsph: object = None
    name: str = "dam_break_3D"
    # monitoring lines of the deck's "control lines" section:
    # (label, first point (x,y,z), second point, number of points)
    control_lines: tuple = ()

    def __post_init__(self):
        if self.bulk is None:
            g = math.sqrt(sum(c * c for c in self.grav))
            c0 = 10.0 * math.sqrt(g * self.column[2])
            self.bulk = self.rho0 * c0 * c0

    @property
    def frontier_slip_input(self):
        """The frontier zone's BC_shear_stress_input: the slip coefficient
        (mode 1) or the roughness length 2 eps/5 + H/5 (mode 2,
        ReadInputBoundaries.f90:207-208)."""
        if self.slip_mode == 2:
            return 2.0 * self.roughness[0] / 5.0 + self.roughness[1] / 5.0
        return self.slip_coeff if self.slip_mode == 1 else 0.0

    # -- derived scalars (ReadInput.f90:198-205, Gest_Input.f90:266-289) ----
    @property
    def h(self):
        return self.dx * self.trunc

    def scalars(self):
        h = self.h
        s = dict(h=h, doubleh=2.0 * h, squareh=h * h, square_doubleh=(2.0 * h) * (2.0 * h),
                 eta=1.0e-3 * h, eta2=1.0e-2 * (h * h))
        h3 = h * h * h
        s["coefke"] = 0.477464 / h3
        s["coefkacl"] = 0.074603 / h3
        s["nmaxpartj"] = int(self.coefnmaxpartj * (h * 4.0 / self.dx) ** 3) + 1
        s["celerita"] = math.sqrt(self.bulk / self.rho0)
        s["kin_visc"] = self.mu / self.rho0
        s["pvolume"] = self.dx * self.dx * self.dx
        return s

    # -- geometry -----------------------------------------------------------
    def vertices(self):
        """Vertex table (1-based ids): 1-8 tank corners, 9-16 column corners,
        17-24 the block's corners if there is one.  Corner order per box:
        bottom (x0y0, x1y0, x1y1, x0y1), then top."""
        def box(x1, y1, z1, x0=0.0):
            return [(x0, 0.0, 0.0), (x1, 0.0, 0.0), (x1, y1, 0.0), (x0, y1, 0.0),
                    (x0, 0.0, z1), (x1, 0.0, z1), (x1, y1, z1), (x0, y1, z1)]
        v = box(*self.tank) + box(*self.column)
        if self.obstacle is not None:
            x0, x1, zt = self.obstacle
            v += box(x1, self.tank[1], zt, x0)
        return np.array(v, dtype=np.float64)

    def faces(self):
        """Face table rows (n1,n2,n3,n4,zone).  Node order makes the normal
        U1 x U2 (DefineLocalSystemVersors.f90) point into the fluid."""
        def box_faces(o, top):
            f = [(o + 1, o + 5, o + 6, o + 2), (o + 4, o + 3, o + 7, o + 8),
                 (o + 1, o + 4, o + 8, o + 5), (o + 2, o + 6, o + 7, o + 3),
                 (o + 1, o + 2, o + 3, o + 4)]
            if top:
                f.append((o + 7, o + 6, o + 5, o + 8))
            return f
        rows = [fc + (1,) for fc in box_faces(0, False)]
        if self.outlet:  # the x = tank length wall (4th face) opens
            rows[3] = rows[3][:4] + (3,)
        rows += [fc + (2,) for fc in box_faces(8, True)]
        if self.obstacle is not None:
            # the block's upstream (normal -x), top (+z) and downstream (+x) faces
            o = 16
            rows += [(o + 5, o + 8, o + 4, o + 1, 1), (o + 5, o + 6, o + 7, o + 8, 1),
                     (o + 3, o + 7, o + 6, o + 2, 1)]
        return rows

    def zone_types(self):
        z = {1: "fixe", 2: "peri"}
        if self.outlet:
            z[3] = "open"
        return z

    def fluid_zone_box(self):
        """Lower and upper corner of the perimeter zone the lattice fills
        (the column box: vertices 9-16)."""
        col = self.vertices()[8:16]
        return col.min(0), col.max(0)

    def free_surface_height(self):
        """The "qp" height of the perimeter zone's hydrostatic pressure."""
        return float(self.column[2])


# ---------------------------------------------------------------------------
# Grid: CreaGrid.f90:58-74
# ---------------------------------------------------------------------------
@dataclass
class Grid:
    extr_lo: np.ndarray
    extr_hi: np.ndarray
    ncd: np.ndarray
    dcd: np.ndarray

    @property
    def nmax(self):
        return int(self.ncd[0] * self.ncd[1] * self.ncd[2])


def make_grid(domain_lo, domain_hi, dx, doubleh):
    epsi = 0.01 * dx
    lo = np.array([c - doubleh - epsi for c in domain_lo])
    hi = np.array([c + doubleh + epsi for c in domain_hi])
    dextr = hi - lo
    ncd = np.array([int(np.rint(d / doubleh)) if abs(d / doubleh - math.floor(d / doubleh)) != 0.5
                    else int(math.floor(d / doubleh + 0.5)) for d in dextr], dtype=np.int32)
    dcd = dextr / ncd
    return Grid(lo, hi, ncd, dcd)


def cell_number(g: Grid, coord):
    """ParticleCellNumber.f90:52-62 + CellNumber.f90:53-59, vectorised.
    Returns 1-based cell id, 0 outside the grid."""
    c = np.atleast_2d(coord)
    idx = np.ceil((c - g.extr_lo) / g.dcd).astype(np.int64)
    ni, nj, nk = (int(v) for v in g.ncd)
    ok = ((idx[:, 0] >= 1) & (idx[:, 0] <= ni) & (idx[:, 1] >= 1) & (idx[:, 1] <= nj)
          & (idx[:, 2] >= 1) & (idx[:, 2] <= nk))
    cid = ((idx[:, 2] - 1) * nj + (idx[:, 1] - 1)) * ni + idx[:, 0]
    return np.where(ok, cid, 0).astype(np.int32)


# ---------------------------------------------------------------------------
# Faces: DefineLocalSystemVersors.f90:55-133, CompleteBoundaries3D.f90
# ---------------------------------------------------------------------------
@dataclass
class Faces:
    """SA-SPH frontier faces in BFaceList order (faces of non-peri zones)."""
    face_id: np.ndarray      # reference face index (1-based)
    nodes: np.ndarray        # 3 or 4
    node_ids: np.ndarray     # (nf,4) vertex ids (1-based) of the face's nodes
    btype: np.ndarray        # BT_* code of the face's zone
    zone: np.ndarray
    area: np.ndarray
    T: np.ndarray            # (nf,3,3) T[:, :, 2] is the normal
    GX: np.ndarray           # (nf,4,3) node global coords
    LX: np.ndarray           # (nf,4,2) node local coords
    velocity: np.ndarray     # (nf,3) BoundaryFace%velocity
    tratto_velocity: np.ndarray  # (nf,3) Tratto%velocity
    slip_mode: np.ndarray
    slip_coeff: np.ndarray
    laminar_check: np.ndarray


def _local_system(gx):
    nn = len(gx)
    sidek = nn - 2
    nidx = {1: (0, 1, 2), 2: (0, 2, 3)}[sidek]
    u1 = gx[nidx[0]] - gx[nidx[2]]
    u2 = gx[nidx[1]] - gx[nidx[2]]
    u1len = math.sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2])
    rr = u1 / u1len
    w = np.array([u1[1] * u2[2] - u1[2] * u2[1], u1[2] * u2[0] - u1[0] * u2[2],
                  u1[0] * u2[1] - u1[1] * u2[0]])
    wlen = math.sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2])
    area = wlen / float(3 - sidek)
    nrm = w / wlen
    ss = np.array([nrm[1] * rr[2] - nrm[2] * rr[1], nrm[2] * rr[0] - nrm[0] * rr[2],
                   nrm[0] * rr[1] - nrm[1] * rr[0]])
    T = np.stack([rr, ss, nrm], axis=1)
    ref = gx[nn - 1]
    lx = np.zeros((4, 2))
    for n in range(nn - 1):
        for i in range(2):
            acc = 0.0
            for j in range(3):
                acc = acc + T[j, i] * (gx[n][j] - ref[j])
            lx[n, i] = acc
    return T, lx, area


def make_faces(case: DamBreak3D) -> Faces:
    vert = case.vertices()
    rows = case.faces()
    ztype = case.zone_types()
    recs = []
    # CompleteBoundaries3D: faces ordered by zone, then by face index
    for zone in sorted(ztype):
        for fi, row in enumerate(rows, start=1):
            if row[4] != zone or ztype[zone] == "peri":
                continue
            gx = np.array([vert[n - 1] for n in row[:4]])
            T, lx, area = _local_system(gx)
            recs.append((fi, 4, _BT_NAMES[ztype[zone]], zone, area, T, gx, lx, row[:4]))
    nf = len(recs)
    return Faces(
        face_id=np.array([r[0] for r in recs], np.int32),
        nodes=np.array([r[1] for r in recs], np.int32),
        node_ids=np.array([r[8] for r in recs], np.int32).reshape(nf, 4),
        btype=np.array([r[2] for r in recs], np.int32),
        zone=np.array([r[3] for r in recs], np.int32),
        area=np.array([r[4] for r in recs]),
        T=np.array([r[5] for r in recs]).reshape(nf, 3, 3),
        GX=np.array([r[6] for r in recs]).reshape(nf, 4, 3),
        LX=np.array([r[7] for r in recs]).reshape(nf, 4, 2),
        velocity=np.zeros((nf, 3)), tratto_velocity=np.zeros((nf, 3)),
        slip_mode=np.full(nf, case.slip_mode, np.int32),
        slip_coeff=np.full(nf, case.frontier_slip_input),
        laminar_check=np.zeros(nf, np.int32))


@dataclass
class ConvexEdges:
    """BoundaryConvexEdge (Hybrid_allocation_module.f90) of the frontier faces:
    per edge the two faces (0-based indices into Faces), the first node's
    position, the vector to the second node and its length."""
    face: np.ndarray       # (ne,2) int32
    gx: np.ndarray         # (ne,3)
    component: np.ndarray  # (ne,3)
    length: np.ndarray     # (ne,)

    def __len__(self):
        return len(self.length)


def find_convex_edges(faces: Faces, verts) -> ConvexEdges:
    """FindBoundaryConvexEdges3D.f90:58-145 over the fixe/tapi faces in
    BFaceList order.  For face nf and each later face nf1 sharing an edge
    (two node ids) the node of nf1 after the shared edge is tested against
    nf's plane (normal T(:,3), measured from nf's last node): below it, the
    edge is convex.  At most one edge is kept per face nf (the loops over nf1
    end at the first convex edge found)."""
    nxt = {3: (1, 2, 0), 4: (1, 2, 3, 0)}   # iseg: the node after node i (0-based)
    ok = [int(b) in (BT_FIXE, BT_TAPI) for b in faces.btype]
    fl, gx, comp, ln = [], [], [], []
    nfc = len(faces.face_id)
    for a in range(nfc - 1):
        if not ok[a]:
            continue
        na = int(faces.nodes[a])
        ida = faces.node_ids[a]
        for b in range(a + 1, nfc):
            if not ok[b]:
                continue
            nb = int(faces.nodes[b])
            idb = faces.node_ids[b]
            found = False
            for i in range(na):
                nodi, nodj = int(ida[i]), int(ida[nxt[na][i]])
                for ii in range(nb):
                    jj = nxt[nb][ii]
                    nodii, nodjj = int(idb[ii]), int(idb[jj])
                    if (nodi == nodii and nodj == nodjj) or (nodi == nodjj and nodj == nodii):
                        nodkk = int(idb[nxt[nb][jj]])
                        ref = faces.GX[a][na - 1]
                        z = 0.0
                        for sd in range(3):
                            z = z + faces.T[a][sd, 2] * (verts[nodkk - 1][sd] - ref[sd])
                        if z < 0.0:
                            fl.append((a, b))
                            gx.append([float(verts[nodi - 1][sd]) for sd in range(3)])
                            d = [verts[nodj - 1][sd] - verts[nodi - 1][sd] for sd in range(3)]
                            l2 = 0.0
                            for sd in range(3):
                                l2 = l2 + d[sd] * d[sd]
                            comp.append(d)
                            ln.append(math.sqrt(l2))
                            found = True
                            break
                if found:
                    break
            if found:
                break
    ne = len(ln)
    return ConvexEdges(np.array(fl, np.int32).reshape(ne, 2), np.array(gx, np.float64).reshape(ne, 3),
                       np.array(comp, np.float64).reshape(ne, 3), np.array(ln, np.float64))


def grid_face_incidence(g: Grid, faces: Faces, maxclose: int):
    """GridCellBoundaryFacesIntersections3D.f90:76-200 restated with numpy over
    cells.  Returns (count[nmax], start[nmax] 1-based into vector, vector of
    indices into ``faces`` (0-based))."""
    ni, nj, nk = (int(v) for v in g.ncd)
    nmax = g.nmax
    cid = np.arange(1, nmax + 1)
    k = (cid - 1) // (ni * nj) + 1
    r = cid - ni * nj * (k - 1)
    j = (r - 1) // ni + 1
    i = r - ni * (j - 1)
    lo = np.stack([g.extr_lo[0] + g.dcd[0] * (i - 1), g.extr_lo[1] + g.dcd[1] * (j - 1),
                   g.extr_lo[2] + g.dcd[2] * (k - 1)], axis=1)
    hi = np.stack([g.extr_lo[0] + g.dcd[0] * i, g.extr_lo[1] + g.dcd[1] * j,
                   g.extr_lo[2] + g.dcd[2] * k], axis=1)
    hits = np.zeros((nmax, len(faces.face_id)), bool)
    for f in range(len(faces.face_id)):
        nn = int(faces.nodes[f])
        gx = faces.GX[f, :nn]
        found = np.ones(nmax, bool)
        for sd in range(3):
            v = gx[:, sd][None, :]
            below = (v < lo[:, sd:sd + 1]).sum(1)
            eqlo = (v == lo[:, sd:sd + 1]).sum(1)
            eqhi = (v == hi[:, sd:sd + 1]).sum(1)
            above = (v > hi[:, sd:sd + 1]).sum(1)
            # values in (lo,hi) are neither; the reference's elseif chain puts
            # v==lo into -1 and v==hi into +1 only when not below lo
            ex = (((below + eqlo) == nn) & (eqlo != nn)) | (((above + eqhi) == nn) & (eqhi != nn))
            found &= ~ex
        # second phase: the 8 cell corners against the face plane
        ref = faces.GX[f, nn - 1]
        nrm = faces.T[f, :, 2]
        neg = np.zeros(nmax, np.int32)
        zer = np.zeros(nmax, np.int32)
        pos = np.zeros(nmax, np.int32)
        for ii in range(2):
            for jj in range(2):
                for kk in range(2):
                    corner = np.stack([(lo if ii == 0 else hi)[:, 0], (lo if jj == 0 else hi)[:, 1],
                                       (lo if kk == 0 else hi)[:, 2]], axis=1)
                    z = np.zeros(nmax)
                    for sd in range(3):
                        z = z + nrm[sd] * (corner[:, sd] - ref[sd])
                    neg += z < 0.0
                    zer += z == 0.0
                    pos += z > 0.0
        ex2 = (((neg + zer) == 8) & (zer != 4)) | (((pos + zer) == 8) & (zer != 4))
        hits[:, f] = found & ~ex2
    count = hits.sum(1).astype(np.int32)
    if count.max(initial=0) > maxclose:
        raise ValueError("too many faces crossing a grid cell (MAXCLOSEBOUNDFACES)")
    start = np.zeros(nmax, np.int32)
    nz = count > 0
    start[nz] = (np.cumsum(count) - count)[nz] + 1
    cells, fidx = np.nonzero(hits)
    return count, start, fidx.astype(np.int32)


# ---------------------------------------------------------------------------
# SA-SPH IC tables: ComputeKernelTable.f90, IWro2dro.f90, JdWsRn.f90,
# ComputeBoundaryIntegralTab.f90
# ---------------------------------------------------------------------------
def iwro2dro(ro):
    a1, a2, a3, a4 = 0.333333333333333, 0.3, 0.125, 0.25
    b1, b2, b3 = 0.0625, 0.025, 4.16666666666667e-03
    if 0.0 <= ro < 1.0:
        ro2 = ro * ro
        ro3 = ro2 * ro
        v = (a1 - a2 * ro2 + a3 * ro3) * ro3
        return KERNELCONST3D * (a4 - v)
    if 1.0 <= ro < 2.0:
        ro2 = ro * ro
        d = 2.0 - ro
        d2 = d * d
        d4 = d2 * d2
        v = -(b1 * ro2 + b2 * d * ro + b3 * d2) * d4
        return KERNELCONST3D * (-v)
    return 0.0


def jdwsrn(ro, n):
    """Beta-spline branch (kernel=1) of JdWsRn.f90."""
    KS3D = 0.31830989
    ro2 = ro * ro
    ro3 = ro2 * ro
    ro4 = ro2 * ro2
    ro5 = ro4 * ro
    v = 0.0
    if n == 1:
        if 0.0 <= ro < 1.0:
            v = -0.75 + (1.0 - 0.5625 * ro) * ro3
        elif 1.0 <= ro < 2.0:
            v = -1.0 + (1.5 - 1.0 * ro + 0.1875 * ro2) * ro2
    elif n == 2:
        if 0.0 <= ro < 1.0:
            v = -0.7 + (0.75 - 0.45 * ro) * ro4
        elif 1.0 <= ro < 2.0:
            v = -0.8 + (1.0 - 0.75 * ro + 0.15 * ro2) * ro3
    elif n == 3:
        if 0.0 <= ro < 1.0:
            v = -0.75 + (0.6 - 0.375 * ro) * ro5
        elif 1.0 <= ro < 2.0:
            v = -0.8 + (0.75 - 0.6 * ro + 0.125 * ro2) * ro4
    return v * KS3D


def kernel_table(h):
    ktdelta = 2.0 / INT_KERNELTABLE
    squareh = h * h
    unosuh = 1.0 / h
    unosusquareh = 1.0 / squareh
    tab = np.zeros((INT_KERNELTABLE + 1, 5))
    for nr in range(INT_KERNELTABLE + 1):
        rob = ktdelta * nr
        tab[nr, 0] = rob
        tab[nr, 1] = iwro2dro(rob)
        tab[nr, 2] = jdwsrn(rob, 1) * unosusquareh
        tab[nr, 3] = jdwsrn(rob, 2) * unosuh
        tab[nr, 4] = jdwsrn(rob, 3)
    return tab, ktdelta


def boundary_integral_tab(doubleh):
    fi_int = PIGRECO / 2.0
    teta_int = PIGRECO + PIGRECO
    dfi = fi_int / FI_STEPS
    dfimz = 0.5 * dfi
    duesen = 2.0 * math.sin(dfimz)
    dteta = teta_int / TETA_STEPS
    tab = np.zeros((BITROWS, 7))
    p = 0
    fi = -0.5 * dfi
    for _ in range(FI_STEPS):
        fi = fi + dfi
        teta = -0.5 * dteta
        zp = -doubleh * math.cos(fi)
        for _ in range(TETA_STEPS):
            teta = teta + dteta
            xp = doubleh * math.sin(fi) * math.cos(teta)
            yp = doubleh * math.sin(fi) * math.sin(teta)
            da = dteta * duesen * math.sin(fi)
            tab[p] = (xp, yp, zp, da, xp / doubleh, yp / doubleh, zp / doubleh)
            p += 1
    return tab


# ---------------------------------------------------------------------------
# Initial particles: GeneratePart.f90:230-280, SetParticles.f90:86-152,
# SetParticleParameters.f90:54-131, SubCalcPreIdro.f90:60-196 ("qp" branch)
# ---------------------------------------------------------------------------
def _axis_points(xmin, npps, dx, offset=None):
    """SetParticles.f90:77-95: PX = Xmin + offset, then npps steps of dx;
    the offset is -dx/2, or for DB-SPH walls +dx/4 along x and -dx/4 along
    y and z."""
    pts = []
    p = xmin - dx * 0.5 if offset is None else xmin + offset
    for _ in range(npps):
        p = p + dx
        pts.append(p)
    return np.array(pts)


def fortran_nint(x):
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def initial_particles(case: DamBreak3D):
    """Lattice of the column zone, hydrostatic pressure, inverse linear EoS.
    Particle order is the reference's: x outer, y, z inner."""
    s = case.scalars()
    vert = case.vertices()
    xmin_all = vert.min(0)
    zmin, zmax = case.fluid_zone_box()
    npps, xres = [], []
    for i in range(3):
        nbefore = fortran_nint((zmin[i] - xmin_all[i]) / case.dx)
        xr = xmin_all[i] + nbefore * case.dx
        xres.append(xr)
        npps.append(fortran_nint((zmax[i] - xr) / case.dx))
    if getattr(case, "dbsph", None) is not None:
        offs = (0.25 * case.dx, -(0.25 * case.dx), -(0.25 * case.dx))
    else:
        offs = (None, None, None)
    ax = [_axis_points(xres[i], npps[i], case.dx, offs[i]) for i in range(3)]
    X, Y, Z = np.meshgrid(ax[0], ax[1], ax[2], indexing="ij")
    coord = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    inside = np.all((coord > zmin) & (coord < zmax), axis=1)
    coord = np.ascontiguousarray(coord[inside])
    n = len(coord)
    g = case.grav
    gmod = math.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
    coshor = -g[2] / gmod
    senhor = g[0] / gmod
    zq = case.free_surface_height()
    affond = (zq - coord[:, 2]) * coshor + coord[:, 0] * senhor
    if getattr(case, "dbsph", None) is not None:
        # SubCalcPreIdro.f90:184-187 for DB-SPH walls: a zero pressure (signed
        # as the product it is written as), the density rho0
        pres = 0.0 * (affond * case.rho0 * gmod) * (1.0 - coord[:, 0] / 0.5925)
    else:
        pres = (affond * case.rho0 * gmod) + case.p_ref
    dens = case.rho0 * (1.0 + (pres - case.p_ref) / case.bulk)
    mass = np.full(n, s["pvolume"] * case.rho0)
    return dict(coord=coord, vel=np.zeros((n, 3)), dens=dens, pres=pres, mass=mass)


def domain_box(case: DamBreak3D):
    """Domain%coord: bounding box of all vertices (the grid spans it)."""
    v = case.vertices()
    return v.min(0), v.max(0)


def control_line_points(lines):
    """Points of the monitoring lines, as ReadInputControlLines.f90:97-124
    lays them out: line i holds the 1-based points icont[i,0]..icont[i,1];
    the step is (second - first)/(ndiv - 1) and each point adds it to the
    previous one.  Returns (icont int32 (nlines, 2), points (npoints, 3))."""
    icont, pts, npts = [], [], 0
    for _label, first, second, ndiv in lines:
        ndiv = int(ndiv)
        v1 = np.array(first, dtype=np.float64)
        icont.append((npts + 1, npts + ndiv))
        npts += ndiv
        with np.errstate(divide="ignore", invalid="ignore"):
            v3 = (np.array(second, dtype=np.float64) - v1) / float(ndiv - 1)
            for _ in range(ndiv):
                pts.append(v1.copy())
                v1 = v1 + v3
    return (np.array(icont, dtype=np.int32).reshape(-1, 2),
            np.array(pts, dtype=np.float64).reshape(-1, 3))


@dataclass
class CaseTables:
    case: DamBreak3D
    scalars: dict
    grid: Grid
    faces: Faces
    gcbf_count: np.ndarray
    gcbf_start: np.ndarray
    gcbf_vector: np.ndarray
    kerneltab: np.ndarray
    ktdelta: float
    bitab: np.ndarray
    particles: dict = field(default_factory=dict)
    edges: "ConvexEdges" = None
    walls: object = None   # DB-SPH wall elements (sphera_amd.dbsph.WallElements)


def build(case: DamBreak3D, particles=True, incidence=True) -> CaseTables:
    """All set-up tables of a case.  particles=False skips the lattice IC (the
    caller uploads its own particles); incidence=False leaves the grid-cell /
    face incidence empty (for runs without SA-SPH frontiers)."""
    s = case.scalars()
    lo, hi = domain_box(case)
    g = make_grid(lo, hi, case.dx, s["doubleh"])
    f = make_faces(case)
    if incidence:
        cnt, st, vec = grid_face_incidence(g, f, case.maxcloseboundfaces)
    else:
        nmax = int(g.nmax)
        cnt, st, vec = np.zeros(nmax, np.int32), np.ones(nmax, np.int32), np.zeros(0, np.int32)
    kt, ktd = kernel_table(s["h"])
    bit = boundary_integral_tab(s["doubleh"])
    parts = initial_particles(case) if particles else {
        k: np.zeros((0, 3) if k in ("coord", "vel") else 0) for k in ("coord", "vel", "pres", "dens", "mass")}
    walls = None
    if getattr(case, "dbsph", None) is not None:
        from .dbsph import tank_mesh, wall_elements
        mesh = getattr(case, "mesh", None)
        if mesh is None:
            mesh = tank_mesh(case.tank, case.dx)
        walls = wall_elements(*mesh, g, case.dx, case.rho0, case.dbsph)
    return CaseTables(case, s, g, f, cnt, st, vec, kt, ktd, bit, parts,
                      find_convex_edges(f, case.vertices()), walls)
