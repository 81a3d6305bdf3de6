"""Host-side set-up of a 2D (x-z plane) dam-break solve: geometry, grid,
initial particles and the SA-SPH boundary-side tables.

The 2D form of the reference is its -DSPACE_2D build.  Particles carry three
coordinates with y = 0; the grid has one cell row in y (CreaGrid.f90:69-71);
frontiers are polylines of "sides" instead of faces.  Every function restates
one reference unit (cited per function), so a case built here is the case the
reference builds from the equivalent v10 deck (``oracle/decks.py`` writes that
deck from the same :class:`DamBreak2D`).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .case import BT_FIXE, BT_PERI, BT_TAPI, _BT_NAMES, DamBreak3D, Grid, fortran_nint, make_grid

PIGRECO = 3.14159265358979
MAXCLOSEBOUNDSIDES = 4   # Static_allocation_module.f90:38


@dataclass
class DamBreak2D:
    """A rectangular tank (open top, SA-SPH "fixed" walls) in the x-z plane
    with a water column in the corner x=z=0, hydrostatic at rest.  Units SI."""
    tank: tuple = (3.22, 1.8)        # width (x), height (z)
    column: tuple = (1.0, 0.5)
    dx: float = 0.01
    trunc: float = 1.3
    rho0: float = 1000.0
    bulk: float | None = None        # K; None -> rho0*(10*sqrt(g*H))^2
    alfa_mon: float = 0.02
    mu: float = 1.0e-3
    grav: tuple = (0.0, 0.0, -9.81)
    p_ref: float = 0.0
    cfl: float = 0.05
    vsc_coeff: float = 0.05
    pesodt: float = 0.0
    teta_p: float = 0.2
    teta_v: float = 0.2
    coefnmaxpartj: float = 2.0
    slip_coeff: float = 0.0
    slip_mode: int = 1            # 1: slip coefficient from input; 2: SA-SPH wall function
    roughness: tuple = (0.0, 0.0)  # slip_coefficient_mode=2 inputs (eps_sigma_z, H_sgr)
    density_thresholds: int = 0      # 1: Leapfrog_continuity's 0.9/1.1 rho0 clamp
    time_split: int = 1           # 1: leapfrog; 0: explicit RK (RKscheme 1, Euler)
    outlet: bool = False          # the right wall (x = W) an "open" side of zone 3
    name: str = "dam_break_2D"
    control_lines: tuple = ()
    dim: int = 2

    def __post_init__(self):
        if self.bulk is None:
            g = math.sqrt(sum(c * c for c in self.grav))
            c0 = 10.0 * math.sqrt(g * self.column[1])
            self.bulk = self.rho0 * c0 * c0

    @property
    def h(self):
        return self.dx * self.trunc

    frontier_slip_input = property(DamBreak3D.frontier_slip_input.fget)

    def scalars(self):
        """ReadInput.f90:198-205 and Gest_Input.f90:266-289 (2D branches)."""
        h = self.h
        s = dict(h=h, doubleh=2.0 * h, squareh=h * h, square_doubleh=(2.0 * h) * (2.0 * h),
                 eta=1.0e-3 * h, eta2=1.0e-2 * (h * h))
        s["coefke"] = 0.682093 / (h * h)
        s["coefkacl"] = 0.099472 / (h * h)
        q = h * 4.0 / self.dx
        s["nmaxpartj"] = int(self.coefnmaxpartj * (q * q)) + 1
        s["celerita"] = math.sqrt(self.bulk / self.rho0)
        s["kin_visc"] = self.mu / self.rho0
        s["pvolume"] = self.dx * self.dx
        return s

    def vertices(self):
        """Vertex table (1-based ids), (x, z): 1-4 the tank walls from the top
        of the left wall down, along the bottom and up the right wall; 5-8 the
        column polygon."""
        W, H = self.tank
        cw, ch = self.column
        return np.array([(0.0, H), (0.0, 0.0), (W, 0.0), (W, H),
                         (0.0, ch), (0.0, 0.0), (cw, 0.0), (cw, ch)], dtype=np.float64)

    def lines(self):
        """LINES section: line id (= zone id) -> vertex ids.  With an outlet
        the right wall (vertices 3-4) is line 3, an open side whose particles
        leave the domain when they cross it (CancelOutgoneParticles_2D.f90)."""
        if self.outlet:
            return {1: [1, 2, 3], 2: [5, 6, 7, 8, 5], 3: [3, 4]}
        return {1: [1, 2, 3, 4], 2: [5, 6, 7, 8, 5]}

    def zone_types(self):
        z = {1: "fixe", 2: "peri"}
        if self.outlet:
            z[3] = "open"
        return z


@dataclass
class Sides:
    """Boundary sides in BoundarySide order (DefineBoundarySideGeometry2D.f90:
    the lines of every zone, peri zones included, one side per vertex pair)."""
    stretch: np.ndarray          # zone (Tratto) id, 1-based
    btype: np.ndarray            # BT_* code of the zone
    v1: np.ndarray               # (ns, 3) first vertex (x, 0, z)
    v2: np.ndarray
    length: np.ndarray
    angle: np.ndarray
    previous: np.ndarray         # previous_side, 1-based, 0 = none
    T: np.ndarray                # (ns, 3, 3)
    RN: np.ndarray               # (ns, 3, 3) BoundaryMassForceMatrix2D
    velocity: np.ndarray         # (ns, 3) BoundarySide%velocity
    tratto_velocity: np.ndarray  # (ns, 3) Tratto%velocity
    slip_mode: np.ndarray
    slip_coeff: np.ndarray
    laminar_check: np.ndarray


def make_sides(case: DamBreak2D) -> Sides:
    """DefineBoundarySideGeometry2D.f90:60-176 with BoundaryMassForceMatrix2D
    .f90:57-61 and DefineBoundarySideRelativeAngles2D.f90:57-89."""
    vert2 = case.vertices()
    vert = np.stack([vert2[:, 0], np.zeros(len(vert2)), vert2[:, 1]], axis=1)
    ztype = case.zone_types()
    recs = []
    for nt, vids in sorted(case.lines().items()):
        tipo = ztype[nt]
        for a, b in zip(vids[:-1], vids[1:]):
            p1, p2 = vert[a - 1], vert[b - 1]
            Dx = p2[0] - p1[0]
            Dz = p2[2] - p1[2]
            L = math.sqrt(Dx * Dx + Dz * Dz)
            sx, sz = Dx / L, Dz / L
            T = np.zeros((3, 3))
            T[0, 0], T[2, 0], T[0, 2], T[2, 2] = sx, sz, -sz, sx
            fis = fin = 1.0 if tipo in ("fixe", "tapi", "leve", "crit") else 0.0
            if tipo == "peri":
                fis = fin = 0.0
            RN = np.zeros((3, 3))
            RN[0, 0] = fis * T[0, 0] * T[0, 0] + fin * T[2, 0] * T[2, 0]
            RN[0, 2] = (fis - fin) * T[0, 0] * T[2, 0]
            RN[2, 0] = RN[0, 2]
            RN[2, 2] = fis * T[2, 0] * T[2, 0] + fin * T[0, 0] * T[0, 0]
            recs.append(dict(stretch=nt, btype=_BT_NAMES[tipo], v1=p1, v2=p2, v1id=a, v2id=b,
                             length=L, T=T, RN=RN))
    ns = len(recs)
    angle = np.zeros(ns)
    prev = np.zeros(ns, np.int32)
    for i, r in enumerate(recs):
        if r["btype"] == BT_PERI:
            continue
        ksi = 0
        for j, q in enumerate(recs):
            if q["btype"] == BT_PERI:
                continue
            if q["v2id"] == r["v1id"]:
                ksi = j + 1
                break
        if ksi > 0:
            Tm, PT = r["T"], recs[ksi - 1]["T"]
            sinangle = PT[0, 0] * Tm[2, 0] - PT[2, 0] * Tm[0, 0]
            cosangle = PT[0, 0] * Tm[0, 0] + PT[2, 0] * Tm[2, 0]
            angle[i] = math.atan2(sinangle, cosangle)
        prev[i] = ksi
    return Sides(
        stretch=np.array([r["stretch"] for r in recs], np.int32),
        btype=np.array([r["btype"] for r in recs], np.int32),
        v1=np.array([r["v1"] for r in recs]).reshape(ns, 3),
        v2=np.array([r["v2"] for r in recs]).reshape(ns, 3),
        length=np.array([r["length"] for r in recs]),
        angle=angle, previous=prev,
        T=np.array([r["T"] for r in recs]).reshape(ns, 3, 3),
        RN=np.array([r["RN"] for r in recs]).reshape(ns, 3, 3),
        velocity=np.zeros((ns, 3)), tratto_velocity=np.zeros((ns, 3)),
        slip_mode=np.full(ns, case.slip_mode, np.int32),
        slip_coeff=np.full(ns, case.frontier_slip_input),
        laminar_check=np.zeros(ns, np.int32))


def domain_box(case: DamBreak2D):
    """Domain%coord over all vertices, y = 0."""
    v = case.vertices()
    lo = np.array([v[:, 0].min(), 0.0, v[:, 1].min()])
    hi = np.array([v[:, 0].max(), 0.0, v[:, 1].max()])
    return lo, hi


def make_grid2d(case: DamBreak2D, doubleh):
    """CreaGrid.f90:58-74: dcd from NINT(extent/2h) in every direction, then
    ncd(2) forced to 1 (dcd(2) keeps its value)."""
    lo, hi = domain_box(case)
    g = make_grid(lo, hi, case.dx, doubleh)
    ncd = g.ncd.copy()
    ncd[1] = 1
    return Grid(g.extr_lo, g.extr_hi, ncd, g.dcd)


def initial_particles(case: DamBreak2D):
    """GeneratePart.f90:196-221 + SetParticles.f90 (2D: Npps(2) = 0 -> 1),
    IsParticleInternal2D over the column polygon, SetParticleParameters
    .f90:65-67 (y = 0), SubCalcPreIdro.f90 "qp" branch.  Order x, then z."""
    s = case.scalars()
    v = case.vertices()
    xmin_all = v.min(0)
    col = v[4:8]
    cmin, cmax = col.min(0), col.max(0)
    ax = []
    for i in range(2):
        nbefore = fortran_nint((cmin[i] - xmin_all[i]) / case.dx)
        xr = xmin_all[i] + nbefore * case.dx
        npps = fortran_nint((cmax[i] - xr) / case.dx)
        pts, p = [], xr - case.dx * 0.5
        for _ in range(npps):
            p = p + case.dx
            pts.append(p)
        ax.append(np.array(pts))
    X, Z = np.meshgrid(ax[0], ax[1], indexing="ij")
    xz = np.stack([X.ravel(), Z.ravel()], axis=1)
    inside = np.all((xz > cmin) & (xz < cmax), axis=1)
    xz = xz[inside]
    n = len(xz)
    coord = np.zeros((n, 3))
    coord[:, 0] = xz[:, 0]
    coord[:, 2] = xz[:, 1]
    g = case.grav
    gmod = math.sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2])
    coshor = -g[2] / gmod
    senhor = g[0] / gmod
    zq = float(case.column[1])
    affond = (zq - coord[:, 2]) * coshor + coord[:, 0] * senhor
    pres = (affond * case.rho0 * gmod) + case.p_ref
    dens = case.rho0 * (1.0 + (pres - case.p_ref) / case.bulk)
    mass = np.full(n, s["pvolume"] * case.rho0)
    return dict(coord=coord, vel=np.zeros((n, 3)), dens=dens, pres=pres, mass=mass)


# BoundIntegralTab2D (SA_SPH_module.f90:30-54): the reference's tabulated 2D
# wall integrals, 22 rows x (delta_i, Int_Wd0s0, Int_IWd1s0, Int_Wd1s0,
# Int_Wd3s0, Int_Wd1s2) at 7 significant digits.  The values are data of the
# reference (not recomputable to the printed digits) and are kept as given.
# Its DATA statement writes them as default-real (E) literals into a double
# precision array, so each entry is the single-precision value widened.
BIT2D = np.array([
    [-9.900000e-02, 6.916001e-01, 5.671769e-01, 1.614238e+00, 7.329450e-01, 8.813822e-01],
    [1.000000e-03, 6.820633e-01, 4.993001e-01, 1.426603e+00, 7.142813e-01, 7.123606e-01],
    [1.010000e-01, 6.725266e-01, 4.314232e-01, 1.238968e+00, 6.956176e-01, 5.433391e-01],
    [2.010000e-01, 6.452481e-01, 3.653947e-01, 1.057834e+00, 6.492812e-01, 4.085521e-01],
    [3.010000e-01, 6.025034e-01, 3.028878e-01, 8.871524e-01, 5.851007e-01, 3.020341e-01],
    [4.010000e-01, 5.473293e-01, 2.453079e-01, 7.296175e-01, 5.106232e-01, 2.189841e-01],
    [5.010000e-01, 4.832434e-01, 1.937172e-01, 5.872117e-01, 4.319031e-01, 1.553082e-01],
    [6.010000e-01, 4.139976e-01, 1.488267e-01, 4.612826e-01, 3.538107e-01, 1.074685e-01],
    [7.010000e-01, 3.434138e-01, 1.109601e-01, 3.526284e-01, 2.802530e-01, 7.237218e-02],
    [8.010000e-01, 2.751568e-01, 8.006434e-02, 2.615428e-01, 2.142205e-01, 4.732315e-02],
    [9.010000e-01, 2.125249e-01, 5.573915e-02, 1.877770e-01, 1.577964e-01, 2.998003e-02],
    [1.001000e+00, 1.581202e-01, 3.728314e-02, 1.304301e-01, 1.120873e-01, 1.834320e-02],
    [1.101000e+00, 1.131765e-01, 2.379696e-02, 8.756287e-02, 7.679627e-02, 1.076657e-02],
    [1.201000e+00, 7.737532e-02, 1.434262e-02, 5.638309e-02, 5.038764e-02, 5.995379e-03],
    [1.301000e+00, 4.995510e-02, 8.042085e-03, 3.441001e-02, 3.129118e-02, 3.118804e-03],
    [1.401000e+00, 2.994721e-02, 4.104289e-03, 1.956203e-02, 1.807987e-02, 1.482120e-03],
    [1.501000e+00, 1.623613e-02, 1.842642e-03, 1.008625e-02, 9.464607e-03, 6.216584e-04],
    [1.601000e+00, 7.615461e-03, 6.873181e-04, 4.510624e-03, 4.293336e-03, 2.172841e-04],
    [1.701000e+00, 2.842820e-03, 1.913104e-04, 1.609094e-03, 1.552266e-03, 5.682598e-05],
    [1.801000e+00, 6.998383e-04, 3.120105e-05, 3.793334e-04, 3.706037e-04, 8.729655e-06],
    [1.901000e+00, 6.214235e-05, 1.372448e-06, 3.231651e-05, 3.195398e-05, 3.625237e-07],
    [2.001000e+00, 0.0, 0.0, 0.0, 0.0, 0.0],
], dtype=np.float32).astype(np.float64)
DELTAX_BIT = 0.1


@dataclass
class CaseTables2D:
    case: DamBreak2D
    scalars: dict
    grid: Grid
    sides: Sides
    bit2d: np.ndarray
    particles: dict = field(default_factory=dict)
    dim: int = 2


def build(case: DamBreak2D, particles=True) -> CaseTables2D:
    s = case.scalars()
    g = make_grid2d(case, s["doubleh"])
    sd = make_sides(case)
    parts = initial_particles(case) if particles else {
        k: np.zeros((0, 3) if k in ("coord", "vel") else 0) for k in ("coord", "vel", "pres", "dens", "mass")}
    return CaseTables2D(case, s, g, sd, BIT2D.copy(), parts)


__all__ = ["DamBreak2D", "Sides", "CaseTables2D", "build", "make_sides", "make_grid2d",
           "initial_particles", "BIT2D", "DELTAX_BIT", "MAXCLOSEBOUNDSIDES", "BT_FIXE", "BT_TAPI"]
