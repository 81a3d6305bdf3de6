"""SPHERA v10 input decks: a writer for the 3D and 2D dam-break cases and a
parser for the subset of the format those cases and the reference's v10 3D
tutorial use.

The format is the reference's (src/Pre_processing/reading_inp/ReadInput.f90:
102-194 and one reader per section); the reference concatenates
./input/*/*.inp into <case>.inp (src/Main_algorithm/sphera.f90:222) and reads
that.  Integers are written without a decimal point (flang's list-directed
read rejects "0." for an integer); the parser accepts both.

* :func:`deck_text` / :func:`deck_text_2d` write the deck of a
  :class:`~sphera_amd.case.DamBreak3D` / ``DamBreak2D``; the sections and the
  values that do not describe the case (DB-SPH, bed load, monitors) follow the
  v10 3D tutorial deck, with outputs throttled.
* :func:`parse_deck` reads a 3D deck with SA-SPH walls (``semi``): one
  ``fixed`` frontier zone of quadrilateral faces, one ``perimeter`` fluid zone
  with a hydrostatic (``qp``) initial pressure, one liquid medium, and returns
  a :class:`DeckCase` that :func:`sphera_amd.case.build` turns into the same
  set-up tables the reference builds from that deck.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

def _f(x):
    return repr(float(x))


def _fixed_zone(case):
    """Lines of a ``fixed`` frontier zone after its type (ReadInputBoundaries.f90:
    183-208): slip_coefficient_mode and laminar_no_slip_check, then the slip
    coefficient (mode 1) or eps_sigma_z and H_sgr (mode 2)."""
    mode = int(getattr(case, "slip_mode", 1))
    out = [f"{mode} .false."]
    if mode == 1:
        out.append(_f(case.slip_coeff))
    elif mode == 2:
        out.append(f"{_f(case.roughness[0])} {_f(case.roughness[1])}")
    return out


def deck_text(case, itmax=1000000000, tmax=1000.0, results_step=None):
    s = case.scalars()
    v = case.vertices()
    db = getattr(case, "dbsph", None)
    lines = ["10.0.0", "##### TITLE #####", case.name, "##### END TITLE #####",
             "##### DOMAIN #####", "bsph" if db is not None else "semi",
             f"{_f(case.dx)} {_f(case.trunc)}", ".false.",
             "##### END DOMAIN #####", "##### VERTICES #####"]
    for i, p in enumerate(v, start=1):
        lines.append(f"{i} {_f(p[0])} {_f(p[1])} {_f(p[2])}")
    lines += ["##### END VERTICES #####", "##### FACES #####"]
    for i, r in enumerate(case.faces(), start=1):
        lines.append(f"{i} {r[0]} {r[1]} {r[2]} {r[3]} -1 -1 {r[4]}")
    lines += ["##### END FACES #####", "##### BOUNDARIES #####",
              "Boundary", "1", "fixed", *_fixed_zone(case), "111111", "1 8",
              "Reservoir", "2", "perimeter", "1", "uniform 009EA8",
              "std 0. 0. 0. 0. 0. 0. 0.", f"qp {_f(case.column[2])} 0.", "1 0 .false.",
              *(["Outlet", "3", "open", "FF0000"] if getattr(case, "outlet", False) else []),
              "##### END BOUNDARIES #####",
              "##### DBSPH #####",
              *(db.deck_lines() if db is not None else
                ["1. .false. 0.5 0 .false.", ".false. .true.", "0 0", "0 .false.", "0 0 6"]),
              "##### END DBSPH #####",
              "##### BED LOAD TRANSPORT #####", "0 1", "1 0. 0.",
              "##### end BED LOAD TRANSPORT #####",
              "##### medium #####", "liquid", "1", f"{_f(case.rho0)} {_f(case.bulk)}",
              f"{_f(case.alfa_mon)} 0.", _f(case.mu), "##### end medium #####",
              "##### RUN PARAMETERS #####", f"{_f(tmax)} {int(itmax)}",
              f"{_f(case.cfl)} {_f(case.vsc_coeff)} {int(getattr(case, 'time_split', 1))} 1 "
              f"{_f(case.pesodt)} .false.",
              f"{_f(case.teta_p)} {_f(case.teta_v)} A",
              f"1.1 {_f(case.coefnmaxpartj)} 0",
              ".false. .false. .false. .false. .false. 0.614 .false. 50. .false.",
              f"1000000 {int(case.maxcloseboundfaces)} 100 100000000",
              str(int(getattr(case, "density_thresholds", 0))),
              "##### end RUN PARAMETERS #####",
              "##### general physical properties #####",
              f"{_f(case.grav[0])} {_f(case.grav[1])} {_f(case.grav[2])}", _f(case.p_ref),
              "##### end general physical properties #####",
              "##### restart #####", "##### end restart #####",
              "##### output regulation #####",
              f"results step {int(results_step)}" if results_step else "results time 100000.",
              "restart time 100000.", "print partial 100000000", "control time 100000.",
              "level time 100000. medium 1", "depth dt_out 100000.",
              "##### end output regulation #####",
              "##### draw options #####", "vtkconverter any 100000.",
              "##### end draw options #####",
              "##### control points #####", "##### end control points #####",
              "##### control lines #####"]
    for label, first, second, ndiv in case.control_lines:
        lines += [str(label)[:8], " ".join(_f(v) for v in first), " ".join(_f(v) for v in second),
                  str(int(ndiv))]
    lines += ["##### end control lines #####",
              "##### section flow rate #####", "0 999. 1", "##### end section flow rate #####",
              "##### substations #####", "0 999.", "##### end substations #####"]
    del s
    return "\n".join(lines) + "\n"


def deck_text_2d(case, itmax=1000000000, tmax=1000.0, results_step=None):
    """v10 deck of a DamBreak2D for the reference's 2D build: vertices (x, z),
    a LINES section instead of FACES, gravity and monitor points with two
    components, and none of the 3D-only lines (the CLC flag of
    ReadInputDomain.f90:109-113, MAXCLOSEBOUNDFACES of
    ReadInputRunParameters.f90:70-76)."""
    v = case.vertices()
    lines = ["10.0.0", "##### TITLE #####", case.name, "##### END TITLE #####",
             "##### DOMAIN #####", "semi", f"{_f(case.dx)} {_f(case.trunc)}",
             "##### END DOMAIN #####", "##### VERTICES #####"]
    for i, p in enumerate(v, start=1):
        lines.append(f"{i} {_f(p[0])} {_f(p[1])}")
    lines += ["##### END VERTICES #####", "##### LINES #####"]
    for lid, vids in sorted(case.lines().items()):
        lines.append(f"{lid} " + " ".join(str(i) for i in vids))
    lines += ["##### END LINES #####", "##### BOUNDARIES #####",
              "Boundary", "1", "fixed", *_fixed_zone(case), "111111",
              "Reservoir", "2", "perimeter", "1", "uniform 009EA8",
              "std 0. 0. 0. 0. 0. 0.", f"qp {_f(case.column[1])}", "1 0 .false.",
              *(["Outlet", "3", "open", "FF0000"] if getattr(case, "outlet", False) else []),
              "##### END BOUNDARIES #####",
              "##### DBSPH #####", "1. .false. 0.5 0 .false.", ".false. .true.", "0 0",
              "0 .false.", "0 0 6", "##### END DBSPH #####",
              "##### BED LOAD TRANSPORT #####", "0 1", "1 0. 0.",
              "##### end BED LOAD TRANSPORT #####",
              "##### medium #####", "liquid", "1", f"{_f(case.rho0)} {_f(case.bulk)}",
              f"{_f(case.alfa_mon)} 0.", _f(case.mu), "##### end medium #####",
              "##### RUN PARAMETERS #####", f"{_f(tmax)} {int(itmax)}",
              f"{_f(case.cfl)} {_f(case.vsc_coeff)} {int(getattr(case, 'time_split', 1))} 1 "
              f"{_f(case.pesodt)} .false.",
              f"{_f(case.teta_p)} {_f(case.teta_v)} A",
              f"1.1 {_f(case.coefnmaxpartj)} 0",
              ".false. .false. .false. .false. .false. 0.614 .false. 50. .false.",
              str(int(getattr(case, "density_thresholds", 0))),
              "##### end RUN PARAMETERS #####",
              "##### general physical properties #####",
              f"{_f(case.grav[0])} {_f(case.grav[2])}", _f(case.p_ref),
              "##### end general physical properties #####",
              "##### restart #####", "##### end restart #####",
              "##### output regulation #####",
              f"results step {int(results_step)}" if results_step else "results time 100000.",
              "restart time 100000.", "print partial 100000000", "control time 100000.",
              "level time 100000. medium 1", "depth dt_out 100000.",
              "##### end output regulation #####",
              "##### draw options #####", "vtkconverter any 100000.",
              "##### end draw options #####",
              "##### control points #####", "##### end control points #####",
              "##### control lines #####"]
    for label, first, second, ndiv in case.control_lines:
        lines += [str(label)[:8], " ".join(_f(v) for v in first), " ".join(_f(v) for v in second),
                  str(int(ndiv))]
    lines += ["##### end control lines #####"]
    return "\n".join(lines) + "\n"



# ---------------------------------------------------------------------------
# Parser (ReadInput.f90:108-194 and the section readers cited below)
# ---------------------------------------------------------------------------
def _num(tok):
    """A Fortran number token (d/D exponents allowed)."""
    return float(tok.replace("d", "e").replace("D", "e"))


def _int(tok):
    return int(round(_num(tok)))


def _logical(tok):
    t = tok.strip(".").lower()
    if t in ("true", "t"):
        return True
    if t in ("false", "f"):
        return False
    raise ValueError(f"not a logical: {tok}")


def _sections(text):
    """Title and the lines of every ``##### NAME #####`` ... ``##### END NAME
    #####`` section, with comments (from '!') and blank lines removed
    (ReadRiga)."""
    lines = []
    for raw in text.splitlines():
        ln = raw.split("!", 1)[0].strip()
        if ln:
            lines.append(ln)
    if not lines or lines[0] != "10.0.0":
        raise ValueError("a v10 deck starts with the line 10.0.0 (ReadInput.f90:102-106)")
    out, cur, name = {}, None, None
    for ln in lines[1:]:
        if ln.startswith("#####"):
            tag = ln.strip("#").strip().lower()
            if tag.startswith("end "):
                if name is not None:
                    out[name] = cur
                name, cur = None, None
            else:
                name, cur = tag, []
            continue
        if cur is not None:
            cur.append(ln)
    return out


@dataclass
class DeckCase:
    """A case read from a v10 deck, with the interface of DamBreak3D that
    :func:`sphera_amd.case.build` and the C-ABI parameters use."""
    name: str
    dx: float
    trunc: float
    rho0: float
    bulk: float
    alfa_mon: float
    mu: float
    grav: tuple
    p_ref: float
    cfl: float
    vsc_coeff: float
    pesodt: float
    teta_p: float
    teta_v: float
    coefnmaxpartj: float
    maxcloseboundfaces: int
    slip_coeff: float
    density_thresholds: int
    time_split: int
    pool: float
    slip_mode: int
    roughness: tuple
    verts: np.ndarray = field(repr=False)
    face_rows: list = field(repr=False)
    zones: dict = field(repr=False)
    fluid_zone: int = 2
    control_lines: tuple = ()
    dim: int = 3
    # DB-SPH walls (Domain%tipo = "bsph"): sphera_amd.dbsph.DBSPHParams and
    # the surface mesh (vertices, faces) of input/07_DBSPH; None for SA-SPH
    dbsph: object = None
    mesh: object = field(default=None, repr=False)

    @property
    def outlet(self):
        """True when an open zone holds frontier faces."""
        return "open" in self.zones.values()

    @property
    def h(self):
        return self.dx * self.trunc

    def scalars(self):
        """ReadInput.f90:198-205 and Gest_Input.f90:266-289 (3D)."""
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

    @property
    def frontier_slip_input(self):
        """BC_shear_stress_input of the frontier zone (mode 2: 2 eps/5 + H/5)."""
        if self.slip_mode == 2:
            return 2.0 * self.roughness[0] / 5.0 + self.roughness[1] / 5.0
        return self.slip_coeff if self.slip_mode == 1 else 0.0

    def vertices(self):
        return self.verts

    def faces(self):
        return self.face_rows

    def zone_types(self):
        return dict(self.zones)

    def fluid_zone_box(self):
        ids = sorted({n for r in self.face_rows if r[4] == self.fluid_zone for n in r[:4]})
        v = self.verts[np.array(ids) - 1]
        return v.min(0), v.max(0)

    def free_surface_height(self):
        return float(self.pool)


def parse_deck(text, name=None):
    """Reads the deck subset described in the module docstring; raises
    ValueError for anything outside it."""
    sec = _sections(text)
    for req in ("domain", "vertices", "faces", "boundaries", "medium", "run parameters",
                "general physical properties"):
        if req not in sec:
            raise ValueError(f"section {req!r} missing")
    # DOMAIN (ReadInputDomain.f90): tipo; dx trunc; the CLC flag (3D)
    dom = sec["domain"]
    tipo = dom[0].split()[-1].lower()
    if tipo not in ("semi", "bsph"):
        raise ValueError(f"Domain%tipo {tipo!r}: only SA-SPH (semi) and DB-SPH (bsph) walls are read")
    dx, trunc = (_num(t) for t in dom[1].split()[:2])
    # VERTICES (ReadInputVertices.f90): id x y z
    vt = {}
    for ln in sec["vertices"]:
        t = ln.split()
        vt[_int(t[0])] = [_num(v) for v in t[1:4]]
    nv = max(vt)
    if sorted(vt) != list(range(1, nv + 1)):
        raise ValueError("vertices must be numbered 1..n")
    verts = np.array([vt[k] for k in range(1, nv + 1)], dtype=np.float64)
    # FACES (ReadInputFaces.f90): id, six node slots (-1 = none), zone
    rows = []
    for ln in sec["faces"]:
        t = [_int(v) for v in ln.split()]
        nodes = [n for n in t[1:7] if n > 0]
        if len(nodes) != 4:
            raise ValueError("only quadrilateral faces are read")
        rows.append(tuple(nodes) + (t[7],))
    # BOUNDARIES (ReadInputBoundaries.f90): name, zone, type, then per type
    b = sec["boundaries"]
    zones, k = {}, 0
    slip, mode, rough, pool, fluid_zone = 0.0, 1, (0.0, 0.0), None, None
    while k < len(b):
        zone = _int(b[k + 1])
        kind = b[k + 2].lower()
        if kind == "fixed":
            mode = _int(b[k + 3].split()[0])
            if mode not in (0, 1, 2):
                raise ValueError("slip_coefficient_mode must be 0, 1 or 2")
            nval = 0
            if mode == 1:
                slip, nval = _num(b[k + 4].split()[0]), 1
            elif mode == 2:
                rough, nval = tuple(_num(v) for v in b[k + 4].split()[:2]), 1
            zones[zone] = "fixe"
            k += 6 + nval
        elif kind == "perimeter":
            q = b[k + 6].split()
            if q[0].lower() != "qp":
                raise ValueError("only hydrostatic (qp) perimeter zones are read")
            pool = _num(q[1])
            zones[zone] = "peri"
            fluid_zone = zone
            k += 8
        elif kind == "open":  # name, zone, type, colour (ReadInputBoundaries.f90:172-182)
            zones[zone] = "open"
            k += 4
        else:
            raise ValueError(f"boundary type {kind!r} is not read")
    if fluid_zone is None:
        raise ValueError("no perimeter zone")
    # medium (ReadInputMedium.f90): liquid; index; den0 eps; alfaMon ...; mu
    med = sec["medium"]
    if med[0].lower() != "liquid":
        raise ValueError("only a liquid medium is read")
    rho0, bulk = (_num(t) for t in med[2].split()[:2])
    alfa_mon = _num(med[3].split()[0])
    mu = _num(med[4].split()[0])
    # RUN PARAMETERS (ReadInputRunParameters.f90:81-126)
    rp = sec["run parameters"]
    t = rp[1].split()
    cfl, vsc, time_split, rk, pesodt = _num(t[0]), _num(t[1]), _int(t[2]), _int(t[3]), _num(t[4])
    if time_split == 0 and rk != 1:
        raise ValueError("explicit RK with RKscheme > 1 is not read")
    tetap, tetav = (_num(v) for v in rp[2].split()[:2])
    coefj = _num(rp[3].split()[1])
    opts = rp[4].split()
    if any(_logical(v) for v in opts[:3]):
        raise ValueError("ALE3 / C1_BE / C1_divu are not read")
    maxclose = _int(rp[5].split()[1])
    dthr = _int(rp[6].split()[0])
    # general physical properties (ReadInputGeneralPhysical.f90)
    gp = sec["general physical properties"]
    grav = tuple(_num(v) for v in gp[0].split()[:3])
    p_ref = _num(gp[1].split()[0])
    # control lines (ReadInputControlLines.f90): label, first, second, points
    lines = []
    cl = sec.get("control lines", [])
    for i in range(0, len(cl) - 3, 4):
        lines.append((cl[i].split()[0], tuple(_num(v) for v in cl[i + 1].split()[:3]),
                      tuple(_num(v) for v in cl[i + 2].split()[:3]), _int(cl[i + 3].split()[0])))
    title = sec.get("title", ["case"])
    dbsph = None
    if tipo == "bsph":  # ReadDBSPH.f90:35-99, the subset sphera_amd.dbsph runs
        from .dbsph import DBSPHParams
        db = sec.get("dbsph", [])
        if len(db) < 6:
            raise ValueError("DBSPH section missing or short")
        t = db[0].split()
        if _int(t[3]) != 0:
            raise ValueError("DB-SPH slip_ID > 0 (the VSL wall model) is not read")
        n_mon = [_int(v) for v in db[2].split()[:2]]
        kl = db[4].split()
        n_kin = _int(kl[0])
        if len(db) < 6 + n_kin:
            raise ValueError("DBSPH section short")
        kin = tuple(tuple(_num(v) for v in db[5 + j].split()[:7]) for j in range(n_kin))
        n_io = [_int(v) for v in db[5 + n_kin].split()[:2]]
        if any(n_mon) or any(n_io) or _int(db[3].split()[0]) != 1:
            raise ValueError("DB-SPH monitors, inlets/outlets and several meshes are not read")
        neg, fs = (_logical(v) for v in db[1].split()[:2])
        dbsph = DBSPHParams(dx_dxw=_num(t[0]), k_w=_num(t[2]), gamma_limiter=_logical(t[4]),
                            negative_wall_p_allowed=neg, fs_allowed=fs, muscl_boundary=_logical(t[1]),
                            kinematics=kin,
                            rotation_centre=tuple(_num(v) for v in kl[1:4]) if n_kin else (0.0, 0.0, 0.0))
        if time_split != 0:
            raise ValueError("DB-SPH walls run with explicit RK1 (time_split = 0)")
    return DeckCase(dbsph=dbsph, name=name or title[0].split()[0], dx=dx, trunc=trunc, rho0=rho0, bulk=bulk,
                    alfa_mon=alfa_mon, mu=mu, grav=grav, p_ref=p_ref, cfl=cfl, vsc_coeff=vsc,
                    pesodt=pesodt, teta_p=tetap, teta_v=tetav, coefnmaxpartj=coefj,
                    maxcloseboundfaces=maxclose, slip_coeff=slip, density_thresholds=dthr,
                    time_split=time_split, pool=pool, slip_mode=mode, roughness=rough, verts=verts, face_rows=rows, zones=zones,
                    fluid_zone=fluid_zone, control_lines=tuple(lines))


def read_case_dir(root, name=None):
    """A case directory as the reference starts from it: the deck of
    input/00_deck/*.inp, and for DB-SPH walls the surface meshes listed in
    input/07_DBSPH/surface_mesh_list.txt (one mesh file is read)."""
    from pathlib import Path
    root = Path(root)
    decks = sorted((root / "input").glob("*/*.inp"))
    if not decks:
        raise ValueError(f"no deck under {root / 'input'}")
    case = parse_deck("".join(p.read_text() for p in decks), name=name)
    if case.dbsph is not None:
        from .dbsph import read_ply
        lst = (root / "input" / "07_DBSPH" / "surface_mesh_list.txt").read_text().split()
        case.mesh = read_ply(root / "input" / "07_DBSPH" / lst[1])
    return case
