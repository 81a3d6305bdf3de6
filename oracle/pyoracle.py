"""ctypes binding of the CPU FP64 oracle (oracle/sph_oracle.cpp).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_build" / "libsph_oracle.so"
_lib = None

PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)


class OrParams(C.Structure):
    _fields_ = [("n", C.c_int32), ("nmaxpartj", C.c_int32), ("maxclose", C.c_int32),
                ("ncd", C.c_int32 * 3), ("extr_lo", C.c_double * 3), ("dcd", C.c_double * 3),
                ("dx", C.c_double), ("h", C.c_double), ("coefke", C.c_double),
                ("coefkacl", C.c_double), ("eta", C.c_double), ("eta2", C.c_double),
                ("grav", C.c_double * 3), ("prif", C.c_double), ("den0", C.c_double),
                ("eps", C.c_double), ("celerita", C.c_double), ("alfamon", C.c_double),
                ("kin_visc", C.c_double), ("mumx", C.c_double), ("cfl", C.c_double),
                ("vsc_coeff", C.c_double), ("pesodt", C.c_double), ("tetap", C.c_double),
                ("tetav", C.c_double), ("semi", C.c_int32), ("density_thresholds", C.c_int32),
                ("time_split", C.c_int32)]


class OrFaces(C.Structure):
    _fields_ = [("nf", C.c_int32), ("T", PD), ("LX", PD), ("GX", PD), ("area", PD),
                ("velocity", PD), ("tvel", PD), ("slip_coeff", PD), ("nodes", PI),
                ("btype", PI), ("slip_mode", PI), ("lam_check", PI), ("gcbf_count", PI),
                ("gcbf_start", PI), ("gcbf_vec", PI), ("kerneltab", PD), ("ktdelta", C.c_double),
                ("bitab", PD), ("nedges", C.c_int32), ("edge_face", PI), ("edge_gx", PD),
                ("edge_comp", PD), ("edge_len", PD)]


def _edge_arrays(tb, k):
    """The convex-edge arrays of the case tables (at least one element each)
    into k; returns the number of edges."""
    E = getattr(tb, "edges", None)
    ne = len(E) if E is not None else 0
    k["ef"] = np.ascontiguousarray(E.face.reshape(-1) if ne else np.zeros(2), np.int32)
    k["egx"] = np.ascontiguousarray(E.gx.reshape(-1) if ne else np.zeros(3), np.float64)
    k["ecomp"] = np.ascontiguousarray(E.component.reshape(-1) if ne else np.zeros(3), np.float64)
    k["elen"] = np.ascontiguousarray(E.length if ne else np.zeros(1), np.float64)
    return ne


class OrSides2D(C.Structure):
    _fields_ = [("ns", C.c_int32), ("T", PD), ("RN", PD), ("v1", PD), ("length", PD),
                ("angle", PD), ("velocity", PD), ("tvel", PD), ("slip_coeff", PD),
                ("btype", PI), ("previous", PI), ("slip_mode", PI), ("lam_check", PI),
                ("bit2d", PD), ("v2", PD)]


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} missing: make -f oracle/Makefile")
        L = C.CDLL(str(LIB_PATH))
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(OrParams), C.POINTER(OrFaces), PD, PD, PD, PD, PD]
        L.or_create2d.restype = C.c_void_p
        L.or_create2d.argtypes = [C.POINTER(OrParams), C.POINTER(OrSides2D), PD, PD, PD, PD, PD]
        for name in ("or_destroy",):
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = None
        for name in ("or_nag", "or_ordgrid", "or_calcvarlength", "or_boundary_tab", "or_dt",
                     "or_momentum", "or_vsmooth", "or_leapfrog_traj", "or_continuity",
                     "or_leapfrog_cont", "or_eos", "or_psmooth", "or_preloop", "or_overflow",
                     "or_cancel_outgone"):
            getattr(L, name).argtypes = [C.c_void_p]
            getattr(L, name).restype = C.c_int
        L.or_outgone_count.argtypes = [C.c_void_p]
        L.or_outgone_count.restype = C.c_longlong
        L.or_leapfrog_mom.argtypes = [C.c_void_p, C.c_double]
        L.or_step.argtypes = [C.c_void_p, C.c_int]
        L.or_pair_count.argtypes = [C.c_void_p]
        L.or_pair_count.restype = C.c_int64
        L.or_branch_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.or_compare_lists.argtypes = [C.c_void_p, PI, PI, C.c_int32]
        L.or_compare_lists.restype = C.c_int64
        L.or_set_lists_only.argtypes = [C.c_void_p, C.c_int]
        L.or_set_lists_only.restype = None
        L.or_branch_counts.restype = None
        L.or_set_state.argtypes = [C.c_void_p, C.c_int, PD, PD, PD, PD, PD, PD, PI, C.c_double,
                                   C.c_double, C.c_double, C.c_int]
        L.or_get_state.argtypes = [C.c_void_p] + [PD] * 8
        L.or_get_scalars.argtypes = [C.c_void_p, PD]
        L.or_get_lists.argtypes = [C.c_void_p, PI, PI, C.c_int32]
        L.or_get_btab.argtypes = [C.c_void_p, PI, PI, PD, PD, PD]
        L.or_free_surface.argtypes = [C.c_void_p, C.c_int, PI, C.c_int, PD, PD, PD]
        L.or_set_dbsph.argtypes = [C.c_void_p, C.c_int] + [PD] * 6 + [PI]
        L.or_set_dbsph.restype = C.c_int
        L.or_set_dbsph_kinematics.argtypes = [C.c_void_p, C.c_int, PD, PD]
        L.or_set_dbsph_kinematics.restype = C.c_int
        L.or_get_dbsph.argtypes = [C.c_void_p] + [PD] * 6 + [PI, PD, PD, PI]
        L.or_get_dbsph.restype = None
        L.or_get_fw_lists.argtypes = [C.c_void_p, PI, PI, C.c_int32]
        L.or_get_fw_lists.restype = C.c_int
        _lib = L
    return _lib


def _pd(a):
    return a.ctypes.data_as(PD) if a is not None else None


def _pi(a):
    return a.ctypes.data_as(PI) if a is not None else None


def params_from_tables(tb, n, semi=True):
    s, g, case = tb.scalars, tb.grid, tb.case
    p = OrParams()
    p.n = n
    p.nmaxpartj = s["nmaxpartj"]
    p.maxclose = 4 if getattr(tb, "dim", 3) == 2 else case.maxcloseboundfaces
    p.ncd[:] = [int(v) for v in g.ncd]
    p.extr_lo[:] = [float(v) for v in g.extr_lo]
    p.dcd[:] = [float(v) for v in g.dcd]
    p.dx, p.h, p.coefke, p.coefkacl = case.dx, s["h"], s["coefke"], s["coefkacl"]
    p.eta, p.eta2 = s["eta"], s["eta2"]
    p.grav[:] = list(case.grav)
    p.prif = case.p_ref
    p.den0, p.eps, p.celerita = case.rho0, case.bulk, s["celerita"]
    p.alfamon, p.kin_visc, p.mumx = case.alfa_mon, s["kin_visc"], 3.4e38
    p.cfl, p.vsc_coeff, p.pesodt = case.cfl, case.vsc_coeff, case.pesodt
    p.tetap, p.tetav = case.teta_p, case.teta_v
    p.semi = 1 if semi else 0
    p.density_thresholds = int(getattr(case, "density_thresholds", 0))
    p.time_split = int(getattr(case, "time_split", 1))
    return p


class Oracle:
    """One FP64 oracle context over a built case (sphera_amd.case.build)."""

    def __init__(self, tb, coord=None, vel=None, dens=None, pres=None, mass=None, semi=True):
        L = lib()
        P = tb.particles
        self.coord = np.ascontiguousarray(P["coord"] if coord is None else coord, np.float64)
        n = len(self.coord)
        vel = np.ascontiguousarray(P["vel"] if vel is None else vel, np.float64)
        dens = np.ascontiguousarray(P["dens"] if dens is None else dens, np.float64)
        pres = np.ascontiguousarray(P["pres"] if pres is None else pres, np.float64)
        mass = np.ascontiguousarray(P["mass"] if mass is None else mass, np.float64)
        db = getattr(tb.case, "dbsph", None) if getattr(tb, "dim", 3) == 3 else None
        if db is not None:  # DB-SPH walls instead of SA-SPH frontiers
            semi = False
        self.tb = tb
        self.p = params_from_tables(tb, n, semi)
        if getattr(tb, "dim", 3) == 2:
            S = tb.sides
            self._keep = dict(
                T=np.ascontiguousarray(S.T.transpose(0, 2, 1).reshape(-1)),
                RN=np.ascontiguousarray(S.RN.transpose(0, 2, 1).reshape(-1)),
                v1=np.ascontiguousarray(S.v1.reshape(-1)), v2=np.ascontiguousarray(S.v2.reshape(-1)),
                L=np.ascontiguousarray(S.length), ang=np.ascontiguousarray(S.angle), vel=np.ascontiguousarray(S.velocity.reshape(-1)),
                tvel=np.ascontiguousarray(S.tratto_velocity.reshape(-1)),
                slip=np.ascontiguousarray(S.slip_coeff), bt=np.ascontiguousarray(S.btype, np.int32),
                prev=np.ascontiguousarray(S.previous, np.int32),
                mode=np.ascontiguousarray(S.slip_mode, np.int32),
                lam=np.ascontiguousarray(S.laminar_check, np.int32),
                bit=np.ascontiguousarray(tb.bit2d.reshape(-1)))
            k = self._keep
            sd = OrSides2D(len(S.length), _pd(k["T"]), _pd(k["RN"]), _pd(k["v1"]), _pd(k["L"]),
                           _pd(k["ang"]), _pd(k["vel"]), _pd(k["tvel"]), _pd(k["slip"]),
                           _pi(k["bt"]), _pi(k["prev"]), _pi(k["mode"]), _pi(k["lam"]),
                           _pd(k["bit"]), _pd(k["v2"]))
            self.ctx = L.or_create2d(C.byref(self.p), C.byref(sd), _pd(self.coord), _pd(vel),
                                     _pd(dens), _pd(pres), _pd(mass))
            if not self.ctx:
                raise ValueError("or_create2d refused the sides (an open side without v2)")
            return
        F = tb.faces
        self._keep = dict(
            T=np.ascontiguousarray(F.T.transpose(0, 2, 1).reshape(-1)),  # column-major per face
            LX=np.ascontiguousarray(F.LX.reshape(-1)), GX=np.ascontiguousarray(F.GX.reshape(-1)),
            area=np.ascontiguousarray(F.area), vel=np.ascontiguousarray(F.velocity.reshape(-1)),
            tvel=np.ascontiguousarray(F.tratto_velocity.reshape(-1)),
            slip=np.ascontiguousarray(F.slip_coeff), nodes=np.ascontiguousarray(F.nodes, np.int32),
            btype=np.ascontiguousarray(F.btype, np.int32),
            mode=np.ascontiguousarray(F.slip_mode, np.int32),
            lam=np.ascontiguousarray(F.laminar_check, np.int32),
            gc=np.ascontiguousarray(tb.gcbf_count, np.int32),
            gs=np.ascontiguousarray(tb.gcbf_start, np.int32),
            gv=np.ascontiguousarray(tb.gcbf_vector if len(tb.gcbf_vector) else np.zeros(1, np.int32),
                                    np.int32),
            kt=np.ascontiguousarray(tb.kerneltab.reshape(-1)),
            bit=np.ascontiguousarray(tb.bitab.reshape(-1)))
        k = self._keep
        ne = _edge_arrays(tb, k)
        f = OrFaces(len(F.face_id), _pd(k["T"]), _pd(k["LX"]), _pd(k["GX"]), _pd(k["area"]),
                    _pd(k["vel"]), _pd(k["tvel"]), _pd(k["slip"]), _pi(k["nodes"]), _pi(k["btype"]),
                    _pi(k["mode"]), _pi(k["lam"]), _pi(k["gc"]), _pi(k["gs"]), _pi(k["gv"]),
                    _pd(k["kt"]), tb.ktdelta, _pd(k["bit"]), ne, _pi(k["ef"]), _pd(k["egx"]),
                    _pd(k["ecomp"]), _pd(k["elen"]))
        self.ctx = L.or_create(C.byref(self.p), C.byref(f), _pd(self.coord), _pd(vel), _pd(dens),
                               _pd(pres), _pd(mass))
        if db is not None:
            W = tb.walls
            self._wk = [np.ascontiguousarray(a, np.float64).reshape(-1)
                        for a in (W.coord, W.normal, W.vel, W.weight, W.volume, W.mass)]
            self._wflags = np.array([db.gamma_limiter, db.fs_allowed, db.negative_wall_p_allowed,
                                     db.muscl_boundary], np.int32)
            if L.or_set_dbsph(self.ctx, len(W), *[_pd(a) for a in self._wk], _pi(self._wflags)):
                raise ValueError("a DB-SPH wall element lies outside the grid")
            if db.kinematics:
                self._kin = np.ascontiguousarray(db.kinematics, np.float64).reshape(-1, 7)
                self._kc = np.ascontiguousarray(db.rotation_centre, np.float64)
                if L.or_set_dbsph_kinematics(self.ctx, len(self._kin), _pd(self._kin), _pd(self._kc)):
                    raise ValueError("DB-SPH kinematics: the first record's time exceeds 0")

    def __del__(self):
        if getattr(self, "ctx", None) and _lib is not None:
            _lib.or_destroy(self.ctx)
            self.ctx = None

    @property
    def n(self):
        return lib().or_nag(self.ctx)

    def call(self, name, *args):
        rc = getattr(lib(), "or_" + name)(self.ctx, *args)
        if rc:
            raise RuntimeError(f"oracle or_{name} returned {rc}")
        return rc

    def preloop(self):
        return self.call("preloop")

    def step(self, k=1):
        return self.call("step", int(k))

    def set_state(self, st):
        n = int(st["nag"])
        a = {k: np.ascontiguousarray(st[k], np.float64) for k in ("coord", "vel", "var", "acc",
                                                                   "dens", "pres")}
        lam = np.ascontiguousarray(st["laminar_flag"], np.int32)
        lib().or_set_state(self.ctx, n, _pd(a["coord"]), _pd(a["vel"]), _pd(a["var"]),
                           _pd(a["acc"]), _pd(a["dens"]), _pd(a["pres"]), _pi(lam), st["dt"],
                           st["dt_average"], st["time"], int(st["it"]))

    def state(self):
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("coord", "vel", "var", "acc")}
        out.update({k: np.zeros(n) for k in ("dens", "pres", "dden", "vpres")})
        lib().or_get_state(self.ctx, *[_pd(out[k]) for k in ("coord", "vel", "var", "acc", "dens",
                                                              "pres", "dden", "vpres")])
        sc = np.zeros(4)
        lib().or_get_scalars(self.ctx, _pd(sc))
        out.update(dt=sc[0], dt_average=sc[1], time=sc[2], it=int(sc[3]), nag=n)
        return out

    def lists(self):
        n = self.n
        stride = int(self.p.nmaxpartj)
        cnt = np.zeros(n, np.int32)
        ids = np.zeros((n, stride), np.int32)
        if lib().or_get_lists(self.ctx, _pi(cnt), _pi(ids), stride):
            raise RuntimeError("lists not built")
        return cnt, ids

    def first_list_mismatch(self, count, ids):
        """First particle whose list differs from the oracle's (-1: none)."""
        count = np.ascontiguousarray(count, np.int32)
        ids = np.ascontiguousarray(ids, np.int32)
        return int(lib().or_compare_lists(self.ctx, _pi(count), _pi(ids), ids.shape[1]))

    def dbsph_state(self):
        """DB-SPH state: per particle gamma, sigma, rho_sph (rhoSPH_new), dshep,
        drho (n,3), dvel (n,3,3) = dvel(i,j), fs; per wall element wall_dens,
        wall_pres, wall_wet."""
        n, nw = self.n, len(self.tb.walls)
        o = {k: np.zeros(n) for k in ("gamma", "sigma", "rho_sph", "dshep")}
        o["drho"] = np.zeros((n, 3))
        o["dvel"] = np.zeros((n, 3, 3))
        o["fs"] = np.zeros(n, np.int32)
        o["wall_dens"], o["wall_pres"] = np.zeros(nw), np.zeros(nw)
        o["wall_wet"] = np.zeros(nw, np.int32)
        lib().or_get_dbsph(self.ctx, *[_pd(o[k]) for k in ("gamma", "sigma", "rho_sph", "dshep", "drho",
                                                           "dvel")], _pi(o["fs"]), _pd(o["wall_dens"]),
                           _pd(o["wall_pres"]), _pi(o["wall_wet"]))
        return o

    def fw_lists(self):
        """Fluid-wall lists: counts and 0-based wall ids in the reference's order."""
        n = self.n
        stride = int(self.p.nmaxpartj)
        cnt = np.zeros(n, np.int32)
        ids = np.zeros((n, stride), np.int32)
        lib().or_get_fw_lists(self.ctx, _pi(cnt), _pi(ids), stride)
        return cnt, ids

    def btab(self):
        n = self.n
        mc = int(self.p.maxclose)
        nc = np.zeros(n, np.int32)
        face = np.zeros((n, mc), np.int32)
        loc = np.zeros((n, mc, 3))
        bi = np.zeros((n, mc, 8))
        ig = np.zeros((n, mc, 9))
        lib().or_get_btab(self.ctx, _pi(nc), _pi(face), _pd(loc), _pd(bi), _pd(ig))
        return nc, face, loc, bi, ig

    def free_surface(self, icont, points):
        """calc_pelo on the current cell lists: (pelolib, zlft), each (nlines, 3)."""
        icont = np.ascontiguousarray(icont, np.int32).reshape(-1, 2)
        points = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        pel = np.zeros((len(icont), 3))
        zl = np.zeros((len(icont), 3))
        self.call("free_surface", len(icont), _pi(icont), len(points), _pd(points), _pd(pel),
                  _pd(zl))
        return pel, zl

    def pair_count(self):
        return int(lib().or_pair_count(self.ctx))

    def lists_only(self):
        """Build the neighbour lists alone (no rag / PartKernel) on the next
        list build: list checks of large clouds; the passes cannot run."""
        lib().or_set_lists_only(self.ctx, 1)
        return self.call("calcvarlength")

    def outgone_count(self):
        """Particles removed through open faces so far (CancelOutgoneParticles_3D)."""
        return int(lib().or_outgone_count(self.ctx))

    def branch_counts(self):
        """How often the rare branches ran: elastic wall reactions applied,
        zpmin clamps of the SA-SPH volume integrals, density_thresholds clamps,
        wall-function interactions below z0 e and on the logarithmic branch,
        convex-edge reactions applied."""
        v = (C.c_int64 * 6)()
        lib().or_branch_counts(self.ctx, v)
        return dict(elastic=int(v[0]), zpmin=int(v[1]), density_clamp=int(v[2]),
                    wall_fn_near=int(v[3]), wall_fn_log=int(v[4]), convex_edge=int(v[5]))
