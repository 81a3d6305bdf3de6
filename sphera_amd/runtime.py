"""ctypes binding of the device library (include/sphera_hip.h).

This is the host-side mirror of the reference's hot-path interface: one method
per time_step_loop phase (same names as the Fortran units they replace), plus
the state exchange a Fortran ISO_C_BINDING bridge would do.  There is no CPU
fallback: if libsphera_hip.so is missing or no GPU is present, every call
raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# the in-tree build; SPHERA_HIP_LIB names another build of the same sources
LIB_PATH = Path(os.environ.get("SPHERA_HIP_LIB") or
                Path(__file__).resolve().parent / "_build" / "libsphera_hip.so")

FP64, FP32_MIXED = 0, 1
# deck options (SPH_OPT_* of sphera_hip.h)
OPT_DENSITY_THRESHOLDS, OPT_ALE3, OPT_C1, OPT_KTGF, OPT_CONVEX_EDGES, OPT_RK1, OPT_RK2 = \
    1, 2, 4, 8, 16, 32, 64
PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int32)

EXPORTS = [
    "sph_ctx_create", "sph_ctx_create2d", "sph_ctx_destroy", "sph_state_upload", "sph_particle_count",
    "sph_state_download", "sph_get_scalars", "sph_get_laminar", "sph_pair_counter", "sph_global_quantities",
    "sph_global_quantities_noalloc",
    "sph_monitor_lines", "sph_free_surface", "sph_get_neighbour_lists", "sph_get_boundary_table", "sph_dt", "sph_momentum",
    "sph_leapfrog_mom", "sph_velocity_smoothing", "sph_leapfrog_traj", "sph_sort", "sph_neighbours",
    "sph_boundary_tab", "sph_continuity", "sph_leapfrog_cont", "sph_eos", "sph_pressure_smoothing",
    "sph_preloop", "sph_step", "sph_step_async", "sph_step_graph", "sph_sync", "sph_get_launch", "sph_set_timing",
    "sph_get_timing", "sph_ctx_bytes", "sph_device_memory", "sph_query",
    "sph_batch_create", "sph_batch_bytes", "sph_segment_upload", "sph_segment_download",
    "sph_segment_scalars", "sph_segment_count", "sph_dbsph_walls", "sph_dbsph_state", "sph_dbsph_kinematics",
    "sph_last_error",
]

PHASES = ["dt", "momentum", "leapfrog_mom", "velocity_smoothing", "leapfrog_traj", "sort",
          "neighbours", "boundary_tab", "continuity", "leapfrog_cont", "eos", "pressure_smoothing"]


class SphParams(C.Structure):
    _fields_ = [("precision", C.c_int32), ("capacity", C.c_int32), ("nmaxpartj", C.c_int32),
                ("maxclose", C.c_int32), ("ncd", C.c_int32 * 3), ("extr_lo", C.c_double * 3),
                ("dcd", C.c_double * 3), ("dx", C.c_double), ("h", C.c_double),
                ("coefke", C.c_double), ("coefkacl", C.c_double), ("eta", C.c_double),
                ("eta2", C.c_double), ("grav", C.c_double * 3), ("prif", C.c_double),
                ("den0", C.c_double), ("eps", C.c_double), ("celerita", C.c_double),
                ("alfamon", C.c_double), ("kin_visc", C.c_double), ("cfl", C.c_double),
                ("vsc_coeff", C.c_double), ("pesodt", C.c_double), ("tetap", C.c_double),
                ("tetav", C.c_double), ("semi", C.c_int32), ("wg_size", C.c_int32),
                ("chunk", C.c_int32), ("chunk_v", C.c_int32), ("chunk_c", C.c_int32),
                ("chunk_p", C.c_int32), ("nb_shape", C.c_int32), ("options", C.c_int32)]


class SphFaces(C.Structure):
    _fields_ = [("nf", C.c_int32), ("T", PD), ("LX", PD), ("GX", PD), ("area", PD),
                ("velocity", PD), ("tvel", PD), ("slip_coeff", PD), ("nodes", PI),
                ("btype", PI), ("slip_mode", PI), ("lam_check", PI), ("gcbf_count", PI),
                ("gcbf_start", PI), ("gcbf_vec", PI), ("kerneltab", PD), ("ktdelta", C.c_double),
                ("bitab", PD), ("nedges", C.c_int32), ("edge_face", PI), ("edge_gx", PD),
                ("edge_comp", PD), ("edge_len", PD)]


class SphSides2D(C.Structure):
    _fields_ = [("ns", C.c_int32), ("T", PD), ("RN", PD), ("v1", PD), ("length", PD),
                ("angle", PD), ("velocity", PD), ("tvel", PD), ("slip_coeff", PD),
                ("btype", PI), ("previous", PI), ("slip_mode", PI), ("lam_check", PI),
                ("bit2d", PD), ("v2", PD)]


class SphDBSPHWalls(C.Structure):
    _fields_ = [("nw", C.c_int32), ("coord", PD), ("normal", PD), ("vel", PD), ("weight", PD),
                ("volume", PD), ("mass", PD), ("gamma_limiter", C.c_int32), ("fs_allowed", C.c_int32),
                ("negative_wall_p_allowed", C.c_int32), ("muscl_boundary", C.c_int32)]


_lib = None


def load():
    """Load libsphera_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    L = C.CDLL(str(LIB_PATH))
    L.sph_ctx_create.argtypes = [C.POINTER(SphParams), C.POINTER(SphFaces), C.POINTER(C.c_void_p)]
    L.sph_ctx_create2d.argtypes = [C.POINTER(SphParams), C.POINTER(SphSides2D), C.POINTER(C.c_void_p)]
    L.sph_ctx_destroy.argtypes = [C.c_void_p]
    L.sph_state_upload.argtypes = [C.c_void_p, C.c_int32, PD, PD, PD, PD, PD]
    L.sph_particle_count.argtypes = [C.c_void_p, PI]
    L.sph_state_download.argtypes = [C.c_void_p, PD, PD, PD, PD, PD, PD, PD, PI]
    L.sph_get_scalars.argtypes = [C.c_void_p, PD]
    L.sph_get_laminar.argtypes = [C.c_void_p, PI]
    L.sph_get_neighbour_lists.argtypes = [C.c_void_p, PI, PI, C.c_int32]
    L.sph_get_boundary_table.argtypes = [C.c_void_p, PI, PI, PD, PD, PD, C.c_int32]
    L.sph_pair_counter.argtypes = [C.c_void_p, C.POINTER(C.c_uint64)]
    L.sph_global_quantities.argtypes = [C.c_void_p, PD]
    L.sph_global_quantities_noalloc.argtypes = [C.c_void_p, PD]
    L.sph_dbsph_walls.argtypes = [C.c_void_p, C.POINTER(SphDBSPHWalls)]
    L.sph_dbsph_kinematics.argtypes = [C.c_void_p, C.c_int32, PD, PD]
    L.sph_dbsph_state.argtypes = [C.c_void_p, PD, PD, PD, PD, PI, PD, PD, PI]
    L.sph_monitor_lines.argtypes = [C.c_void_p, C.c_int32, PI, C.c_int32, PD]
    L.sph_free_surface.argtypes = [C.c_void_p, PD, PD]
    for ph in PHASES + ["preloop", "sync"]:
        getattr(L, "sph_" + ph).argtypes = [C.c_void_p]
    L.sph_step.argtypes = [C.c_void_p, C.c_int32]
    L.sph_step_async.argtypes = [C.c_void_p, C.c_int32]
    L.sph_step_graph.argtypes = [C.c_void_p, C.c_int32]
    L.sph_set_timing.argtypes = [C.c_void_p, C.c_int32]
    L.sph_get_launch.argtypes = [C.c_void_p, PI]
    L.sph_get_timing.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, PD, PI, C.c_int32]
    L.sph_ctx_bytes.argtypes = [C.POINTER(SphParams), C.POINTER(C.c_uint64)]
    L.sph_device_memory.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.sph_query.argtypes = [C.c_void_p, PI]
    L.sph_batch_create.argtypes = [C.c_int32, C.POINTER(SphParams), C.POINTER(C.POINTER(SphFaces)),
                                   C.POINTER(C.c_void_p)]
    L.sph_batch_bytes.argtypes = [C.c_int32, C.POINTER(SphParams), C.POINTER(C.c_uint64)]
    L.sph_segment_upload.argtypes = [C.c_void_p, C.c_int32, C.c_int32, PD, PD, PD, PD, PD]
    L.sph_segment_download.argtypes = [C.c_void_p, C.c_int32, PD, PD, PD, PD, PD, PD, PD, PI]
    L.sph_segment_scalars.argtypes = [C.c_void_p, C.c_int32, PD]
    L.sph_segment_count.argtypes = [C.c_void_p, C.c_int32, PI]
    L.sph_last_error.restype = C.c_char_p
    for name in EXPORTS:
        if name != "sph_last_error":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


class SphError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


def _pd(a):
    return a.ctypes.data_as(PD)


def _pi(a):
    return a.ctypes.data_as(PI)


def faces_struct(tb):
    """Pack CaseTables frontier data in the C layout; returns (struct, keepalive)."""
    F = tb.faces
    k = dict(
        T=np.ascontiguousarray(F.T.transpose(0, 2, 1).reshape(-1)),
        LX=np.ascontiguousarray(F.LX.reshape(-1)), GX=np.ascontiguousarray(F.GX.reshape(-1)),
        area=np.ascontiguousarray(F.area, np.float64),
        vel=np.ascontiguousarray(F.velocity.reshape(-1)),
        tvel=np.ascontiguousarray(F.tratto_velocity.reshape(-1)),
        slip=np.ascontiguousarray(F.slip_coeff, np.float64),
        nodes=np.ascontiguousarray(F.nodes, np.int32), btype=np.ascontiguousarray(F.btype, np.int32),
        mode=np.ascontiguousarray(F.slip_mode, np.int32),
        lam=np.ascontiguousarray(F.laminar_check, np.int32),
        gc=np.ascontiguousarray(tb.gcbf_count, np.int32),
        gs=np.ascontiguousarray(tb.gcbf_start, np.int32),
        gv=np.ascontiguousarray(tb.gcbf_vector if len(tb.gcbf_vector) else np.zeros(1), np.int32),
        kt=np.ascontiguousarray(tb.kerneltab.reshape(-1)), bit=np.ascontiguousarray(tb.bitab.reshape(-1)))
    ne = edge_arrays(tb, k)
    s = SphFaces(len(F.face_id), _pd(k["T"]), _pd(k["LX"]), _pd(k["GX"]), _pd(k["area"]),
                 _pd(k["vel"]), _pd(k["tvel"]), _pd(k["slip"]), _pi(k["nodes"]), _pi(k["btype"]),
                 _pi(k["mode"]), _pi(k["lam"]), _pi(k["gc"]), _pi(k["gs"]), _pi(k["gv"]),
                 _pd(k["kt"]), tb.ktdelta, _pd(k["bit"]), ne, _pi(k["ef"]), _pd(k["egx"]),
                 _pd(k["ecomp"]), _pd(k["elen"]))
    return s, k


def edge_arrays(tb, k):
    """Adds the convex-edge arrays of CaseTables (contiguous, at least one
    element each) to the keepalive dict k; returns the number of edges."""
    E = getattr(tb, "edges", None)
    ne = len(E) if E is not None else 0
    k["ef"] = np.ascontiguousarray(E.face.reshape(-1) if ne else np.zeros(2), np.int32)
    k["egx"] = np.ascontiguousarray(E.gx.reshape(-1) if ne else np.zeros(3), np.float64)
    k["ecomp"] = np.ascontiguousarray(E.component.reshape(-1) if ne else np.zeros(3), np.float64)
    k["elen"] = np.ascontiguousarray(E.length if ne else np.zeros(1), np.float64)
    return ne


def sides_struct(tb):
    """Pack CaseTables2D boundary sides in the C layout; returns (struct, keepalive)."""
    S = tb.sides
    k = dict(T=np.ascontiguousarray(S.T.transpose(0, 2, 1).reshape(-1)),
             RN=np.ascontiguousarray(S.RN.transpose(0, 2, 1).reshape(-1)),
             v1=np.ascontiguousarray(S.v1.reshape(-1)), v2=np.ascontiguousarray(S.v2.reshape(-1)),
             L=np.ascontiguousarray(S.length, np.float64),
             ang=np.ascontiguousarray(S.angle, np.float64),
             vel=np.ascontiguousarray(S.velocity.reshape(-1)),
             tvel=np.ascontiguousarray(S.tratto_velocity.reshape(-1)),
             slip=np.ascontiguousarray(S.slip_coeff, np.float64),
             bt=np.ascontiguousarray(S.btype, np.int32), prev=np.ascontiguousarray(S.previous, np.int32),
             mode=np.ascontiguousarray(S.slip_mode, np.int32),
             lam=np.ascontiguousarray(S.laminar_check, np.int32),
             bit=np.ascontiguousarray(tb.bit2d.reshape(-1), np.float64))
    s = SphSides2D(len(S.length), _pd(k["T"]), _pd(k["RN"]), _pd(k["v1"]), _pd(k["L"]), _pd(k["ang"]),
                   _pd(k["vel"]), _pd(k["tvel"]), _pd(k["slip"]), _pi(k["bt"]), _pi(k["prev"]),
                   _pi(k["mode"]), _pi(k["lam"]), _pd(k["bit"]), _pd(k["v2"]))
    return s, k


def params_struct(tb, precision=FP64, capacity=None, semi=True, tune=None):
    s, g, case = tb.scalars, tb.grid, tb.case
    p = SphParams()
    p.precision = precision
    p.capacity = int(capacity or len(tb.particles["coord"]))
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
    p.alfamon, p.kin_visc = case.alfa_mon, s["kin_visc"]
    p.cfl, p.vsc_coeff, p.pesodt = case.cfl, case.vsc_coeff, case.pesodt
    p.tetap, p.tetav = case.teta_p, case.teta_v
    p.semi = 1 if semi else 0
    p.options = OPT_DENSITY_THRESHOLDS if int(getattr(case, "density_thresholds", 0)) == 1 else 0
    if int(getattr(case, "time_split", 1)) == 0:
        p.options |= OPT_RK1
    if semi and getattr(tb, "edges", None) is not None and len(tb.edges):
        p.options |= OPT_CONVEX_EDGES
    for k, v in (tune or {}).items():
        setattr(p, k, int(v))
    return p


def ctx_bytes(tb, precision=FP64, capacity=None, semi=True):
    """Device bytes a context for this case would hold."""
    L = load()
    v = C.c_uint64()
    p = params_struct(tb, precision, capacity, semi)
    rc = L.sph_ctx_bytes(C.byref(p), C.byref(v))
    if rc:
        raise SphError(rc, "sph_ctx_bytes", L.sph_last_error().decode())
    return int(v.value)


def batch_bytes(tbs, precision=FP32_MIXED, semi=True):
    """Device bytes a batch context holding these cases would hold."""
    L = load()
    v = C.c_uint64()
    ps = (SphParams * len(tbs))(*[params_struct(tb, precision, None, semi) for tb in tbs])
    rc = L.sph_batch_bytes(len(tbs), ps, C.byref(v))
    if rc:
        raise SphError(rc, "sph_batch_bytes", L.sph_last_error().decode())
    return int(v.value)


def device_memory():
    """(free, total) device bytes."""
    L = load()
    f, t = C.c_uint64(), C.c_uint64()
    rc = L.sph_device_memory(C.byref(f), C.byref(t))
    if rc:
        raise SphError(rc, "sph_device_memory", L.sph_last_error().decode())
    return int(f.value), int(t.value)


class SphContext:
    """One device solve (one C-ABI context)."""

    def __init__(self, tb, precision=FP64, capacity=None, semi=True, tune=None):
        self.L = load()
        self.tb = tb
        self.precision = precision
        db = getattr(tb.case, "dbsph", None) if getattr(tb, "dim", 3) == 3 else None
        if db is not None:  # DB-SPH walls instead of SA-SPH frontiers
            semi = False
        self.params = params_struct(tb, precision, capacity, semi, tune)
        h = C.c_void_p()
        if db is not None:
            self._check(self.L.sph_ctx_create(C.byref(self.params), None, C.byref(h)), "sph_ctx_create")
            self.h = h
            W = tb.walls
            self._wkeep = [np.ascontiguousarray(a, np.float64).reshape(-1)
                           for a in (W.coord, W.normal, W.vel, W.weight, W.volume, W.mass)]
            ws = SphDBSPHWalls(len(W), *[_pd(a) for a in self._wkeep], int(db.gamma_limiter),
                               int(db.fs_allowed), int(db.negative_wall_p_allowed), int(db.muscl_boundary))
            self._check(self.L.sph_dbsph_walls(self.h, C.byref(ws)), "sph_dbsph_walls")
            if db.kinematics:
                self._kin = np.ascontiguousarray(db.kinematics, np.float64).reshape(-1, 7)
                self._kc = np.ascontiguousarray(db.rotation_centre, np.float64)
                self._check(self.L.sph_dbsph_kinematics(self.h, len(self._kin), _pd(self._kin), _pd(self._kc)),
                            "sph_dbsph_kinematics")
            return
        if getattr(tb, "dim", 3) == 2:
            ss, self._keep = sides_struct(tb)
            self._check(self.L.sph_ctx_create2d(C.byref(self.params), C.byref(ss), C.byref(h)),
                        "sph_ctx_create2d")
        else:
            fs, self._keep = faces_struct(tb)
            self._check(self.L.sph_ctx_create(C.byref(self.params), C.byref(fs), C.byref(h)),
                        "sph_ctx_create")
        self.h = h

    def _check(self, rc, where):
        if rc:
            raise SphError(rc, where, self.L.sph_last_error().decode())
        return rc

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.sph_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state ------------------------------------------------------------------
    def upload(self, coord=None, vel=None, dens=None, pres=None, mass=None):
        P = self.tb.particles
        a = [np.ascontiguousarray(P[k] if v is None else v, np.float64)
             for k, v in (("coord", coord), ("vel", vel), ("dens", dens), ("pres", pres),
                          ("mass", mass))]
        self._check(self.L.sph_state_upload(self.h, len(a[0]), *[_pd(x) for x in a]),
                    "sph_state_upload")

    @property
    def n(self):
        v = np.zeros(1, np.int32)
        self._check(self.L.sph_particle_count(self.h, _pi(v)), "sph_particle_count")
        return int(v[0])

    def dbsph_state(self):
        """DB-SPH state in pg order: gamma, sigma, rho_sph, dshep, fs; wall_dens,
        wall_pres, wall_wet per wall element."""
        n, nw = self.n, len(self.tb.walls)
        o = {k: np.zeros(n) for k in ("gamma", "sigma", "rho_sph", "dshep")}
        o["fs"] = np.zeros(n, np.int32)
        o["wall_dens"], o["wall_pres"] = np.zeros(nw), np.zeros(nw)
        o["wall_wet"] = np.zeros(nw, np.int32)
        self._check(self.L.sph_dbsph_state(self.h, *[_pd(o[k]) for k in ("gamma", "sigma", "rho_sph", "dshep")],
                                           _pi(o["fs"]), _pd(o["wall_dens"]), _pd(o["wall_pres"]),
                                           _pi(o["wall_wet"])), "sph_dbsph_state")
        return o

    def state(self):
        n = self.n
        out = {k: np.zeros((n, 3)) for k in ("coord", "vel", "acc")}
        out.update({k: np.zeros(n) for k in ("dens", "pres", "dden", "vpres")})
        ids = np.zeros(n, np.int32)
        self._check(self.L.sph_state_download(self.h, *[_pd(out[k]) for k in (
            "coord", "vel", "acc", "dens", "pres", "dden", "vpres")], _pi(ids)), "sph_state_download")
        out["ids"] = ids
        out["var"] = out["vel"]
        sc = self.scalars()
        out.update(sc)
        out["nag"] = n
        return out

    def laminar(self):
        """Laminar flags in pg order (pg%laminar_flag)."""
        v = np.zeros(self.n, np.int32)
        self._check(self.L.sph_get_laminar(self.h, _pi(v)), "sph_get_laminar")
        return v

    def scalars(self):
        sc = np.zeros(5)
        self._check(self.L.sph_get_scalars(self.h, _pd(sc)), "sph_get_scalars")
        return dict(dt=sc[0], dt_average=sc[1], time=sc[2], it=int(sc[3]), pairs=int(sc[4]))

    def pair_counter(self):
        v = C.c_uint64(0)
        self._check(self.L.sph_pair_counter(self.h, C.byref(v)), "sph_pair_counter")
        return int(v.value)

    def lists(self):
        n = self.n
        stride = int(self.params.nmaxpartj)
        cnt = np.zeros(n, np.int32)
        ids = np.zeros((n, stride), np.int32)
        self._check(self.L.sph_get_neighbour_lists(self.h, _pi(cnt), _pi(ids), stride),
                    "sph_get_neighbour_lists")
        return cnt, ids

    def btab(self, stride=8):
        """Boundary table in pg order: ncbf, face, loc, bi, ig (oracle layout)."""
        n = self.n
        nc = np.zeros(n, np.int32)
        face = np.zeros((n, stride), np.int32)
        loc = np.zeros((n, stride, 3))
        bi = np.zeros((n, stride, 8))
        ig = np.zeros((n, stride, 9))
        self._check(self.L.sph_get_boundary_table(self.h, _pi(nc), _pi(face), _pd(loc), _pd(bi),
                                                  _pd(ig), stride), "sph_get_boundary_table")
        return nc, face, loc, bi, ig

    # phases -----------------------------------------------------------------
    def phase(self, name):
        return self._check(getattr(self.L, "sph_" + name)(self.h), "sph_" + name)

    def preloop(self):
        return self.phase("preloop")

    def step(self, k=1):
        return self._check(self.L.sph_step(self.h, int(k)), "sph_step")

    def step_async(self, k=1):
        return self._check(self.L.sph_step_async(self.h, int(k)), "sph_step_async")

    def step_graph(self, k=1):
        """k steps as one replayed HIP graph (asynchronous; sync() to wait)."""
        return self._check(self.L.sph_step_graph(self.h, int(k)), "sph_step_graph")

    def sync(self):
        return self.phase("sync")

    def global_quantities(self, noalloc=False):
        """mass, linear momentum (3), angular momentum (3), volume
        (noalloc: through sph_global_quantities_noalloc, which allocates no
        device memory)."""
        v = np.zeros(8)
        name = "sph_global_quantities_noalloc" if noalloc else "sph_global_quantities"
        self._check(getattr(self.L, name)(self.h, _pd(v)), name)
        return dict(mass=v[0], lin_mom=v[1:4].copy(), ang_mom=v[4:7].copy(), volume=v[7])

    def monitor_lines(self, icont, points):
        """Sets the free-surface monitoring lines: icont (nlines, 2) 1-based
        first/last point of each line, points (npoints, 3)."""
        self._icont = np.ascontiguousarray(icont, np.int32).reshape(-1, 2)
        pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        self._check(self.L.sph_monitor_lines(self.h, len(self._icont), _pi(self._icont), len(pts),
                                             _pd(pts)), "sph_monitor_lines")

    def free_surface(self):
        """calc_pelo: (pelolib, zlft), each (nlines, 3)."""
        nl = len(getattr(self, "_icont", ()))
        pel, zl = np.zeros((nl, 3)), np.zeros((nl, 3))
        self._check(self.L.sph_free_surface(self.h, _pd(pel), _pd(zl)), "sph_free_surface")
        return pel, zl

    def busy(self):
        """True while work enqueued on this context's stream is running."""
        v = np.zeros(1, np.int32)
        self._check(self.L.sph_query(self.h, _pi(v)), "sph_query")
        return bool(v[0])

    def launch(self):
        v = np.zeros(3, np.int32)
        self._check(self.L.sph_get_launch(self.h, _pi(v)), "sph_get_launch")
        return dict(wg_size=int(v[0]), chunk=int(v[1]), grid=int(v[2]))

    def set_timing(self, on=True):
        self._check(self.L.sph_set_timing(self.h, 1 if on else 0), "sph_set_timing")

    def timing(self):
        names = C.create_string_buffer(4096)
        ms = np.zeros(64)
        calls = np.zeros(64, np.int32)
        k = self.L.sph_get_timing(self.h, names, 4096, _pd(ms), _pi(calls), 64)
        nm = names.value.decode().split(";")[:k]
        return {nm[i]: (float(ms[i]), int(calls[i])) for i in range(k)}


class SphBatch(SphContext):
    """Several independent 3D solves concatenated in one context (segments):
    each kernel of a step is launched once over all of them.  Segment k is
    case tbs[k]; state methods take the segment index."""

    def __init__(self, tbs, precision=FP32_MIXED, semi=True, tune=None):
        self.L = load()
        self.tbs = list(tbs)
        self.tb = self.tbs[0]
        self.precision = precision
        n = len(self.tbs)
        self.params_all = (SphParams * n)(*[params_struct(tb, precision, None, semi, tune)
                                            for tb in self.tbs])
        self.params = self.params_all[0]
        packed = [faces_struct(tb) for tb in self.tbs]
        self._keep = packed
        self._faces = [f for f, _ in packed]
        fptrs = (C.POINTER(SphFaces) * n)(*[C.pointer(f) for f in self._faces])
        h = C.c_void_p()
        self._check(self.L.sph_batch_create(n, self.params_all, fptrs, C.byref(h)), "sph_batch_create")
        self.h = h
        self.nseg = n

    def upload(self, coords=None):
        """Uploads every segment (initial state of each case; coords[k]
        replaces case k's positions when given)."""
        for k, tb in enumerate(self.tbs):
            P = tb.particles
            c = P["coord"] if coords is None or coords[k] is None else coords[k]
            a = [np.ascontiguousarray(v, np.float64)
                 for v in (c, P["vel"], P["dens"], P["pres"], P["mass"])]
            self._check(self.L.sph_segment_upload(self.h, k, len(a[0]), *[_pd(x) for x in a]),
                        "sph_segment_upload")

    def count(self, k):
        v = np.zeros(1, np.int32)
        self._check(self.L.sph_segment_count(self.h, k, _pi(v)), "sph_segment_count")
        return int(v[0])

    @property
    def n(self):
        return sum(self.count(k) for k in range(self.nseg))

    def seg_scalars(self, k):
        sc = np.zeros(5)
        self._check(self.L.sph_segment_scalars(self.h, k, _pd(sc)), "sph_segment_scalars")
        return dict(dt=sc[0], dt_average=sc[1], time=sc[2], it=int(sc[3]), pairs=int(sc[4]))

    def seg_state(self, k):
        n = self.count(k)
        out = {q: np.zeros((n, 3)) for q in ("coord", "vel", "acc")}
        out.update({q: np.zeros(n) for q in ("dens", "pres", "dden", "vpres")})
        ids = np.zeros(n, np.int32)
        self._check(self.L.sph_segment_download(self.h, k, *[_pd(out[q]) for q in (
            "coord", "vel", "acc", "dens", "pres", "dden", "vpres")], _pi(ids)), "sph_segment_download")
        out["ids"] = ids
        out["var"] = out["vel"]
        out.update(self.seg_scalars(k))
        out["nag"] = n
        return out

    def state(self):
        return self.seg_state(0)

    def scalars(self):
        return self.seg_scalars(0)
