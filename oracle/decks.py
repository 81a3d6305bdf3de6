"""Oracle driver: writes a SPHERA v10 input deck for a DamBreak3D case and runs
the flang-built reference harness on it.  Test infrastructure only.

Deck syntax follows the v10 readers (src/Pre_processing/reading_inp/
ReadInput.f90:102-194 and one reader per section); the section set and the
values not describing the case (DB-SPH, bed load, monitors) mirror the v10
3D tutorial deck input/tutorials/release_v_10_0_0/42_dam_break_3D_demo, with
outputs throttled.  Integers are written without a decimal point (flang's
list-directed read rejects "0." for an integer).
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REF_DIR = HERE / "_ref"
HARNESS = REF_DIR / "harness3d.x"
HARNESS_HIP = REF_DIR / "harness3d_hip.x"
REFERENCE = REF_DIR / "sphera3d.x"
HARNESS2D = REF_DIR / "harness2d.x"
HARNESS2D_HIP = REF_DIR / "harness2d_hip.x"
REFERENCE2D = REF_DIR / "sphera2d.x"


from sphera_amd.deck import deck_text, deck_text_2d  # noqa: E402,F401  (the v10 deck writer)


def write_case_dir(case, root, itmax=1000000000, results_step=None, text=None):
    """Lay out a run directory the reference can start from: it concatenates
    ./input/*/*.inp into <case>.inp (src/Main_algorithm/sphera.f90:222).
    text: a deck to run instead of the one written from `case`."""
    root = Path(root)
    (root / "input" / "00_deck").mkdir(parents=True, exist_ok=True)
    if text is None:
        text = deck_text_2d(case, itmax=itmax, results_step=results_step) if getattr(case, "dim", 3) == 2 \
            else deck_text(case, itmax=itmax, results_step=results_step)
    (root / "input" / "00_deck" / "deck.inp").write_text(text)
    if getattr(case, "dbsph", None) is not None:  # the tank's surface mesh (input/07_DBSPH)
        from sphera_amd.dbsph import tank_mesh, write_ply
        (root / "input" / "07_DBSPH").mkdir(parents=True, exist_ok=True)
        write_ply(root / "input" / "07_DBSPH" / "tank.ply", *tank_mesh(case.tank, case.dx))
        (root / "input" / "07_DBSPH" / "surface_mesh_list.txt").write_text(
            "surface_mesh_list.txt\ntank.ply\n")
    (root / "output").mkdir(exist_ok=True)
    (root / "restart").mkdir(exist_ok=True)
    return root


def run_harness(case, nsteps, every=1, lists=True, ic_coord=None, threads=None,
                workdir=None, timeout=3600, exe=None, extra_env=None, ic_vel=None, text=None):
    """Run the reference harness (or ``exe``, e.g. the device drop-in harness
    HARNESS_HIP); returns the path of its dump file."""
    exe = Path(exe or (HARNESS2D if getattr(case, "dim", 3) == 2 else HARNESS))
    if not exe.exists():
        raise FileNotFoundError(f"{exe} not built (make -f oracle/Makefile.ref)")
    wd = Path(workdir or tempfile.mkdtemp(prefix="sph_ref_")).resolve()
    write_case_dir(case, wd, text=text)
    env = dict(os.environ)
    env["SPH_HARNESS_OUT"] = str(wd / "dump.bin")
    env["SPH_HARNESS_STEPS"] = str(int(nsteps))
    env["SPH_HARNESS_EVERY"] = str(int(every))
    env["SPH_HARNESS_LISTS"] = "1" if lists else "0"
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    env.update(extra_env or {})
    if ic_coord is not None:
        icf = wd / "ic.bin"
        with open(icf, "wb") as fh:
            fh.write(struct.pack("<i", len(ic_coord)))
            fh.write(np.ascontiguousarray(ic_coord, dtype="<f8").tobytes())
        env["SPH_HARNESS_IC"] = str(icf)
    if ic_vel is not None:
        ivf = wd / "icv.bin"
        with open(ivf, "wb") as fh:
            fh.write(struct.pack("<i", len(ic_vel)))
            fh.write(np.ascontiguousarray(ic_vel, dtype="<f8").tobytes())
        env["SPH_HARNESS_ICV"] = str(ivf)
    res = subprocess.run([str(exe), case.name], cwd=wd, env=env, capture_output=True,
                         text=True, timeout=timeout)
    if res.returncode != 0 or not (wd / "dump.bin").exists():
        raise RuntimeError(f"reference harness failed ({res.returncode}):\n{res.stdout[-3000:]}"
                           f"\n{res.stderr[-3000:]}")
    return wd / "dump.bin"


def run_reference(case, itmax, results_step, workdir=None, threads=None, timeout=3600):
    """Run the unmodified reference program (sphera3d.x, or sphera2d.x for a
    2D case; its own time_step_loop) on the case's deck for itmax steps, writing results every
    results_step steps; returns the run directory (the results are in
    <case>.ris there)."""
    exe = REFERENCE2D if getattr(case, "dim", 3) == 2 else REFERENCE
    if not exe.exists():
        raise FileNotFoundError(f"{exe} not built (make -f oracle/Makefile.ref)")
    wd = Path(workdir or tempfile.mkdtemp(prefix="sph_ris_")).resolve()
    write_case_dir(case, wd, itmax=itmax, results_step=results_step)
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    res = subprocess.run([str(exe), case.name], cwd=wd, env=env, capture_output=True,
                         text=True, timeout=timeout)
    if res.returncode != 0 or not (wd / f"{case.name}.ris").exists():
        raise RuntimeError(f"reference failed ({res.returncode}):\n{res.stdout[-3000:]}"
                           f"\n{res.stderr[-3000:]}")
    return wd


from sphera_amd.results import read_results as read_ris  # noqa: E402,F401
from sphera_amd.results import records as fortran_records  # noqa: E402,F401


def read_pelo(workdir, case):
    """The reference's free-surface monitor files <case>.plb and <case>.zlft
    (calc_pelo.f90:132-133: time, then pelolib(3, nlines) in g14.7 records of
    30 values).  Returns (time (k,), pelolib (k, nlines, 3), zlft (k, nlines, 3))."""
    nl = len(case.control_lines)
    out = []
    for ext in ("plb", "zlft"):
        tok = (Path(workdir) / f"{case.name}.{ext}").read_text().split("\n", 1)[1].split()
        v = np.array([float(t) for t in tok]).reshape(-1, 1 + 3 * nl)
        out.append(v)
    return out[0][:, 0], out[0][:, 1:].reshape(-1, nl, 3), out[1][:, 1:].reshape(-1, nl, 3)


class _Reader:
    def __init__(self, buf):
        self.b = buf
        self.o = 0

    def arr(self, dt, n):
        a = np.frombuffer(self.b, dtype=dt, count=n, offset=self.o)
        self.o += a.nbytes
        return a.copy()

    def i4(self, n=1):
        a = self.arr("<i4", n)
        return int(a[0]) if n == 1 else a

    def f8(self, n=1):
        a = self.arr("<f8", n)
        return float(a[0]) if n == 1 else a

    def raw(self, n):
        s = self.b[self.o:self.o + n]
        self.o += n
        return s

    def eof(self):
        return self.o >= len(self.b)


def read_dump(path):
    """Parse the harness stream (layout: time_step_loop_dump.f90 write_header,
    write_step).  Arrays are in the reference's pg order; neighbour ids and face
    ids are the reference's 1-based indices."""
    r = _Reader(Path(path).read_bytes())
    if r.i4() != 1213417555:
        raise ValueError("not a harness dump")
    hdr = {}
    ver, nag, nmaxpartj, nmax = r.i4(4)
    ncd = r.i4(3)
    hdr.update(version=ver, nag=nag, nmaxpartj=nmaxpartj, nmax=nmax, ncd=ncd)
    hdr["extr_lo"] = r.f8(3)
    hdr["extr_hi"] = r.f8(3)
    hdr["dcd"] = r.f8(3)
    (hdr["h"], hdr["dx"], hdr["coefke"], hdr["coefkacl"], hdr["prif"]) = r.f8(5)
    hdr["grav"] = r.f8(3)
    hdr["eta"], hdr["eta2"], hdr["pvolume"] = r.f8(3)
    (hdr["den0"], hdr["eps"], hdr["celerita"], hdr["alfamon"], hdr["kin_visc"],
     hdr["mumx"]) = r.f8(6)
    hdr["cfl"], hdr["vsc_coeff"], hdr["pesodt"], hdr["tetap"], hdr["tetav"] = r.f8(5)
    hdr["dt_alfa_mon"], hdr["maxclose"], numfacce, hdr["numbedges"] = r.i4(4)
    if ver == 2:
        return _read_dump_2d(r, hdr, numfacce, nmaxpartj)
    if ver == 3:
        return _read_dump_bsph(r, hdr, numfacce)
    faces = []
    for _ in range(numfacce):
        nf = r.i4()
        if nf == 0:
            continue
        nodes, st = r.i4(2)
        tipo = r.raw(4).decode()
        vals = r.f8(1 + 9 + 3 + 3)
        nd = []
        for _ in range(6):
            name = r.i4()
            gx = r.f8(3)
            lx = r.f8(3)
            nd.append((name, gx, lx))
        mode = r.i4()
        coeff = r.f8()
        lam = r.i4()
        faces.append(dict(nf=nf, nodes=nodes, stretch=st, tipo=tipo, area=vals[0],
                          T=vals[1:10].reshape(3, 3, order="F"), velocity=vals[10:13],
                          tratto_velocity=vals[13:16], node=nd, slip_mode=mode,
                          slip_coeff=coeff, laminar_check=lam))
    hdr["faces"] = faces
    hdr["gcbf_count"] = r.i4(nmax)
    hdr["gcbf_start"] = r.i4(nmax)
    nv = r.i4()
    hdr["gcbf_vector"] = r.i4(nv) if nv > 0 else np.zeros(0, np.int32)
    hdr["kerneltab"] = r.f8(41 * 5).reshape(41, 5, order="F")
    hdr["ktdelta"] = r.f8()
    hdr["bitab"] = r.f8(256 * 7).reshape(256, 7, order="F")
    steps = []
    while not r.eof():
        if r.i4() != 1347704147:
            raise ValueError("bad step tag")
        it, n = r.i4(2)
        t, dt, dtavg = r.f8(3)
        st = dict(it=it, nag=n, time=t, dt=dt, dt_average=dtavg)
        for k in ("coord", "vel", "var", "acc"):
            st[k] = r.f8(3 * n).reshape(n, 3)
        for k in ("dens", "pres", "dden", "vpres", "mass"):
            st[k] = r.f8(n)
        for k in ("cella", "kodvel", "koddens", "laminar_flag"):
            st[k] = r.i4(n)
        lists = r.i4()
        if lists:
            cnt = r.i4(n)
            tot = int(cnt.sum())
            st["nbr_count"] = cnt
            st["nbr"] = r.i4(tot) if tot else np.zeros(0, np.int32)
            ncb = r.i4(n)
            st["ncbf"] = ncb
            nb = int(ncb.sum())
            rec = np.dtype([("face", "<i4"), ("loc", "<f8", 3), ("bi", "<f8", 8),
                            ("ig", "<f8", 9)])
            st["btab"] = np.frombuffer(r.raw(nb * rec.itemsize), dtype=rec).copy()
        steps.append(st)
    return hdr, steps


def _read_dump_2d(r, hdr, nsides, nmaxpartj):
    """2D layout (harness_dump_io.f90 under SPACE_2D): the boundary sides in
    the header; per step Ncbs and IntNcbs per particle, then IntNcbs records."""
    sides = []
    for _ in range(nsides):
        st, prev = r.i4(2)
        vtx = r.i4(2)
        tipo = r.raw(4).decode()
        vals = r.f8(2 + 27 + 3 + 3 + 3 + 3)
        mode = r.i4()
        coeff = r.f8()
        lam = r.i4()
        sides.append(dict(stretch=st, previous=prev, vertex=vtx, tipo=tipo, length=vals[0],
                          angle=vals[1], T=vals[2:11].reshape(3, 3, order="F"),
                          R=vals[11:20].reshape(3, 3, order="F"),
                          RN=vals[20:29].reshape(3, 3, order="F"), velocity=vals[29:32],
                          tratto_velocity=vals[32:35], v1=vals[35:38], v2=vals[38:41],
                          slip_mode=mode, slip_coeff=coeff, laminar_check=lam))
    hdr["sides"] = sides
    steps = []
    while not r.eof():
        if r.i4() != 1347704147:
            raise ValueError("bad step tag")
        it, n = r.i4(2)
        t, dt, dtavg = r.f8(3)
        st = dict(it=it, nag=n, time=t, dt=dt, dt_average=dtavg)
        for k in ("coord", "vel", "var", "acc"):
            st[k] = r.f8(3 * n).reshape(n, 3)
        for k in ("dens", "pres", "dden", "vpres", "mass"):
            st[k] = r.f8(n)
        for k in ("cella", "kodvel", "koddens", "laminar_flag"):
            st[k] = r.i4(n)
        lists = r.i4()
        if lists:
            cnt = r.i4(n)
            tot = int(cnt.sum())
            st["nbr_count"] = cnt
            st["nbr"] = r.i4(tot) if tot else np.zeros(0, np.int32)
            st["ncbs"] = r.i4(n)
            nint_ = r.i4(n)
            st["intncbs"] = nint_
            nb = int(nint_.sum())
            rec = np.dtype([("face", "<i4"), ("loc", "<f8", 3), ("bi", "<f8", 8),
                            ("ig", "<f8", 9)])
            st["btab"] = np.frombuffer(r.raw(nb * rec.itemsize), dtype=rec).copy()
        steps.append(st)
    return hdr, steps


def _read_dump_bsph(r, hdr, nw):
    """DB-SPH layout (harness_dump_io.f90 with Domain%tipo = "bsph"): the wall
    elements in the header; per step the DB-SPH particle and wall state, the
    fluid lists and the fluid-wall lists."""
    hdr["dx_dxw"], hdr["k_w"] = r.f8(2)
    (hdr["slip_id"], hdr["muscl_boundary"], hdr["gamma_limiter"], hdr["negative_wall_p"],
     hdr["fs_allowed"], hdr["n_inlet"], hdr["n_outlet"]) = r.i4(7)
    w = {k: [] for k in ("coord", "normal", "vel", "weight", "volume", "mass", "k_d", "adjacent", "cell")}
    for _ in range(nw):
        v = r.f8(13)
        w["coord"].append(v[0:3])
        w["normal"].append(v[3:6])
        w["vel"].append(v[6:9])
        w["weight"].append(v[9])
        w["volume"].append(v[10])
        w["mass"].append(v[11])
        w["k_d"].append(v[12])
        a = r.i4(4)
        w["adjacent"].append(a[:3])
        w["cell"].append(a[3])
    hdr["walls"] = {k: np.array(v) for k, v in w.items()}
    steps = []
    while not r.eof():
        if r.i4() != 1347704147:
            raise ValueError("bad step tag")
        it, n = r.i4(2)
        t, dt, dtavg = r.f8(3)
        st = dict(it=it, nag=n, time=t, dt=dt, dt_average=dtavg)
        for k in ("coord", "vel", "var", "acc"):
            st[k] = r.f8(3 * n).reshape(n, 3)
        for k in ("dens", "pres", "dden", "vpres", "mass"):
            st[k] = r.f8(n)
        for k in ("cella", "kodvel", "koddens", "laminar_flag"):
            st[k] = r.i4(n)
        for k in ("gamma", "sigma", "rho_sph", "dshep"):
            st[k] = r.f8(n)
        st["drho"] = r.f8(3 * n).reshape(n, 3)
        st["dvel"] = r.f8(9 * n).reshape(n, 3, 3).transpose(0, 2, 1)  # dvel(i,j) at [:, i, j]
        st["fs"] = r.i4(n)
        st["wall_dens"] = r.f8(nw)
        st["wall_pres"] = r.f8(nw)
        st["wall_wet"] = r.i4(nw)
        lists = r.i4()
        if lists:
            cnt = r.i4(n)
            tot = int(cnt.sum())
            st["nbr_count"] = cnt
            st["nbr"] = r.i4(tot) if tot else np.zeros(0, np.int32)
            cw = r.i4(n)
            tw = int(cw.sum())
            st["fw_count"] = cw
            st["fw"] = r.i4(tw) if tw else np.zeros(0, np.int32)
        steps.append(st)
    return hdr, steps
