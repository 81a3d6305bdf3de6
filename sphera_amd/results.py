"""Output files of the reference: the results file (<case>.ris) and the
Paraview particle file (.vtu).

The results file holds Fortran sequential unformatted records, 4-byte
little-endian length markers around each record.

Memo_Results.f90:279-287 writes, at every results step, a visualisation
record pair (restartcode = 0):

    it (int32), simulation_time, dt (float64), nag, restartcode (int32),
    emission_time (float64)
    coord(1..3), vel(1..3), pres, dens, mass, kin_visc (nag float64 each),
    imed, icol (nag int32 each)

:func:`write_results` writes such pairs from a device state (pg order), so a
device run produces the records the reference writes at the same steps;
:func:`read_results` reads them back (also from a file the reference wrote,
skipping its header records and its full-state restart records, restartcode =
1, which hold the compiler's layout of the particle type and are not
written here).
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

_HEAD = struct.Struct("<iddiid")


def zone_colour(hexrgb):
    """icol of a zone coloured "RRGGBB" in the deck ("uniform 009EA8"): the
    three bytes in the reverse order, as the reference stores them."""
    v = int(hexrgb, 16)
    r, g, b = (v >> 16) & 0xFF, (v >> 8) & 0xFF, v & 0xFF
    return (b << 16) | (g << 8) | r


def _record(fh, payload):
    n = len(payload)
    fh.write(struct.pack("<i", n))
    fh.write(payload)
    fh.write(struct.pack("<i", n))


def write_results(fh, it, time, dt, state, kin_visc, imed=1, icol=zone_colour("009EA8"),
                  emission_time=0.0):
    """Appends one visualisation record pair to the open binary file fh.
    state: coord (n,3), vel (n,3), pres, dens, mass (n) in pg order; kin_visc,
    imed and icol are per particle or scalars."""
    c = np.asarray(state["coord"], np.float64)
    n = len(c)
    v = np.asarray(state["vel"], np.float64)
    cols = [c[:, 0], c[:, 1], c[:, 2], v[:, 0], v[:, 1], v[:, 2],
            np.asarray(state["pres"], np.float64), np.asarray(state["dens"], np.float64),
            np.asarray(state["mass"], np.float64), np.broadcast_to(np.float64(kin_visc), (n,))]
    f = np.ascontiguousarray(np.stack(cols), dtype="<f8")
    i = np.ascontiguousarray(np.stack([np.broadcast_to(np.int32(imed), (n,)),
                                       np.broadcast_to(np.int32(icol), (n,))]), dtype="<i4")
    _record(fh, _HEAD.pack(int(it), float(time), float(dt), n, 0, float(emission_time)))
    _record(fh, f.tobytes() + i.tobytes())


def records(path):
    """The raw records of a Fortran sequential unformatted file."""
    return records_from_bytes(Path(path).read_bytes())


def records_from_bytes(b):
    """The raw records of Fortran sequential unformatted data."""
    o, out = 0, []
    while o < len(b):
        (n,) = struct.unpack_from("<i", b, o)
        rec = b[o + 4:o + 4 + n]
        (m,) = struct.unpack_from("<i", b, o + 4 + n)
        if m != n:
            raise ValueError("bad record markers")
        out.append(rec)
        o += 8 + n
    return out


def read_results(path):
    """The visualisation record pairs of a results file: a list of dicts with
    it, time, dt, nag and the per-particle fields."""
    recs = records(path)
    out = []
    for k, r in enumerate(recs):
        if len(r) != _HEAD.size or k + 1 >= len(recs):
            continue
        it, t, dt, nag, code, emis = _HEAD.unpack(r)
        if code != 0 or len(recs[k + 1]) != nag * (10 * 8 + 2 * 4):
            continue
        d = recs[k + 1]
        f = np.frombuffer(d, "<f8", 10 * nag).reshape(10, nag)
        i = np.frombuffer(d, "<i4", 2 * nag, offset=80 * nag).reshape(2, nag)
        out.append(dict(it=it, time=t, dt=dt, nag=nag, emission_time=emis,
                        coord=f[0:3].T.copy(), vel=f[3:6].T.copy(), pres=f[6].copy(),
                        dens=f[7].copy(), mass=f[8].copy(), kin_visc=f[9].copy(),
                        imed=i[0].copy(), icol=i[1].copy()))
    return out


# ---------------------------------------------------------------------------
# Paraview file of the particles (result_converter.f90:85-265 and :595-604,
# the fields a Domain%tipo = "semi" case without C1_BE, ALE3 or KTGF writes)
# ---------------------------------------------------------------------------
def _es(v):
    """Fortran ES12.4E3: one digit, point, four digits, E, sign, three digits,
    right-aligned in 12 characters."""
    v = float(v)
    if v == 0.0:
        s = "0.0000E+000"
    else:
        m, e = f"{v:.4E}".split("E")
        s = f"{m}E{'-' if int(e) < 0 else '+'}{abs(int(e)):03d}"
    return s.rjust(12)


def _rows(vals, per, fmt, lead):
    out = []
    for i in range(0, len(vals), per):
        out.append(" " * 8 + lead(vals[i:i + per], fmt))
    return out


def write_vtu(fh, state, kin_visc, laminar, imed=1):
    """Writes the particle file (VTKConverter_<case>_block_<k>.vtu) the
    reference's result_converter writes for a semi case: state holds coord
    (n,3), vel (n,3), pres, dens, mass in pg order (every particle inside the
    grid), laminar the particles' laminar flag.  fh is a text file."""
    c = np.asarray(state["coord"], np.float64)
    n = len(c)
    v = np.asarray(state["vel"], np.float64)
    L = []
    w = L.append
    w('<?xml version="1.0"?>')
    w('<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" '
      'compressor="vtkZLibDataCompressor">')
    w('<UnstructuredGrid>')
    w(f'  <Piece NumberOfPoints="{n}"        NumberOfCells="1"      >')
    w('    <Points>')
    w('      <DataArray type="Float32" NumberOfComponents="3" format="ascii" >')
    for i in range(0, n, 6):  # '(6(3(ES12.4E3,1x)))', adjustl, len_trim
        w(" " * 8 + "".join(_es(x) + " " for x in c[i:i + 6].ravel()).strip())
    w('      </DataArray>')
    w('    </Points>')
    w('    <Cells>')
    w('      <DataArray type="Int32" Name="connectivity" format="ascii" >')
    for i in range(0, n, 30):  # k = i .. min(i + 29, n): the last row reaches n
        w(" " * 8 + "".join(f"{k:8d}" for k in range(i, min(i + 29, n) + 1)).strip())
    w('      </DataArray>')
    w('      <DataArray type="Int32" Name="offsets" format="ascii" >')
    w(" " * 8 + f"{n:8d}".rstrip())
    w('      </DataArray>')
    w('      <DataArray type="UInt8" Name="types" format="ascii" >')
    w(" " * 8 + "     2")
    w('      </DataArray>')
    w('    </Cells>')
    w('<PointData>')

    def vec(name, a):
        w(f'      <DataArray type="Float32" Name="{name}"  NumberOfComponents="3"  format="ascii" >')
        for i in range(0, n, 6):
            w(" " * 8 + "".join(" " + _es(x) for x in a[i:i + 6].ravel()))
        w('      </DataArray>')

    def sca(name, a, per=16, fmt=None, typ="Float32"):
        w(f'      <DataArray type="{typ}" Name="{name}" format="ascii" >')
        for i in range(0, n, per):
            w(" " * 8 + "".join(" " + (fmt(x) if fmt else _es(x)) for x in a[i:i + per]))
        w('      </DataArray>')

    vec("Velocity vectors", v)
    sca("Pressure (Pa)", np.asarray(state["pres"], np.float64))
    sca("Density (kg/mc)", np.asarray(state["dens"], np.float64))
    sca("Mass (kg)", np.asarray(state["mass"], np.float64))
    sca("Viscosity (m^2/s)", np.broadcast_to(np.float64(kin_visc), (n,)))
    sca("z(m)", c[:, 2])
    i6 = lambda x: f"{int(x):6d}"  # noqa: E731
    sca("Medium", np.broadcast_to(np.int32(imed), (n,)), 24, i6)
    sca("Status", np.ones(n, np.int32), 24, i6)  # "flu" particles of a liquid
    sca("Finger", np.arange(1, n + 1), 24, lambda x: f"{int(x):8d}")
    sca("laminar_flag", np.asarray(laminar), 16, lambda x: f"{int(x):8d}", "Int32")
    w('    </PointData>')
    w('    <CellData>')
    w('    </CellData>')
    w('  </Piece>')
    w('</UnstructuredGrid>')
    w('</VTKFile>')
    fh.write("\n".join(L) + "\n")
