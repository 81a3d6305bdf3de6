"""Host set-up (sphera_amd/case.py) against the reference's own set-up tables,
recorded in tests/golden/ref_tiny.npz by the flang build: grid, scalars,
SA-SPH faces, grid-cell/face incidence, kernel and direction tables, and the
initial lattice with its hydrostatic pressure and density.  All bit-exact."""
from pathlib import Path

import numpy as np

from sphera_amd.case import DamBreak3D, build, cell_number, iwro2dro, jdwsrn

G = np.load(Path(__file__).resolve().parent / "golden" / "ref_tiny.npz")


def _tb():
    return build(DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny"))


def test_scalars_and_grid():
    tb = _tb()
    s = tb.scalars
    assert s["nmaxpartj"] == int(G["hdr_nmaxpartj"])
    for k in ("h", "coefke", "coefkacl", "eta", "eta2", "celerita", "kin_visc"):
        assert s[k] == float(G["hdr_" + k]), k
    assert np.array_equal(tb.grid.ncd, G["hdr_ncd"])
    assert np.array_equal(tb.grid.extr_lo, G["hdr_extr_lo"])
    assert np.array_equal(tb.grid.dcd, G["hdr_dcd"])


def test_faces_and_incidence():
    tb = _tb()
    F = tb.faces
    assert np.array_equal(F.face_id, G["hdr_face_id"])
    assert np.array_equal(F.T, G["hdr_face_T"])
    assert np.array_equal(F.area, G["hdr_face_area"])
    assert np.array_equal(F.LX, G["hdr_face_LX"])
    assert np.array_equal(tb.gcbf_count, G["hdr_gcbf_count"])
    ref_start, ref_vec = G["hdr_gcbf_start"], G["hdr_gcbf_vector"]
    for c in np.nonzero(tb.gcbf_count)[0]:
        n = tb.gcbf_count[c]
        mine = F.face_id[tb.gcbf_vector[tb.gcbf_start[c] - 1:tb.gcbf_start[c] - 1 + n]]
        assert np.array_equal(mine, ref_vec[ref_start[c] - 1:ref_start[c] - 1 + n]), c
    # every normal points into the tank: the water column's centre is on the
    # positive side of each frontier face
    centre = np.array(tb.case.column) * 0.5
    for f in range(len(F.face_id)):
        assert np.dot(F.T[f, :, 2], centre - F.GX[f, 3]) > 0


def test_tables():
    tb = _tb()
    assert np.array_equal(tb.kerneltab, G["hdr_kerneltab"])
    assert tb.ktdelta == float(G["hdr_ktdelta"])
    assert np.array_equal(tb.bitab, G["hdr_bitab"])


def test_initial_particles():
    tb = _tb()
    P = tb.particles
    assert np.array_equal(P["coord"], G["s0_coord"])
    assert np.array_equal(P["pres"], G["s0_pres"])
    assert np.array_equal(P["dens"], G["s0_dens"])


def test_cell_number_closed_form():
    tb = _tb()
    g = tb.grid
    # the centre of cell (i,j,k) maps to ((k-1)nj + (j-1))ni + i; outside -> 0
    for (i, j, k) in [(1, 1, 1), (3, 2, 5), tuple(int(v) for v in g.ncd)]:
        x = g.extr_lo + (np.array([i, j, k]) - 0.5) * g.dcd
        assert cell_number(g, x)[0] == ((k - 1) * g.ncd[1] + (j - 1)) * g.ncd[0] + i
    assert cell_number(g, g.extr_lo - 1.0)[0] == 0


def test_kernel_integrals_closed_forms():
    """IWro2dro(0) is the normalisation of the cubic spline, 1/(4 pi) per
    steradian (so the kernel integrates to 1); JdWsRn vanishes at q = 2 and
    equals the moments int_q0^2 dW/dq q^n dq (scipy quadrature of the
    textbook spline)."""
    from scipy.integrate import quad
    assert abs(iwro2dro(0.0) - 1.0 / (4.0 * np.pi)) < 1e-14
    for n in (1, 2, 3):
        assert abs(jdwsrn(2.0, n)) < 1e-16

    def dwdq(q):  # d/dq of the 3D cubic spline times pi (Monaghan & Lattanzio 1985)
        if q < 1.0:
            return -3.0 * q + 2.25 * q * q
        if q < 2.0:
            return -0.75 * (2.0 - q) ** 2
        return 0.0

    for q0 in (0.0, 0.3, 1.0, 1.7):
        for n in (1, 2, 3):
            val, _ = quad(lambda q: dwdq(q) * q ** n, q0, 2.0, points=[1.0])
            # JdWsRn is the integral from q0 to 2 times 1/pi (0.31830989, the
            # reference's 8-digit constant)
            assert abs(jdwsrn(q0, n) - val * 0.31830989) < 1e-12, (q0, n)
