"""Checks of the oracle that do not go through the reference: brute-force
neighbour sets under the reference's +-1-cell rule, list symmetry and the
lattice count, the kernel's normalisation, Galilean invariance of the viscous
terms, a still-water tank, and the linear EoS round trip.  A dropped term, a
wrong sign or a transposed index in the restatement fails one of them."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from sphera_amd.case import DamBreak3D, build, cell_number


def _cloud_case():
    # a closed box of walls, the particles below replace the lattice
    return DamBreak3D(tank=(0.2, 0.2, 0.2), column=(0.2, 0.2, 0.2), dx=0.01, name="cloud")


def _brute_sets(tb, coord):
    """Neighbour sets from scratch: j is a neighbour of i iff their reference
    grid cells differ by at most one in every index and r^2 <= (2h)^2."""
    g = tb.grid
    c = cell_number(g, coord).astype(np.int64) - 1
    ni, nj = int(g.ncd[0]), int(g.ncd[1])
    ci = np.stack([c % ni, (c // ni) % nj, c // (ni * nj)], axis=1)
    sq = (2.0 * tb.scalars["h"]) ** 2
    d = coord[:, None, :] - coord[None, :, :]
    r2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    near = np.all(np.abs(ci[:, None, :] - ci[None, :, :]) <= 1, axis=2)
    ok = near & (r2 <= sq)
    return [set(np.nonzero(ok[i])[0].tolist()) for i in range(len(coord))]


def test_lists_equal_brute_force_on_random_cloud():
    case = _cloud_case()
    tb = build(case)
    rng = np.random.default_rng(12345)
    n = 1500
    coord = rng.uniform(0.005, 0.195, (n, 3))
    o = Oracle(tb, coord=coord, vel=np.zeros((n, 3)), dens=np.full(n, 1000.0),
               pres=np.zeros(n), mass=np.full(n, 1e-3))
    o.preloop()
    cnt, ids = o.lists()
    brute = _brute_sets(tb, coord)
    for i in range(n):
        got = set(ids[i, :cnt[i]].tolist())
        assert got == brute[i], i
        assert i in got  # self included
    # symmetry j in N(i) <=> i in N(j)
    for i in range(n):
        for j in ids[i, :cnt[i]]:
            assert i in set(ids[j, :cnt[j]].tolist())


def test_interior_lattice_count_is_81():
    tb = build(DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.01))
    o = Oracle(tb)
    o.preloop()
    cnt, _ = o.lists()
    x = tb.particles["coord"]
    two_h = 2.0 * tb.scalars["h"]
    inner = np.all((x > two_h + 1e-9) & (x < np.array([0.12, 0.1, 0.12]) - two_h - 1e-9), axis=1)
    assert inner.sum() > 0
    # 81 lattice points within 2.6 dx, self included (h/dx = 1.3)
    assert np.all(cnt[inner] == 81)


def test_shepard_sum_and_kernel_normalisation():
    """sum_j m_j W_ij / rho_j ~ 1 at interior lattice particles: the cubic
    spline integrates to one (to the reference's 6-digit coefke)."""
    tb = build(DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.01))
    h, coefke = tb.scalars["h"], tb.scalars["coefke"]
    x = tb.particles["coord"]
    dx = tb.case.dx
    i = int(np.argmin(np.sum((x - np.array([0.06, 0.05, 0.06])) ** 2, axis=1)))
    r = np.sqrt(np.sum((x - x[i]) ** 2, axis=1)) / h
    r = r[r < 2.0]
    w = np.where(r < 1.0, 2.0 / 3.0 - r * r + 0.5 * r ** 3, (2.0 - r) ** 3 / 6.0) * coefke
    s = float(np.sum(w) * dx ** 3)
    assert abs(s - 1.0) < 0.01


def test_viscous_terms_vanish_for_uniform_velocity():
    """Monaghan and Morris terms depend on v_j - v_i only: a uniform velocity
    leaves the inner accelerations unchanged (walls off)."""
    tb = build(DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.01))
    n = len(tb.particles["coord"])
    rng = np.random.default_rng(3)
    coord = tb.particles["coord"] + rng.uniform(-0.05, 0.05, (n, 3)) * tb.case.dx
    acc = []
    for u in (0.0, 1.0):
        o = Oracle(tb, coord=coord, vel=np.tile([u, -0.5 * u, 0.25 * u], (n, 1)), semi=False)
        o.preloop()
        o.call("momentum")
        acc.append(o.state()["acc"])
    assert np.allclose(acc[0], acc[1], rtol=0, atol=1e-9 * np.abs(acc[0]).max())


def test_still_water_tank_stays_at_rest():
    """Water filling the tank footprint, hydrostatic: velocities stay small,
    the free surface stays put, mass is exact."""
    case = DamBreak3D(tank=(0.2, 0.1, 0.2), column=(0.2, 0.1, 0.1), dx=0.01, name="still")
    tb = build(case)
    o = Oracle(tb)
    o.preloop()
    o.step(100)
    st = o.state()
    c0 = 10.0 * np.sqrt(9.81 * 0.1)
    assert np.abs(st["vel"]).max() < 0.02 * c0
    assert abs(st["coord"][:, 2].max() - tb.particles["coord"][:, 2].max()) < 0.5 * case.dx
    assert o.n == len(tb.particles["coord"])


def test_eos_round_trip():
    tb = build(DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01))
    c = tb.case
    p = tb.particles["pres"]
    rho = c.rho0 * (1.0 + (p - c.p_ref) / c.bulk)  # inverse EoS (EoS_barotropic_linear.f90:53)
    p2 = c.bulk / c.rho0 * (rho - c.rho0) + c.p_ref
    assert np.allclose(p2, p, rtol=1e-12, atol=1e-9)
    assert np.array_equal(rho, tb.particles["dens"])


@pytest.mark.parametrize("k", [1, 2])
def test_neighbour_truncation_like_reference(k):
    """With a tiny NMAXPARTJ every list stops at the cap (the reference
    truncates, CalcVarLength.f90:200-207), keeping its first entries."""
    tb = build(DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01))
    full = Oracle(tb)
    full.preloop()
    cnt_full, ids_full = full.lists()
    tb.scalars["nmaxpartj"] = 10 * k
    o = Oracle(tb)
    o.preloop()
    cnt, ids = o.lists()
    assert np.all(cnt == np.minimum(cnt_full, 10 * k))
    for i in range(len(cnt)):
        assert np.array_equal(ids[i, :cnt[i]], ids_full[i, :cnt[i]])
    from oracle.pyoracle import lib
    # particles that met a candidate once full (the reference warns for each)
    assert lib().or_overflow(o.ctx) >= int(np.sum(cnt_full > 10 * k)) > 0


def test_lists_only_mode_and_list_comparison():
    """The oracle's lists-only build gives the lists of the full
    CalcVarLength, and the list comparison finds a planted difference (one
    entry, one count) at the first particle it touches."""
    import numpy as np
    from oracle.pyoracle import Oracle
    from sphera_amd.microbench import cloud
    tb, st = cloud(20_000)
    full = Oracle(tb, semi=False, **st)
    full.preloop()
    lo = Oracle(tb, semi=False, **st)
    lo.lists_only()
    c, ids = full.lists()
    c2, ids2 = lo.lists()
    assert np.array_equal(c, c2) and np.array_equal(ids, ids2)
    assert lo.first_list_mismatch(c, ids) == -1
    ids[17, 2] += 1
    assert lo.first_list_mismatch(c, ids) == 17
    ids[17, 2] -= 1
    c[3] -= 1
    assert lo.first_list_mismatch(c, ids) == 3
