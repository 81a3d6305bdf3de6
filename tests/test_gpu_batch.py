"""Segmented batches on the device: several dam breaks of uneven sizes, each
with its own dx, grid and tables, concatenated in one context and stepped by
one launch per kernel.  Every segment must end bit for bit in the state the
same solve reaches alone (same arithmetic, same list order), in both
precisions, eagerly and by graph replay."""
import numpy as np
import pytest

from sphera_amd.case import DamBreak3D, build

pytestmark = pytest.mark.gpu

FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")


def _cases():
    # uneven particle counts and dx (hence h, grid, NMAXPARTJ and SA-SPH tables)
    return [
        build(DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="b0")),
        build(DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.2, 0.15), dx=0.012, name="b1")),
        build(DamBreak3D(tank=(0.35, 0.15, 0.25), column=(0.1, 0.15, 0.1), dx=0.0085, name="b2")),
    ]


def _jitter(tb, seed):
    rng = np.random.default_rng(seed)
    c = tb.particles["coord"]
    return c + rng.uniform(-0.01, 0.01, c.shape) * tb.case.dx


def _alone(tb, prec, coord, steps):
    from sphera_amd import runtime as rt
    d = rt.SphContext(tb, precision=prec)
    try:
        d.upload(coord=coord)
        d.preloop()
        d.step(steps)
        return d.state(), d.pair_counter()
    finally:
        d.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_segments_equal_alone(precision, graph):
    from sphera_amd import runtime as rt
    prec = rt.FP64 if precision == "fp64" else rt.FP32_MIXED
    tbs = _cases()
    coords = [_jitter(tb, 7 + k) for k, tb in enumerate(tbs)]
    steps = 12
    print(f"[gpu] batch {precision} graph={graph} n={[len(c) for c in coords]}", flush=True)
    b = rt.SphBatch(tbs, precision=prec)
    try:
        b.upload(coords)
        b.preloop()
        if graph:
            for _ in range(steps):
                b.step_graph(1)
            b.sync()
        else:
            b.step(steps)
        got = [b.seg_state(k) for k in range(len(tbs))]
        pairs = b.pair_counter()
    finally:
        b.close()
    total = 0
    for k, tb in enumerate(tbs):
        ref, p = _alone(tb, prec, coords[k], steps)
        total += p
        assert got[k]["nag"] == ref["nag"]
        assert np.array_equal(got[k]["ids"], ref["ids"])
        for f in FIELDS:
            assert np.array_equal(got[k][f], ref[f]), (k, f)
        for s in ("dt", "dt_average", "time", "it", "pairs"):
            assert got[k][s] == ref[s], (k, s)
    assert pairs == total


def test_batch_of_one_is_the_plain_context():
    """A batch of one segment is laid out as a one-solve context."""
    from sphera_amd import runtime as rt
    tb = _cases()[1]
    coord = _jitter(tb, 3)
    b = rt.SphBatch([tb], precision=rt.FP32_MIXED)
    try:
        b.upload([coord])
        b.preloop()
        b.step(5)
        got = b.seg_state(0)
    finally:
        b.close()
    ref, _ = _alone(tb, rt.FP32_MIXED, coord, 5)
    for f in FIELDS:
        assert np.array_equal(got[f], ref[f]), f


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_batch_segment_removal_equals_alone(precision):
    """A segment whose initial state has particles outside its grid (the
    reference's removal fixture) loses them as in OrdGrid1 (swap with the
    last, per segment) while the other segments are untouched: every segment
    still equals its solve alone, ids dense."""
    from pathlib import Path

    from sphera_amd import runtime as rt
    prec = rt.FP64 if precision == "fp64" else rt.FP32_MIXED
    g = np.load(Path(__file__).resolve().parent / "golden" / "ref_removal.npz")
    tiny = build(DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="tiny"))
    other = _cases()[1]
    ic = g["ic_coord"]
    # the removal case in the middle, so that segments before and after it are checked
    tbs = [other, tiny, other]
    coords = [_jitter(other, 11), ic, _jitter(other, 12)]
    print(f"[gpu] batch removal {precision} n={[len(c) for c in coords]}", flush=True)

    assert len(ic) == len(tiny.particles["coord"])  # the fixture moves particles of this case
    b = rt.SphBatch(tbs, precision=prec)
    try:
        b.upload(coords)
        b.preloop()
        b.step(6)
        got = [b.seg_state(k) for k in range(3)]
    finally:
        b.close()
    for k, (tb, c) in enumerate(zip(tbs, coords)):
        ref, _ = _alone(tb, prec, c, 6)
        assert got[k]["nag"] == ref["nag"]
        assert np.array_equal(got[k]["ids"], np.arange(ref["nag"]))
        for f in FIELDS:
            assert np.array_equal(got[k][f], ref[f]), (k, f)
    assert got[1]["nag"] == len(ic) - 6
