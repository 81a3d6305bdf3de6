"""Random-cloud microbenchmark set-up: at 200k particles the device's
neighbour lists equal the oracle's (order and truncation), and one step's
fields match within the precision's bar."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_cloud_lists_match_oracle():
    from oracle.pyoracle import Oracle
    from sphera_amd import runtime as rt
    from sphera_amd.microbench import cloud
    n = 200_000
    tb, st = cloud(n)
    print(f"[gpu] cloud n={n} grid={tuple(tb.grid.ncd)}", flush=True)
    o = Oracle(tb, coord=st["coord"], vel=st["vel"], dens=st["dens"], pres=st["pres"], mass=st["mass"],
               semi=False)
    o.preloop()
    d = rt.SphContext(tb, precision=rt.FP64, capacity=n, semi=False)
    d.upload(**st)
    d.preloop()
    dc, di = d.lists()
    oc, oi = o.lists()
    assert np.array_equal(dc, oc)
    assert np.array_equal(di, oi)
    assert 40 < dc.mean() < 80
    d.step(1)
    o.step(1)
    ds, os_ = d.state(), o.state()
    for f in ("coord", "vel", "dens", "pres"):
        assert np.array_equal(ds[f], os_[f]), f
    d.close()


@pytest.mark.timeout(900)
@pytest.mark.parametrize("n", [2_097_152, 16_777_216])
def test_cloud_lists_match_oracle_large(n):
    """The uniform cloud of the microbenchmark at 2,097,152 particles and at
    its full 16,777,216: the device's neighbour lists (FP64 and FP32-mixed
    contexts, the default search shape) equal the oracle's, entry for entry
    in the reference's order (the oracle builds lists only at these sizes)."""
    from oracle.pyoracle import Oracle
    from sphera_amd import runtime as rt
    from sphera_amd.microbench import cloud
    tb, st = cloud(n)
    print(f"[gpu] cloud n={n} grid={tuple(int(v) for v in tb.grid.ncd)}", flush=True)
    o = Oracle(tb, semi=False, **st)
    o.lists_only()
    for prec in (rt.FP64, rt.FP32_MIXED):
        d = rt.SphContext(tb, precision=prec, capacity=n, semi=False)
        d.upload(**st)
        d.preloop()
        dc, di = d.lists()
        d.close()
        assert 40 < dc.mean() < 80
        assert o.first_list_mismatch(dc, di) == -1, prec
