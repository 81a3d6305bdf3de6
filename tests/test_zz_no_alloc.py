"""sph_global_quantities_noalloc (the entry point that allocates no device
memory) against sph_global_quantities, through the C ABI: bit-identical
results on two calls in a row, the second on a smaller solve."""
import numpy as np
import pytest

from sphera_amd.case import DamBreak3D, build

pytestmark = pytest.mark.gpu


def _ctx(case, seed):
    from sphera_amd import runtime as rt
    tb = build(case)
    rng = np.random.default_rng(seed)
    c = tb.particles["coord"]
    coord = c + rng.uniform(-0.01, 0.01, c.shape) * case.dx
    d = rt.SphContext(tb, precision=rt.FP64)
    d.upload(coord=coord)
    d.preloop()
    d.step(3)
    return d


def test_noalloc_global_quantities_match():
    big = _ctx(DamBreak3D(tank=(0.4, 0.2, 0.3), column=(0.12, 0.1, 0.12), dx=0.005, name="big"), 1)
    small = _ctx(DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01, name="small"), 2)
    assert big.n > 4 * small.n
    print(f"[gpu] noalloc n={big.n} then n={small.n}", flush=True)
    want = [big.global_quantities(), small.global_quantities()]
    got = [big.global_quantities(noalloc=True), small.global_quantities(noalloc=True)]
    for w, g in zip(want, got):
        for k in ("mass", "lin_mom", "ang_mom", "volume"):
            assert np.array_equal(np.asarray(w[k]), np.asarray(g[k])), k
    # and once more on the larger solve after the smaller one
    again = big.global_quantities(noalloc=True)
    for k in ("mass", "lin_mom", "ang_mom", "volume"):
        assert np.array_equal(np.asarray(want[0][k]), np.asarray(again[k])), k
    big.close()
    small.close()
