"""Ensemble server on the device: a request served among others (own context
and stream, overlapping kernels, admission and release around it) ends in
exactly the state it reaches when run alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("coord", "vel", "acc", "dens", "pres", "dden", "vpres")


@pytest.mark.parametrize("precision", ["fp32_mixed", "fp64"])
def test_mixed_equals_alone(precision):
    from sphera_amd import runtime as rt
    from sphera_amd import serve
    prec = rt.FP32_MIXED if precision == "fp32_mixed" else rt.FP64
    reqs = serve.make_ensemble(6, steps=12, rate=0.0, seed0=100, dx_range=(0.014, 0.02))
    for r in reqs[::2]:
        r.keep_state = True
    print(f"[gpu] serving {precision}: n={[r.n for r in reqs]}", flush=True)
    out = serve.EnsembleServer(precision=prec, max_active=3, chunk=4).run(reqs)
    assert out["requests"] == 6
    for r in reqs[::2]:
        alone = serve.run_alone(r, precision=prec)
        for f in FIELDS:
            assert np.array_equal(r.state[f], alone[f]), (r.rid, f)
    print("[gpu] serving summary:", {k: v for k, v in out.items() if k != "particles"}, flush=True)


@pytest.mark.parametrize("precision", ["fp32_mixed", "fp64"])
def test_batched_serving_equals_alone(precision):
    """Requests served as segments of batches (several batches at once) end
    in exactly the state each reaches alone."""
    from sphera_amd import runtime as rt
    from sphera_amd import serve
    prec = rt.FP32_MIXED if precision == "fp32_mixed" else rt.FP64
    reqs = serve.make_ensemble(7, steps=10, rate=0.0, seed0=200, dx_range=(0.014, 0.02))
    for r in reqs[::2]:
        r.keep_state = True
    print(f"[gpu] batched serving {precision}: n={[r.n for r in reqs]}", flush=True)
    out = serve.BatchServer(precision=prec, batch=3, max_batches=2, chunk=4).run(reqs)
    assert out["requests"] == 7
    for r in reqs[::2]:
        alone = serve.run_alone(r, precision=prec)
        for f in FIELDS:
            assert np.array_equal(r.state[f], alone[f]), (r.rid, f)
