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


@pytest.mark.timeout(600)
def test_ensemble_requests_served_as_one_batch_match_oracle():
    """Four requests of the 512-request ensemble's own draw (seeds 249, 28,
    281 and 60: its largest, 323,680 particles, the next largest, 253,440, a
    median one and its smallest), served together as one segmented batch for
    20 steps: each request's FP64 state equals the oracle's bit for bit, and
    each FP32-mixed state is within 1e-4 of it (largest per-particle
    difference relative to the field's largest |value|)."""
    from oracle.pyoracle import Oracle
    from sphera_amd import runtime as rt
    from sphera_amd import serve
    from sphera_amd.case import build
    seeds = (249, 28, 281, 60)
    steps = 20
    refs = {}
    for s in seeds:
        tb = build(serve.ensemble_case(s))
        o = Oracle(tb)
        o.preloop()
        o.step(steps)
        refs[s] = o.state()
        del o
    for prec in (rt.FP64, rt.FP32_MIXED):
        reqs = []
        for s in seeds:
            case = serve.ensemble_case(s)
            tb = build(case)
            reqs.append(serve.Request(rid=s, case=case, steps=steps, arrival=0.0, tb=tb,
                                      n=len(tb.particles["coord"]), keep_state=True))
        print(f"[gpu] ensemble batch precision={prec} n={[r.n for r in reqs]}", flush=True)
        out = serve.BatchServer(precision=prec, batch=len(seeds), max_batches=1, chunk=5).run(reqs)
        assert out["requests"] == len(seeds)
        for r in reqs:
            ref = refs[r.rid]
            for f in FIELDS:
                if prec == rt.FP64:
                    assert np.array_equal(r.state[f], ref[f]), (r.rid, f)
                else:
                    e = float(np.abs(r.state[f] - ref[f]).max() / max(np.abs(ref[f]).max(), 1e-300))
                    assert e <= 1e-4, (r.rid, f, e)
