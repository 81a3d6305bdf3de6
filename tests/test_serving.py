"""Ensemble server host logic (sphera_amd/serve.py) with a stand-in device:
FIFO admission in arrival order, the active-request cap, admission by free
device memory with release of finished contexts, continuous batching (a
waiting request starts while others still run) and the latency summary."""
import numpy as np
import pytest

from sphera_amd import serve
from sphera_amd.case import DamBreak3D


class FakeCtx:
    """Runs `k` enqueued steps in `k` polls; memory comes from a shared pool."""
    pool = None
    log = []

    def __init__(self, tb, precision=None):
        self.tb = tb
        self.need = FakeCtx.bytes_for(tb)
        FakeCtx.pool["free"] -= self.need
        self.pending = 0
        self.steps = 0
        FakeCtx.log.append(("start", tb.rid))

    @staticmethod
    def bytes_for(tb):
        return tb.n * 1000

    def upload(self):
        pass

    def preloop(self):
        pass

    def busy(self):
        if self.pending:
            self.pending -= 1
            self.steps += 1
            return True
        return False

    def step_async(self, k):
        self.pending += k

    def step_graph(self, k):
        self.pending += k

    def sync(self):
        pass

    def pair_counter(self):
        return self.steps * self.tb.n

    def state(self):
        return {}

    def close(self):
        FakeCtx.pool["free"] += self.need
        FakeCtx.log.append(("close", self.tb.rid))


class FakeTb:
    def __init__(self, rid, n):
        self.rid, self.n = rid, n


@pytest.fixture
def fake_device(monkeypatch):
    FakeCtx.pool = {"free": 10_000_000, "total": 10_000_000}
    FakeCtx.log = []
    monkeypatch.setattr(serve.rt, "SphContext", FakeCtx)
    monkeypatch.setattr(serve.rt, "ctx_bytes", lambda tb, prec: FakeCtx.bytes_for(tb))
    monkeypatch.setattr(serve.rt, "device_memory", lambda: (FakeCtx.pool["free"], FakeCtx.pool["total"]))
    return FakeCtx


def _reqs(ns, steps=10):
    case = DamBreak3D(tank=(0.3, 0.1, 0.2), column=(0.08, 0.1, 0.1), dx=0.01)
    return [serve.Request(rid=k, case=case, steps=steps, arrival=0.0, tb=FakeTb(k, n), n=n)
            for k, n in enumerate(ns)]


def test_fifo_cap_and_completion(fake_device):
    reqs = _reqs([1000, 2000, 3000, 4000, 5000])
    out = serve.EnsembleServer(max_active=2, chunk=3, reserve_frac=0.0).run(reqs)
    assert out["requests"] == 5
    starts = [rid for ev, rid in fake_device.log if ev == "start"]
    assert starts == [0, 1, 2, 3, 4]  # FIFO
    assert out["pair_interactions"] == sum(10 * n for n in (1000, 2000, 3000, 4000, 5000))
    assert out["p50_latency_s"] <= out["p99_latency_s"]
    assert all(r.t_admit <= r.t_done for r in reqs)


def test_continuous_batching(fake_device):
    """A short request leaves and the next one joins while a long one runs."""
    reqs = _reqs([1000, 1000, 1000])
    reqs[0].steps, reqs[1].steps, reqs[2].steps = 3, 60, 3
    serve.EnsembleServer(max_active=2, chunk=2, reserve_frac=0.0).run(reqs)
    assert reqs[0].t_done < reqs[2].t_admit < reqs[1].t_done


def test_admission_waits_for_memory_and_releases_parked(fake_device):
    # 10 MB device, 4 MB requests: two fit at a time; a third needs a release
    fake_device.pool = {"free": 10_000_000, "total": 10_000_000}
    reqs = _reqs([4000, 4000, 4000, 4000])
    out = serve.EnsembleServer(max_active=8, chunk=5, reserve_frac=0.0).run(reqs)
    assert out["requests"] == 4
    log = fake_device.log
    # request 2 starts only after a finished context was closed
    i_start2 = log.index(("start", 2))
    assert any(ev == "close" for ev, _ in log[:i_start2])
    assert fake_device.pool["free"] == 10_000_000  # everything released


def test_request_larger_than_device_fails(fake_device):
    fake_device.pool = {"free": 1_000_000, "total": 1_000_000}
    with pytest.raises(MemoryError):
        serve.EnsembleServer(max_active=2).run(_reqs([5000]))


def test_ensemble_cases_span_sizes():
    reqs = serve.make_ensemble(8, steps=1, rate=100.0, seed0=0)
    ns = [r.n for r in reqs]
    assert min(ns) > 0 and max(ns) > 4 * min(ns)
    arr = [r.arrival for r in reqs]
    assert arr[0] == 0.0 and all(np.diff(arr) >= 0)


class FakeBatch(FakeCtx):
    """A batch context over several requests' cases (one pool charge)."""

    def __init__(self, tbs, precision=None):
        self.tbs = list(tbs)
        self.need = sum(FakeCtx.bytes_for(tb) for tb in self.tbs)
        FakeCtx.pool["free"] -= self.need
        self.pending = 0
        self.steps = 0
        FakeCtx.log.append(("batch", tuple(tb.rid for tb in self.tbs)))

    def pair_counter(self):
        return self.steps * sum(tb.n for tb in self.tbs)

    def seg_state(self, k):
        return {"rid": self.tbs[k].rid}

    def close(self):
        FakeCtx.pool["free"] += self.need
        FakeCtx.log.append(("close", tuple(tb.rid for tb in self.tbs)))


@pytest.fixture
def fake_batches(fake_device, monkeypatch):
    monkeypatch.setattr(serve.rt, "SphBatch", FakeBatch)
    monkeypatch.setattr(serve.rt, "batch_bytes",
                        lambda tbs, prec: sum(FakeCtx.bytes_for(tb) for tb in tbs))
    return fake_device


def test_batches_group_fifo_and_equal_steps(fake_batches):
    reqs = _reqs([1000, 2000, 3000, 4000, 5000])
    reqs[2].steps = 7  # a different step count: a batch of its own, in its FIFO turn
    for r in reqs:
        r.keep_state = True
    out = serve.BatchServer(batch=2, max_batches=1, chunk=3, reserve_frac=0.0).run(reqs)
    assert out["requests"] == 5
    batches = [ids for ev, ids in fake_batches.log if ev == "batch"]
    assert batches == [(0, 1), (2,), (3, 4)]
    assert out["pair_interactions"] == sum(r.steps * r.n for r in reqs)
    assert all(r.state == {"rid": r.rid} for r in reqs)
    assert fake_batches.pool["free"] == 10_000_000


def test_batch_shrinks_to_fit_memory(fake_batches):
    # 10 MB device; 3 x 4 MB requests do not fit one batch, two do; the rest
    # waits (FIFO) until the first batch is released
    reqs = _reqs([4000, 4000, 4000, 1000])
    serve.BatchServer(batch=3, max_batches=2, chunk=5, reserve_frac=0.0).run(reqs)
    batches = [ids for ev, ids in fake_batches.log if ev == "batch"]
    assert batches == [(0, 1), (2, 3)]
    assert reqs[2].t_admit >= reqs[0].t_done
    assert fake_batches.pool["free"] == 10_000_000


def test_batches_run_at_once(fake_batches):
    """A later batch starts while an earlier one still runs (continuous
    batching across batches)."""
    reqs = _reqs([1000, 1000, 1000])
    reqs[0].steps = 60
    reqs[1].steps = reqs[2].steps = 3
    serve.BatchServer(batch=2, max_batches=2, chunk=2, reserve_frac=0.0).run(reqs)
    batches = [ids for ev, ids in fake_batches.log if ev == "batch"]
    assert batches == [(0,), (1, 2)]
    assert reqs[1].t_admit < reqs[0].t_done and reqs[1].t_done < reqs[0].t_done
