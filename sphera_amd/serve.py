"""Ensemble serving: many independent dam-break solves on one GPU.

The reference runs one case per process (src/Main_algorithm/sphera.f90).  An
ensemble here is a stream of requests, each a whole solve of its own case for
a given number of time steps.  The server keeps one device context per
admitted request, each with its own HIP stream, so the kernels of different
requests overlap on the GPU:

* admission is FIFO in arrival order, by free device memory: a request joins
  when the bytes its context needs (sph_ctx_bytes) fit the free HBM above a
  reserve, and at most `max_active` requests run at once;
* continuous batching: every active request gets `chunk` more steps enqueued
  (sph_step_async, or with graphs=True replays of the request's captured
  one-step HIP graph, sph_step_graph) as soon as its stream is idle
  (sph_query), and a finished
  request leaves while the others keep running, so waiting requests join
  without a batch boundary;
* contexts of finished requests are released when memory is needed or at the
  end (releasing device memory synchronises the device, which would stall the
  streams of the running requests).

With batch > 1 (BatchServer) admitted requests are instead grouped into
segmented batches: up to `batch` waiting requests of equal step count join one
batch context (sph_batch_create), whose every kernel is launched once over all
of its segments, and whose step is one replayed HIP graph.  Several batches
run at once on their own streams, and a batch is formed from the requests
waiting whenever a batch slot and memory are free, so requests still join
while others run.

Latency of a request = completion time - arrival time (time to solution);
the run reports p50/p99 latency and the aggregate pair-interactions/s.
"""
from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import runtime as rt
from .case import DamBreak3D, build


@dataclass
class Request:
    rid: int
    case: DamBreak3D
    steps: int
    arrival: float  # seconds after the start of the run
    keep_state: bool = False
    tb: object = None
    n: int = 0
    t_admit: float = float("nan")
    t_done: float = float("nan")
    pairs: int = 0
    state: dict | None = None
    ctx: object = field(default=None, repr=False)
    done_steps: int = 0


ENSEMBLE_TANK = (1.6, 0.6, 0.6)


def ensemble_case(seed, size_range=(0.2, 0.6), dx_range=(0.008, 0.02), tank=ENSEMBLE_TANK):
    """Dam break of request `seed` in the fixed 1.6 x 0.6 x 0.6 m tank: water
    column width, depth and height ~U(size_range) m each (clipped to the
    tank), dx ~U(dx_range) m; from 1k particles (0.2 m cube at dx = 0.02 m)
    to 420k (0.6 m cube at dx = 0.008 m)."""
    r = np.random.default_rng(seed)
    W, D, H = r.uniform(*size_range, 3)
    dx = float(r.uniform(*dx_range))
    col = (min(W, tank[0]), min(D, tank[1]), min(H, tank[2]))
    return DamBreak3D(tank=tuple(tank), column=col, dx=dx, name=f"req{seed}")


def make_ensemble(count, steps, rate, seed0=0, arrival_seed=12345, **kw):
    """`count` requests with seeds seed0.., Poisson arrivals of `rate` per second
    (rate <= 0: all at t=0).  Case tables are built here, before the run."""
    gaps = np.random.default_rng(arrival_seed).exponential(1.0 / rate, count) if rate > 0 else np.zeros(count)
    t = np.cumsum(gaps) - (gaps[0] if rate > 0 else 0.0)
    reqs = []
    for k in range(count):
        case = ensemble_case(seed0 + k, **kw)
        tb = build(case)
        reqs.append(Request(rid=seed0 + k, case=case, steps=steps, arrival=float(t[k]), tb=tb,
                            n=len(tb.particles["coord"])))
    return reqs


class EnsembleServer:
    def __init__(self, precision=rt.FP32_MIXED, max_active=16, chunk=25, reserve_frac=0.05,
                 graphs=False):
        self.precision = precision
        self.graphs = graphs
        self.max_active = max_active
        self.chunk = chunk
        self.reserve_frac = reserve_frac

    def _start(self, r):
        d = rt.SphContext(r.tb, precision=self.precision)
        d.upload()
        d.preloop()
        r.ctx = d
        r.done_steps = 0

    def _finish(self, r, now):
        d = r.ctx
        d.sync()  # checks the device status flags of this request
        r.pairs = d.pair_counter()
        if r.keep_state:
            r.state = d.state()
        r.t_done = now()

    def run(self, requests):
        """Serves `requests`; returns a summary dict (latencies in seconds)."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        active, finished, parked = [], [], []
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        free, total = rt.device_memory()
        while pending or waiting or active:
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())
            # admission
            while waiting and len(active) < self.max_active:
                r = waiting[0]
                need = rt.ctx_bytes(r.tb, self.precision)
                free, total = rt.device_memory()
                if free - need < self.reserve_frac * total:
                    if parked:  # release finished contexts, then retry
                        for q in parked:
                            q.ctx.close()
                            q.ctx = None
                        parked.clear()
                        continue
                    if not active:
                        raise MemoryError(f"request {r.rid} needs {need} B, {free} B free")
                    break
                waiting.popleft()
                r.t_admit = now()
                self._start(r)
                active.append(r)
            # progress of the running requests
            progressed = False
            for r in list(active):
                if r.ctx.busy():
                    continue
                progressed = True
                if r.done_steps < r.steps:
                    k = min(self.chunk, r.steps - r.done_steps)
                    if self.graphs:  # the one-step graph, captured once per request
                        for _ in range(k):
                            r.ctx.step_graph(1)
                    else:
                        r.ctx.step_async(k)
                    r.done_steps += k
                else:
                    self._finish(r, now)
                    active.remove(r)
                    finished.append(r)
                    parked.append(r)
            if not progressed:
                time.sleep(20e-6)
        makespan = now()
        for r in parked:
            r.ctx.close()
            r.ctx = None
        lat = np.array([r.t_done - r.arrival for r in finished])
        wait = np.array([r.t_admit - r.arrival for r in finished])
        pairs = sum(r.pairs for r in finished)
        return {
            "requests": len(finished),
            "makespan_s": makespan,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "p50_queue_s": float(np.percentile(wait, 50)),
            "p99_queue_s": float(np.percentile(wait, 99)),
            "pair_interactions": int(pairs),
            "pair_interactions_per_s": pairs / makespan if makespan > 0 else 0.0,
            "particles": [int(r.n) for r in finished],
        }


def run_alone(req, precision=rt.FP32_MIXED):
    """One request on an otherwise idle device: its state after its steps."""
    d = rt.SphContext(req.tb, precision=precision)
    try:
        d.upload()
        d.preloop()
        d.step(req.steps)
        return d.state()
    finally:
        d.close()


@dataclass
class _Batch:
    reqs: list
    ctx: object
    steps: int
    done_steps: int = 0


class BatchServer:
    """Serves requests in segmented batches (one context per batch of up to
    `batch` requests); at most `max_batches` batches run at once."""

    def __init__(self, precision=rt.FP32_MIXED, batch=8, max_batches=4, chunk=25, reserve_frac=0.05,
                 graphs=True):
        self.precision = precision
        self.batch = batch
        self.max_batches = max_batches
        self.chunk = chunk
        self.reserve_frac = reserve_frac
        self.graphs = graphs

    def _form(self, waiting, parked):
        """Takes the next batch from the head of `waiting`: up to `batch`
        requests with the head's step count.  If it does not fit the free
        memory, finished batches are released first, then the batch shrinks.
        Returns the list, or [] when not even one request fits."""
        head = waiting[0]
        group = [r for r in list(waiting)[: 4 * self.batch] if r.steps == head.steps][: self.batch]

        def fits(g):
            free, total = rt.device_memory()
            return free - rt.batch_bytes([r.tb for r in g], self.precision) >= self.reserve_frac * total

        if not fits(group) and parked:
            for b in parked:
                b.ctx.close()
            parked.clear()
        while group and not fits(group):
            group = group[:-1]
        return group

    def run(self, requests):
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting = deque()
        active, finished, parked = [], [], []
        pairs_total = 0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        while pending or waiting or active:
            t = now()
            while pending and pending[0].arrival <= t:
                waiting.append(pending.popleft())
            while waiting and len(active) < self.max_batches:
                group = self._form(waiting, parked)
                if not group:
                    if not active:
                        raise MemoryError(f"request {waiting[0].rid} does not fit the free device memory")
                    break
                for r in group:
                    waiting.remove(r)
                    r.t_admit = now()
                ctx = rt.SphBatch([r.tb for r in group], precision=self.precision)
                ctx.upload()
                ctx.preloop()
                active.append(_Batch(group, ctx, group[0].steps))
            progressed = False
            for b in list(active):
                if b.ctx.busy():
                    continue
                progressed = True
                if b.done_steps < b.steps:
                    k = min(self.chunk, b.steps - b.done_steps)
                    if self.graphs:
                        for _ in range(k):
                            b.ctx.step_graph(1)
                    else:
                        b.ctx.step_async(k)
                    b.done_steps += k
                else:
                    b.ctx.sync()
                    pairs_total += b.ctx.pair_counter()
                    tf = now()
                    for k, r in enumerate(b.reqs):
                        if r.keep_state:
                            r.state = b.ctx.seg_state(k)
                        r.t_done = tf
                        finished.append(r)
                    active.remove(b)
                    parked.append(b)
            if not progressed:
                time.sleep(20e-6)
        makespan = now()
        for b in parked:
            b.ctx.close()
        lat = np.array([r.t_done - r.arrival for r in finished])
        wait = np.array([r.t_admit - r.arrival for r in finished])
        return {
            "requests": len(finished),
            "batch": self.batch,
            "max_batches": self.max_batches,
            "makespan_s": makespan,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "p50_queue_s": float(np.percentile(wait, 50)),
            "p99_queue_s": float(np.percentile(wait, 99)),
            "pair_interactions": int(pairs_total),
            "pair_interactions_per_s": pairs_total / makespan if makespan > 0 else 0.0,
            "particles": [int(r.n) for r in finished],
        }


def _reset(reqs):
    for r in reqs:
        r.t_admit = r.t_done = float("nan")
        r.pairs, r.state, r.ctx, r.done_steps = 0, None, None, 0


def main(argv=None):
    """Ensemble sweep: the same seeded requests served at each arrival rate,
    in segmented batches and/or one context per request; one JSON line per
    run on stdout, progress on stderr."""
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--rates", default="4,8,16",
                    help="Poisson arrival rates, requests per second (0: all at once)")
    ap.add_argument("--modes", default="batched,contexts")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--max-batches", type=int, default=4)
    ap.add_argument("--max-active", type=int, default=16)
    ap.add_argument("--precision", choices=("fp32_mixed", "fp64"), default="fp32_mixed")
    a = ap.parse_args(argv)
    prec = rt.FP32_MIXED if a.precision == "fp32_mixed" else rt.FP64
    t = time.perf_counter()
    base = make_ensemble(a.requests, a.steps, 1.0)
    print(f"[serve] {a.requests} cases built in {time.perf_counter() - t:.1f} s, "
          f"{min(r.n for r in base)}-{max(r.n for r in base)} particles", file=sys.stderr, flush=True)
    unit_gaps = np.diff([r.arrival for r in base], prepend=0.0)  # exponential(1) gaps
    for rate in [float(x) for x in a.rates.split(",")]:
        for mode in a.modes.split(","):
            _reset(base)
            t_arr = np.cumsum(unit_gaps / rate) if rate > 0 else np.zeros(len(base))
            for r, ta in zip(base, t_arr):
                r.arrival = float(ta)
            if mode == "batched":
                srv = BatchServer(precision=prec, batch=a.batch, max_batches=a.max_batches)
            else:
                srv = EnsembleServer(precision=prec, max_active=a.max_active)
            out = srv.run(base)
            ps = out.pop("particles")
            line = dict(mode=mode, rate_per_s=rate, steps=a.steps, precision=a.precision,
                        particles_min=min(ps), particles_max=max(ps), **out)
            print(f"[serve] {mode} rate {rate}/s: p50 {out['p50_latency_s']:.2f} s, p99 "
                  f"{out['p99_latency_s']:.2f} s, {out['pair_interactions_per_s'] / 1e9:.2f} G pairs/s",
                  file=sys.stderr, flush=True)
            print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
