#!/usr/bin/env python3
"""Headline benchmark: pair-interactions/s of the 3D dam break on one MI355X.

Workload: tank 1.6 x 0.6 x 0.6 m, water column 0.4 x 0.6 x 0.3 m, dx = 0.005 m
(576,000 particles), h = 1.3 dx, linear EoS with c0 = 10 sqrt(g H), SA-SPH
walls, lattice jitter U(+-0.01 dx) from seed 20260101, leapfrog.  A step is one
whole time step on the device (dt, momentum, velocity smoothing, trajectories,
cell sort, neighbour lists, SA-SPH boundary table, continuity + EoS, pressure
smoothing), inputs resident in HBM.  value = pair interactions / s of the
FP32-mixed path, sum_steps sum_i (nPartIntorno(i)-1) / t; the plain FP64 path
is timed beside it.

The served ensemble ("ensemble" in the output line): 512 seeded dam-break
requests (seeds 0-511; column sides and height U[0.2, 0.6] m, dx U[0.008,
0.02] m, 2k-324k particles) of 500 FP32-mixed steps each, Poisson arrivals at
16 and 32 requests/s, each served as segmented batches and with one context
per request; p50/p99 time to solution and aggregate pair-interactions/s per
run, the batched form at the lower rate as the headline.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per particle of one pair pass: each cell stages its 27
# neighbour cells once (record size S per particle) plus the per-particle output
REC_BYTES = {  # pass: (S fp64, out fp64, S fp32, out fp32)
    "continuity": (68, 8, 36, 4),
    "pressure_smoothing": (52, 8, 28, 4),
    "momentum": (84, 24, 44, 12),
    "velocity_smoothing": (68, 24, 36, 12),
}
# neighbour search: each cell stages the FP32 positions (float4, 16 B) of its 27
# neighbour cells once; output 4 B per list entry (self included) + 4 B count
NB_STAGE_BYTES = 16


def alg_bytes(kernel, n, pairs_per_step, fp32=True):
    """Algorithmic bytes of one launch of `kernel` on n particles."""
    if kernel in REC_BYTES:
        S, out = REC_BYTES[kernel][2:] if fp32 else REC_BYTES[kernel][:2]
        return n * (27 * S + out)
    if kernel == "neighbours":
        return n * 27 * NB_STAGE_BYTES + 4 * (pairs_per_step + n) + 4 * n
    return None


TRAFFIC_FILE = ROOT / "profiles" / "hbm_traffic_damb3d_576k.md"


def source_hash():
    """Hash of the device sources; a measured traffic figure is used only for
    the sources it was measured on."""
    h = hashlib.sha1()
    for f in ("sph_device.h", "sph_kernels.hip", "sph_api.hip"):
        h.update((ROOT / "sphera_amd" / "csrc" / f).read_bytes())
    return h.hexdigest()[:12]


def traffic_bytes(kernel, n, fp32):
    """HBM bytes per launch of `kernel` measured by the two rocprofv3 counter
    runs (FETCH_SIZE, WRITE_SIZE, scaled by a copy of known size) recorded in
    TRAFFIC_FILE for this workload and these sources, else None."""
    prec = "fp32_mixed" if fp32 else "fp64"
    try:
        lines = TRAFFIC_FILE.read_text().splitlines()
    except OSError:
        return None
    src = source_hash()
    for ln in lines:
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if (len(cells) == 7 and cells[0] == kernel and cells[1] == prec and cells[2] == str(n)
                and cells[6] == src):
            return int(cells[5])
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_case(args):
    from sphera_amd.case import DamBreak3D, build
    case = DamBreak3D(tank=(1.6, 0.6, 0.6), column=(0.4, 0.6, 0.3), dx=args.dx, name="dam_break_3D")
    tb = build(case)
    rng = np.random.default_rng(20260101)
    c = tb.particles["coord"]
    coord = c + rng.uniform(-0.01, 0.01, c.shape) * case.dx
    return case, tb, coord


def run_device(tb, coord, precision, steps, warmup, tune=None, graph=True):
    from sphera_amd import runtime as rt
    d = rt.SphContext(tb, precision=precision, tune=tune)
    d.upload(coord=coord)
    d.preloop()
    if warmup:
        d.step(warmup)
    if graph:  # capture (and instantiate) the graph outside the timed region
        d.step_graph(1)
        d.sync()
    p0 = d.pair_counter()  # synchronises the context stream
    t0 = time.perf_counter()
    if graph:
        for _ in range(steps):  # one graph launch per step
            d.step_graph(1)
        d.sync()
    else:
        d.step(steps)       # enqueues every step, then synchronises
    t1 = time.perf_counter()
    pairs = d.pair_counter() - p0
    # per-kernel durations from HIP events on the context stream, in a run of
    # its own so that the events do not load the timed steps above
    d.set_timing(True)
    d.step(max(5, min(steps, 20)))
    timing = d.timing()
    n = d.n
    sc = d.scalars()
    launch = d.launch()
    d.close()
    return dict(seconds=t1 - t0, pairs=pairs, timing=timing, n=n, it=sc["it"], dt=sc["dt"],
                launch=launch)


def cpu_baseline(tb, coord, max_steps, budget_s):
    """The FP64 OpenMP restatement on the host cores, on a bounded sample: whole
    time steps of the same case until `budget_s` seconds or `max_steps`."""
    from oracle.pyoracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    o = Oracle(tb, coord=coord)
    o.preloop()
    o.step(1)
    # sum over steps of the list pairs built in each step
    pairs = 0
    steps = 0
    t0 = time.perf_counter()
    while steps < max_steps and time.perf_counter() - t0 < budget_s:
        o.step(1)
        pairs += o.pair_count()
        steps += 1
    t1 = time.perf_counter()
    return dict(value=pairs / (t1 - t0), unit="pair-interactions/s", cores=threads,
                kind="port",
                sample=f"oracle FP64 OpenMP restatement, full {len(coord)}-particle case, "
                       f"{steps} whole time steps after 1 untimed ({(t1 - t0):.1f} s)")


def cpu_baseline_reference(case, coord, pairs_per_step, n1=1, n2=25):
    """The reference itself (its flang OpenMP build driven by the dump harness
    in oracle/_ref, same jittered positions): two runs of n1 and n2 time steps,
    the difference of their wall times timing n2 - n1 whole steps without the
    pre-processing both share.  pairs_per_step: the list pairs of one step of
    this case (the reference's lists are the device's).  None if the harness
    is not built."""
    import tempfile
    from oracle.decks import HARNESS, run_harness
    if not HARNESS.exists():
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS") or min(os.cpu_count() or 1, 16))
    times = []
    for n in (n1, n2):
        with tempfile.TemporaryDirectory(prefix="sph_refbase_") as wd:
            t0 = time.perf_counter()
            run_harness(case, n, every=n, lists=False, ic_coord=coord, threads=threads, workdir=wd)
            times.append(time.perf_counter() - t0)
    dt = times[1] - times[0]
    if dt <= 0:
        return None
    return dict(value=pairs_per_step * (n2 - n1) / dt, unit="pair-interactions/s", cores=threads,
                kind="reference",
                sample=f"the reference (flang -O2 OpenMP build, oracle/_ref/harness3d.x), full "
                       f"{len(coord)}-particle case: runs of {n1} and {n2} time steps took "
                       f"{times[0]:.1f} s and {times[1]:.1f} s; {n2 - n1} steps in {dt:.1f} s at the "
                       f"device's {pairs_per_step:.4g} list pairs per step")


def copy_probe(mib):
    """A plain device copy of known size (mib MiB read, mib MiB written), run
    before the benchmark so that rocprofv3 counter runs of this command can
    scale FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--dx", type=float, default=0.005)
    ap.add_argument("--cpu-steps", type=int, default=int(os.environ.get("SPH_CPU_STEPS", "200")),
                    help="at most this many CPU baseline steps")
    ap.add_argument("--cpu-seconds", type=float, default=20.0,
                    help="CPU baseline sample length (whole steps until this many seconds)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ref-cpu", action="store_true",
                    help="time only the oracle's CPU restatement, not the reference's own build")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="launch every kernel of a step instead of replaying the step's HIP graph")
    ap.add_argument("--wg", type=int, default=0,
                    help="fixed workgroup size of the particle kernels (skips the autotuner)")
    ap.add_argument("--chunk", type=int, default=0, help="fixed FP32 list-walk chunk (4 or 8)")
    ap.add_argument("--nb-shape", type=int, default=0,
                    help="fixed neighbour-search launch shape (sph_params.nb_shape), with --wg")
    ap.add_argument("--only", choices=("fp32_mixed", "fp64"), default=None,
                    help="time one precision only (profiling runs)")
    ap.add_argument("--ens-requests", type=int, default=512,
                    help="ensemble: number of requests (0 skips the ensemble)")
    ap.add_argument("--ens-steps", type=int, default=500, help="ensemble: time steps per request")
    ap.add_argument("--ens-rates", default="16,32",
                    help="ensemble: Poisson arrival rates, requests per second (one run of each "
                         "serving form per rate)")
    ap.add_argument("--ens-active", type=int, default=16, help="ensemble: concurrent requests")
    ap.add_argument("--ens-batch", type=int, default=8,
                    help="ensemble: requests per segmented batch (0 skips the batched run)")
    ap.add_argument("--ens-batches", type=int, default=4, help="ensemble: concurrent segmented batches")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a device copy of this many MiB first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU decomposition "
              "(every run uses one MI355X); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    from sphera_amd import runtime as rt
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
    case, tb, coord = make_case(args)
    log(f"[bench] case n={len(coord)} grid={tuple(tb.grid.ncd)} steps={args.steps} warmup={args.warmup}")
    res = {}
    tunes = {}
    for name, prec in (("fp32_mixed", rt.FP32_MIXED), ("fp64", rt.FP64)):
        if args.only and name != args.only:
            continue
        tune = dict(wg_size=args.wg, chunk=args.chunk, nb_shape=args.nb_shape) if args.wg else None
        if not args.no_autotune and not args.wg:
            from sphera_amd.autotune import autotune
            tune, _ = autotune(tb, coord, prec, log=log)
            log(f"[bench] {name}: autotuned launch {tune}")
        tunes[name] = tune
        log(f"[bench] running {name}")
        res[name] = run_device(tb, coord, prec, args.steps, args.warmup, tune, graph=not args.no_graph)
        r = res[name]
        log(f"[bench] {name}: {r['seconds'] / args.steps * 1e3:.3f} ms/step, "
            f"{r['pairs'] / r['seconds'] / 1e9:.3f} G pairs/s")
    main_name = args.only or "fp32_mixed"
    main_r = res[main_name]
    fp32 = main_name == "fp32_mixed"
    value = main_r["pairs"] / main_r["seconds"]
    n = main_r["n"]
    pairs_per_step = main_r["pairs"] / args.steps
    # dominant kernel of the step (by time, among the kernels with a byte
    # model) and its roofline; every such kernel is listed under "kernels"
    per_k = {}
    for k, (ms_tot, calls) in main_r["timing"].items():
        b = alg_bytes(k, n, pairs_per_step, fp32)
        if b is None or calls == 0:
            continue
        avg_s = ms_tot / calls / 1e3
        per_k[k] = dict(avg_launch_ms=avg_s * 1e3, bytes_alg_per_launch=int(b),
                        achieved_gbs=b / avg_s / 1e9, frac=b / avg_s / 1e9 / HBM_PEAK_GBS)
    dom = max(per_k, key=lambda k: per_k[k]["avg_launch_ms"])
    kern = {k: round(v[0] / v[1], 4) for k, v in main_r["timing"].items()}
    # pair kernels only: the four passes' average launch times, one launch each per step
    pair_step_s = sum(v[0] / v[1] for k, v in main_r["timing"].items() if k in REC_BYTES and v[1]) / 1e3
    fp64 = None
    if "fp64" in res and main_name != "fp64":
        r64 = res["fp64"]
        fp64 = {"value": r64["pairs"] / r64["seconds"],
                "ms_per_step": r64["seconds"] / args.steps * 1e3,
                "kernel_ms": {k: round(v[0] / v[1], 4) for k, v in r64["timing"].items()},
                "launch": r64["launch"]}
    ens = None
    if args.ens_requests > 0:
        from sphera_amd import serve
        t_b = time.perf_counter()
        reqs = serve.make_ensemble(args.ens_requests, args.ens_steps, 1.0)
        unit_gaps = np.diff([r.arrival for r in reqs], prepend=0.0)  # exponential(1) gaps
        log(f"[bench] ensemble: {len(reqs)} requests x {args.ens_steps} steps, "
            f"{min(r.n for r in reqs)}-{max(r.n for r in reqs)} particles, built in "
            f"{time.perf_counter() - t_b:.1f} s")
        prec = rt.FP32_MIXED if fp32 else rt.FP64
        runs = []
        for rate in [float(x) for x in args.ens_rates.split(",")]:
            forms = [("segmented batches", lambda: serve.BatchServer(
                precision=prec, batch=args.ens_batch, max_batches=args.ens_batches))] \
                if args.ens_batch > 0 else []
            forms.append(("one context per request",
                          lambda: serve.EnsembleServer(precision=prec, max_active=args.ens_active)))
            for mode, make in forms:
                for r in reqs:
                    r.t_admit = r.t_done = float("nan")
                    r.pairs, r.state, r.ctx, r.done_steps = 0, None, None, 0
                t_arr = np.cumsum(unit_gaps / rate)
                for r, ta in zip(reqs, t_arr):
                    r.arrival = float(ta)
                out = make().run(reqs)
                run = dict(mode=mode, arrival_rate_per_s=rate, p50_latency_s=out["p50_latency_s"],
                           p99_latency_s=out["p99_latency_s"], mean_latency_s=out["mean_latency_s"],
                           p99_queue_s=out["p99_queue_s"], makespan_s=out["makespan_s"],
                           pair_interactions_per_s=out["pair_interactions_per_s"])
                if mode == "segmented batches":
                    run.update(batch=args.ens_batch, max_batches=args.ens_batches)
                else:
                    run.update(max_active=args.ens_active)
                runs.append(run)
                log(f"[bench] ensemble {mode} at {rate}/s: p50 {run['p50_latency_s']:.3f} s, p99 "
                    f"{run['p99_latency_s']:.3f} s, {run['pair_interactions_per_s'] / 1e9:.2f} G pairs/s")
        head = runs[0]  # the served form at the first (lowest) rate
        ens = dict(workload=f"{len(reqs)} seeded dam-break requests (seeds 0-{len(reqs) - 1}) in the "
                            "1.6x0.6x0.6 m tank, column sides and height U[0.2,0.6] m, dx U[0.008,0.02] m, "
                            f"{args.ens_steps} steps each, Poisson arrivals",
                   requests=len(reqs), steps=args.ens_steps,
                   particles_min=min(r.n for r in reqs), particles_max=max(r.n for r in reqs),
                   mode=head["mode"], arrival_rate_per_s=head["arrival_rate_per_s"],
                   p50_latency_s=head["p50_latency_s"], p99_latency_s=head["p99_latency_s"],
                   pair_interactions_per_s=head["pair_interactions_per_s"], runs=runs)
    cpu = cpu_port = None
    if not args.no_cpu and args.cpu_steps > 0:
        log(f"[bench] cpu baseline (up to {args.cpu_steps} steps / {args.cpu_seconds} s)")
        cpu_port = cpu_baseline(tb, coord, args.cpu_steps, args.cpu_seconds)
        log(f"[bench] cpu baseline: {cpu_port['value'] / 1e6:.2f} M pairs/s ({cpu_port['sample']})")
        if not args.no_ref_cpu:
            cpu = cpu_baseline_reference(case, coord, pairs_per_step)
            if cpu:
                log(f"[bench] reference cpu baseline: {cpu['value'] / 1e6:.2f} M pairs/s ({cpu['sample']})")
        if cpu is None:
            cpu, cpu_port = cpu_port, None
    out_line = {
        "metric": "pair-interactions/s mixed vs FP64 (3D dam break); p50/p99 ensemble latency",
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_r["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (fp64 positions/integration)" if fp32 else "f64",
        "data": "synthetic: jittered lattice dam break, seed 20260101",
        "config": {"workload": "3D dam break, tank 1.6x0.6x0.6 m, column 0.4x0.6x0.3 m, "
                               f"dx={case.dx} m, h/dx=1.3, SA-SPH walls", "particles": n,
                   "grid_cells": int(tb.grid.nmax)},
        "fp64": fp64,
        "launch": dict(main_r["launch"], graph=not args.no_graph, tune=tunes.get(main_name)),
        "mixed_over_fp64": value / fp64["value"] if fp64 else None,
        "pair_kernels_only_value": pairs_per_step / pair_step_s if pair_step_s > 0 else None,
        "kernel_ms": kern,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": per_k[dom]["achieved_gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": per_k[dom]["frac"],
                     "traffic": traffic_bytes(dom, n, fp32),
                     "bytes_alg_per_launch": per_k[dom]["bytes_alg_per_launch"],
                     "avg_launch_ms": per_k[dom]["avg_launch_ms"]},
        "kernels": per_k,
        "ensemble": ens,
        "cpu_baseline": cpu,
        "cpu_baseline_port": cpu_port,
    }
    print(json.dumps(out_line), flush=True)


if __name__ == "__main__":
    main()
