"""Served request stream on one GPU: p50/p99 latency and throughput of the serving loop.

A pool of seeded boxes (workloads.served_pool) is registered with an od_server;
Poisson arrivals (workloads.served_requests) each advance one object 1..8 Newmark
steps with an impulse on its first step.  Latency is finish - arrival (queueing
included); throughput counts Newmark steps over the span from the first arrival
to the last finish.  Before the streams every object advances one step, 128
objects per batch, so that the objects the budget holds are resident (--no-warm
starts with nothing resident).

    python -m opendeformer_amd.serving --objects 1024 --tets 1000 500000 \\
        --rates 50,200,1000,2000 --duration 60 --budget-gb 200
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np

from . import capi, workloads as W


def build_server(pool, budget_bytes, slots=4, precision=capi.OD_MIXED, tol=1e-6):
    from .core import Server
    srv = Server(budget_bytes=budget_bytes, slots=slots, precision=precision, tol=tol)
    for w in pool:
        srv.add_object(w.X, w.tets, w.fixed_dofs, w.kind, w.E, w.nu, w.rho, w.body, w.dt)
    return srv


def run_stream(srv, pool, rate, duration, seed=6):
    obj, steps, arrival, impulses = W.served_requests(pool, rate, duration, seed=seed)
    out = srv.run(obj, steps, arrival, impulses)
    lat = out["finish"] - arrival
    wait = out["start"] - arrival
    span = float(out["finish"].max() - arrival.min()) if len(arrival) else 0.0
    return {
        "rate_req_s": rate, "requests": int(len(obj)), "duration_s": duration,
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                       "max": round(float(lat.max()) * 1e3, 3)},
        "queue_wait_ms_p50": round(float(np.percentile(wait, 50)) * 1e3, 3),
        "latency_per_step_ms_p50": round(float(np.percentile(lat / steps, 50)) * 1e3, 3),
        "steps_per_s": round(float(steps.sum()) / span, 2) if span > 0 else None,
        "pcg_iters_per_step": round(float(out["iters"].sum()) / float(steps.sum()), 2),
        "server": out["stats"],
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--objects", type=int, default=1024)
    ap.add_argument("--tets", type=float, nargs=2, default=[1000, 500000])
    ap.add_argument("--rates", default="50,200,1000,2000")
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--no-warm", action="store_true",
                    help="start the streams with nothing resident (default: every object first advances one step, "
                         "in batches, so the objects that fit the budget are resident before the timed streams)")
    ap.add_argument("--durations", default=None,
                    help="per-rate stream durations in s (comma-separated; default --duration for every rate)")
    args = ap.parse_args()
    t0 = time.perf_counter()
    pool = W.served_pool(args.objects, seed=args.seed, tet_lo=args.tets[0], tet_hi=args.tets[1])
    srv = build_server(pool, int(args.budget_gb * 2 ** 30), slots=args.slots)
    print(json.dumps({"pool_objects": len(pool), "pool_tets": int(sum(w.n_tets for w in pool)),
                      "pool_dof": int(sum(3 * w.X.shape[0] - len(w.fixed_dofs) for w in pool)),
                      "build_s": round(time.perf_counter() - t0, 1)}), flush=True)
    if not args.no_warm:
        for b0 in range(0, len(pool), 128):
            ids = np.arange(b0, min(b0 + 128, len(pool)), dtype=np.int32)
            t1 = time.perf_counter()
            out = srv.run(ids, np.ones(ids.size, np.int32), np.zeros(ids.size))
            print(json.dumps({"warm_objects": [int(ids[0]), int(ids[-1])], "s": round(time.perf_counter() - t1, 2),
                              "server": out["stats"]}), flush=True)
    rates = [float(r) for r in args.rates.split(",")]
    durs = [float(d) for d in args.durations.split(",")] if args.durations else [args.duration] * len(rates)
    for i, (rate, dur) in enumerate(zip(rates, durs)):
        print(json.dumps(run_stream(srv, pool, rate, dur, seed=args.seed + 1 + i)), flush=True)


if __name__ == "__main__":
    main()
