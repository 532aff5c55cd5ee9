"""Benchmark of the MI355X particle-history tracker.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config NAME] [--no-cpu]

One step is one pass of the hot path over one batch of synthetic input:
  c2  continuous-energy fixed-source slab: one batch of 1e6 histories
      (the full workload is 1e7 histories = 10 steps)
Prints one JSON line (rank 0).  The timed region starts with all inputs in
HBM and is bracketed by device synchronisations.  Kernel durations come from
HIP events recorded by the library on its own stream.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s

# Algorithmic bytes per event of each kernel (particle state read/written
# once plus the table entries strictly needed; search probes and queue
# compaction traffic excluded).
#   advance: read pos 24 + dir 24 + E 8 + rng 4 + cell 4 + steps 4 + slot 4,
#            write pos 24 + rng 4 + steps 4 + slot 4 = 108
#            + cross-section bracketing points (per table: 2 keys + 2 values
#            = 32) + surface coefficients (8 per plane, 32 per quadric)
#   collide: read pos 24 + dir 24 + E 8 + rng 4 + cell 4 + hist 4 + slot 4,
#            write dir 24 + E 8 + rng 4 + slot 4 = 112 + tables
#   cross:   read pos 24 + dir 24 + E 8 + surf 4 + cell 4 + hist 4 + slot 4,
#            write cell 4 + slot 4 + tally record 16 = 96
BYTES = {
    "c2": {"advance": 108 + 32 + 2 * 8,        # total xs + 2 planes
           "collide": 112 + 3 * 32,            # total, capture, elastic
           "cross": 96,
           "tsl": 0},
}


def cpu_baseline(cfg, threads, target_s=12.0):
    """The oracle (CPU restatement, `threads` workers) on a bounded sample."""
    from oracle.oracle import Oracle
    if cfg == "c2":
        from minimc_amd import configs
        pr = configs.c2_slab()
        o = Oracle(pr)
        n = 2000
        t0 = time.perf_counter()
        o.run_fixed_source(n, threads=threads)
        dt = time.perf_counter() - t0
        n2 = int(max(2000, min(2_000_000, n * target_s / max(dt, 1e-3))))
        o.reset_tallies()
        t0 = time.perf_counter()
        o.run_fixed_source(n2, seed0=pr.seed, first=n, threads=threads)
        dt = time.perf_counter() - t0
        return {"value": n2 / dt, "unit": "histories/s", "cores": threads,
                "kind": "port",
                "sample": "%d histories of the 1e7-history CE slab, %.1f s" % (n2, dt)}
    raise ValueError(cfg)


def run_c2(args):
    from minimc_amd import configs
    from minimc_amd.tracker import Tracker
    pr = configs.c2_slab()
    batch = 1_000_000
    t = Tracker(pr, max_in_flight=batch)
    # warmup batches use histories beyond the timed ones
    for w in range(args.warmup):
        t.run_fixed_source(batch, first=(args.steps + w) * batch)
    t.reset_tallies()
    c0 = t.counters()
    t.set_profiling(True)
    t0 = time.perf_counter()
    for s in range(args.steps):
        t.run_fixed_source(batch, first=s * batch)
    dt = time.perf_counter() - t0
    kt = t.kernel_times()
    c1 = t.counters()
    ev = {k: c1[k] - c0[k] for k in ("advances", "collisions", "crossings",
                                     "tsl_collisions", "histories")}
    n = args.steps * batch
    leak = sum(t.tallies(e)[0].sum() for e in range(2))
    p = leak / n
    rel_var = (1 - p) / (p * n) if 0 < p < 1 else float("nan")
    fom = 1.0 / (rel_var * dt)
    counts = {"advance": ev["advances"], "collide": ev["collisions"],
              "cross": ev["crossings"], "tsl": ev["tsl_collisions"]}
    dom = max(kt, key=lambda k: kt[k][0])
    bytes_dom = counts[dom] * BYTES["c2"][dom]
    ms_dom, launches = kt[dom]
    achieved = bytes_dom / (ms_dom * 1e-3) / 1e9
    out = {
        "metric": "FOM 1/(sigma^2 T) + histories/s, surrogate-TSL vs POD-TSL tracker, 1 MI355X",
        "value": n / dt, "unit": "histories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "c2: continuous-energy fixed-source slab, 60k-point "
                   "synthetic XS, 2 MeV point source, 1e6 histories per step",
                   "histories_per_step": batch},
        "fom": fom, "fom_tally": "total leakage", "leakage_probability": p,
        "collisions_per_s": ev["collisions"] / dt,
        "events": counts,
        "kernel_ms": {k: v[0] for k, v in kt.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_event": BYTES["c2"][dom],
                     "avg_launch_ms": ms_dom / max(launches, 1)},
    }
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one MI355X; "
              "the multi-GPU design (history sharding + one all-reduce of FP64 "
              "tallies) is described in DESIGN.md, not run.", file=sys.stderr)
        sys.exit(2)
    rank = int(os.environ.get("RANK", "0"))
    out = run_c2(args)
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(args.config, threads)
    else:
        out["cpu_baseline"] = None
    if rank == 0:
        print(json.dumps(out))


if __name__ == "__main__":
    main()
