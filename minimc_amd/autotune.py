"""Autotuner for the tracker's launch parameters.

The event-based tracker has two knobs whose best values depend on the
problem (history length distribution, table sizes) and on the device:

  tail_threshold  live-particle count below which the history-based tail
                  kernel finishes the batch (-1: never)
  graph_iters     event iterations launched between two host reads of the
                  live-queue length
  compact_interval  bank compaction policy (0: when half empty, k: every k
                  iterations, -1: never)
  graphs          captured batches (k-eigenvalue)
  split_collide   separate advance and collision kernels (k-eigenvalue)
  tail_block, tsl_block  threads per block of the tail kernel and of the POD
                  thermal-scattering kernel
and, for the surrogate tracker (tune_tail_network),
  tail_network    live count down to which the tail parks thermal
                  collisions for the network (0: POD in the tail)

`tune()` times short runs of the actual problem for each candidate
(coordinate descent: the tail threshold, the readback interval, the
compaction policy, captured batches, then the block sizes),
and returns the fastest option set (a later candidate must beat the best so
far by MARGIN).  Results are cached per problem in
minimc_amd/_build/autotune.json (that folder is built per machine).

python -m minimc_amd.autotune [--config c5] [--cycles 3]
"""
import hashlib
import json
import os
import time

from . import abi

TAIL_CANDIDATES = (-1, 4096, 16384, 32768, 65536, 131072, 262144)
ITER_CANDIDATES = (4, 8, 16, 32)
COMPACT_CANDIDATES = (-1, 1, 2, 4)  # 0 (compact when half empty) is the start
# threads per block beside the default 256: the tail kernel (two waves per
# SIMD) and the POD thermal-scattering kernel (three waves per SIMD)
TAIL_BLOCK_CANDIDATES = (512,)
TSL_BLOCK_CANDIDATES = (384,)
# surrogate mode: live count down to which the tail runs in rounds that park
# thermal collisions for the network (0: the tail samples them with POD)
TAIL_NETWORK_CANDIDATES = (0, 16384, 4096)
# A candidate replaces the best so far only if it is this much faster: run to
# run timing noise is about 1 %, and near-ties keep the earlier candidate.
MARGIN = 0.015
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "autotune.json")


def _digest(h, obj):
    """Feeds a problem's plain data (dicts, lists, numbers, arrays) to h."""
    import numpy as np
    if isinstance(obj, dict):
        for k in sorted(obj, key=str):
            h.update(str(k).encode())
            _digest(h, obj[k])
    elif isinstance(obj, (list, tuple)):
        for v in obj:
            _digest(h, v)
    elif isinstance(obj, np.ndarray):
        h.update(np.ascontiguousarray(obj).tobytes())
    else:
        h.update(repr(obj).encode())


def _signature(problem, tsl_mode, n):
    h = hashlib.sha1()
    for part in (problem.surfaces, problem.fields, problem.cells, problem.materials,
                 problem.nuclides, problem.tsl, problem.estimators, problem.source):
        _digest(h, part)
    _digest(h, (problem.problem, problem.tracking, problem.seed, n, tsl_mode))
    return h.hexdigest()[:16]


def _time_run(problem, opts, n, cycles, tsl_mode, prepare):
    from .tracker import Tracker
    t = Tracker(problem, tsl_mode=tsl_mode, max_in_flight=n, **opts)
    try:
        if prepare:
            prepare(t)
        if problem.problem == "keigenvalue":
            t.set_entropy_mesh(*problem.entropy_mesh)
            t.keig_init(n)
            t.run_cycle(active=False)  # untimed: first-touch and source spread
            t0 = time.perf_counter()
            for _ in range(cycles):
                t.run_cycle(active=False)
            return (time.perf_counter() - t0) / cycles
        t.run_fixed_source(n, first=0)
        t0 = time.perf_counter()
        for c in range(cycles):
            t.run_fixed_source(n, first=(c + 1) * n)
        return (time.perf_counter() - t0) / cycles
    finally:
        t.close()


def tune(problem, n=None, cycles=3, tsl_mode=abi.MMC_TSL_POD, prepare=None,
         use_cache=True, log=None):
    """Returns ({"tail_threshold": .., "graph_iters": .., ...}, table of timings)."""
    n = n or problem.histories
    key = _signature(problem, tsl_mode, n)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        with open(CACHE) as f:
            cache = json.load(f)
        if key in cache:
            return cache[key]["best"], cache[key]["table"]
    table = []
    best, best_s = {"tail_threshold": 0, "graph_iters": 8}, float("inf")
    for tail in TAIL_CANDIDATES:
        opts = {"tail_threshold": tail, "graph_iters": 8}
        s = _time_run(problem, opts, n, cycles, tsl_mode, prepare)
        table.append((dict(opts), s))
        if log:
            log("tail %6d iters %2d: %.2f ms" % (tail, 8, s * 1e3))
        if s < best_s * (1 - MARGIN):
            best, best_s = dict(opts), s
    for it in ITER_CANDIDATES:
        if it == best["graph_iters"]:
            continue
        opts = dict(best, graph_iters=it)
        s = _time_run(problem, opts, n, cycles, tsl_mode, prepare)
        table.append((dict(opts), s))
        if log:
            log("tail %6d iters %2d: %.2f ms" % (opts["tail_threshold"], it, s * 1e3))
        if s < best_s * (1 - MARGIN):
            best, best_s = dict(opts), s
    for cp in COMPACT_CANDIDATES:
        opts = dict(best, compact_interval=cp)
        s = _time_run(problem, opts, n, cycles, tsl_mode, prepare)
        table.append((dict(opts), s))
        if log:
            log("tail %6d iters %2d compact %2d: %.2f ms" % (
                opts["tail_threshold"], opts["graph_iters"], cp, s * 1e3))
        if s < best_s * (1 - MARGIN):
            best, best_s = dict(opts), s
    # batches replayed as captured graphs (k-eigenvalue; compaction every
    # iteration, at readbacks or never)
    if problem.problem == "keigenvalue" and best.get("compact_interval", 0) <= 1:
        opts = dict(best, graphs=1)
        s = _time_run(problem, opts, n, cycles, tsl_mode, prepare)
        table.append((dict(opts), s))
        if log:
            log("tail %6d iters %2d compact %2d graphs: %.2f ms" % (
                opts["tail_threshold"], opts["graph_iters"],
                opts.get("compact_interval", 0), s * 1e3))
        if s < best_s * (1 - MARGIN):
            best, best_s = dict(opts), s
    # separate advance and collision kernels (k-eigenvalue: by default one
    # kernel does the flight and the collision that ends it)
    if problem.problem == "keigenvalue":
        opts = dict(best, split_collide=1)
        s = _time_run(problem, opts, n, cycles, tsl_mode, prepare)
        table.append((dict(opts), s))
        if log:
            log("split collide: %.2f ms" % (s * 1e3))
        if s < best_s * (1 - MARGIN):
            best, best_s = dict(opts), s
    # block sizes: the tail kernel (when the tail runs) and the POD
    # thermal-scattering kernel (when the problem has a thermal-scattering law)
    blocks = []
    if best.get("tail_threshold", 0) >= 0:
        blocks += [("tail_block", b) for b in TAIL_BLOCK_CANDIDATES]
    if problem.tsl and tsl_mode == abi.MMC_TSL_POD:
        blocks += [("tsl_block", b) for b in TSL_BLOCK_CANDIDATES]
    for name, b in blocks:
        opts = dict(best, **{name: b})
        s = _time_run(problem, opts, n, cycles, tsl_mode, prepare)
        table.append((dict(opts), s))
        if log:
            log("%s %d: %.2f ms" % (name, b, s * 1e3))
        if s < best_s * (1 - MARGIN):
            best, best_s = dict(opts), s
    if use_cache:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        cache[key] = {"best": best, "table": table}
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    return best, table


def tune_tail_network(problem, base, n=None, cycles=3, prepare=None, use_cache=True,
                      log=None):
    """For the surrogate tracker with launch options `base` (from tune()):
    the tail_network value, from TAIL_NETWORK_CANDIDATES, that runs fastest
    (prepare(tracker) must set the network).  Returns (options, table)."""
    n = n or problem.histories
    key = "tail_network:" + _signature(problem, abi.MMC_TSL_SURROGATE, n) + \
        json.dumps(base, sort_keys=True)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        with open(CACHE) as f:
            cache = json.load(f)
        if key in cache:
            return cache[key]["best"], cache[key]["table"]
    table = []
    best, best_s = None, float("inf")
    for tn in TAIL_NETWORK_CANDIDATES:
        opts = dict(base, tail_network=tn)
        s = _time_run(problem, opts, n, cycles, abi.MMC_TSL_SURROGATE, prepare)
        table.append((dict(opts), s))
        if log:
            log("tail_network %6d: %.2f ms" % (tn, s * 1e3))
        if best is None or s < best_s * (1 - MARGIN):
            best, best_s = dict(opts), s
    if use_cache:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        cache[key] = {"best": best, "table": table}
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    return best, table


def main():
    import argparse
    from . import configs
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="c5", choices=("c5", "c2", "c1"))
    ap.add_argument("--cycles", type=int, default=3)
    ap.add_argument("--no-cache", action="store_true")
    args = ap.parse_args()
    pr = {"c5": configs.c5_multizone, "c2": configs.c2_slab, "c1": configs.c1_slab}[args.config]()
    n = pr.histories if args.config != "c2" else 1_000_000
    best, table = tune(pr, n=n, cycles=args.cycles, use_cache=not args.no_cache,
                       log=lambda m: print(m, flush=True))
    print(json.dumps({"config": args.config, "best": best,
                      "table": [[o, s * 1e3] for o, s in table]}))


if __name__ == "__main__":
    main()
