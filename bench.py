"""Benchmark of the MI355X particle-history tracker.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c5|c2] [--no-cpu]
                [--train-epochs E] [--tail-threshold N --graph-iters K]

One step is one pass of the hot path over one batch of synthetic input.

  c5 (default, the headline workload) multi-zone continuous-temperature
     k-eigenvalue problem: 60 cm box of 27 cubes (fuel, bound-hydrogen
     moderator, structure), per-cube temperature 300-1200 K, 1e6 histories
     per cycle.  One step = one active cycle; the warmup steps are the
     inactive cycles (defaults: 20 inactive + 100 active, the full problem).
     The problem is run twice: with the POD thermal-scattering sampler and
     with the learned sampler (BF16-MFMA network, POD outside its box).  The
     figure of merit of each run is 1/(sigma_rel(k)^2 T) over the active
     cycles; `value` is the surrogate run's histories/s.
  c2 continuous-energy fixed-source slab: one step = one batch of 1e6
     histories (the full workload is 1e7 histories = 10 steps).

Prints one JSON line (rank 0).  The timed region starts with all inputs in
HBM and is bracketed by device synchronisations.  Kernel durations come from
HIP events recorded by the library on its own stream.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E, GB/s
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 MFMA, TFLOP/s

# Algorithmic bytes per event of each kernel (particle state read/written
# once plus the table entries strictly needed; search probes and queue
# compaction traffic excluded).
#   advance: read pos 24 + dir 24 + E 8 + rng 4 + cell 4 + steps 4 + slot 4,
#            write pos 24 + rng 4 + steps 4 + slot 4 = 108
#            + cross-section bracketing points (per table: 2 keys + 2 values
#            = 32) + surface coefficients (8 per plane, 32 per quadric)
#   collide: read pos 24 + dir 24 + E 8 + rng 4 + cell 4 + hist 4 + slot 4,
#            write dir 24 + E 8 + rng 4 + slot 4 = 112 + tables
#   cross:   a surface crossing runs inside the advance kernel on the
#            particle's registers: write cell 4, + read hist 4 + tally
#            record 16 where the surface has current estimators (c2) = 24
#            per crossing, added to the advance kernel's bytes
#   tsl:     read dir 24 + E 8 + cell 4 + rng 4 + slot 4, write dir 24 + E 8
#            + rng 4 + slot 4 = 84 + four rank-8 POD reconstructions (two
#            bracketing CDF rows each for beta and alpha; 8 coefficients +
#            8 temperature-basis values = 128 B each)
#   mlp:     FLOP per sample of the 4-256-256-256-2 forward pass
BYTES = {
    "c2": {"advance": 108 + 32 + 2 * 8,        # total xs + 2 planes
           "collide": 112 + 3 * 32,            # total, capture, elastic
           "collide_fused": 36 + 3 * 32,
           "cross": 24, "tsl": 0},
    "c5": {"advance": 108 + 32 + 6 * 8,        # total xs + 6 cube faces
           "collide": 112 + 3 * 32,            # total, reaction split, elastic
           "collide_fused": 36 + 3 * 32,
           "cross": 4,                         # no current estimators
           "tsl": 84 + 4 * 128},
}
# collide_fused: a collision done by the advance kernel right after the flight
# (k-eigenvalue runs): its table reads and the 36 B of direction, energy and
# random state it writes; the 76 B of particle state a separate collision
# kernel reads are the advance's own writes, not counted twice.
MLP_FLOP_PER_SAMPLE = 2 * (4 * 256 + 256 * 256 + 256 * 256 + 256 * 2)


def host_cpu():
    """The host's processor as `nproc` and `lscpu` report it."""
    info = {"nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket",
                             "Thread(s) per core", "CPU(s)", "CPU max MHz"):
                info["lscpu " + k.strip()] = v.strip()
    except (OSError, ValueError):
        pass
    return info


def cpu_baseline(cfg, threads, target_s=12.0, single_s=6.0):
    """The oracle (CPU restatement) on a bounded sample of the same workload:
    `threads` workers (the reference's scheme, FixedSource.cpp:22-77), then
    one worker for the single-thread rate."""
    multi = cpu_rate(cfg, threads, target_s)
    single = cpu_rate(cfg, 1, single_s)
    multi["single_thread"] = {k: single[k] for k in ("value", "unit", "sample")}
    multi["host"] = host_cpu()
    return multi


def cpu_rate(cfg, threads, target_s):
    from oracle.oracle import Oracle
    from minimc_amd import configs
    if cfg == "c2":
        pr = configs.c2_slab()
        o = Oracle(pr)
        n = 200 * threads
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
    if cfg == "c5":
        # POD sampler (the reference's method); 2 untimed inactive cycles, then
        # active cycles until about target_s of work is timed
        pr = configs.c5_multizone()
        n = 100_000 if threads > 4 else 10_000
        o = Oracle(pr)
        o.set_entropy_mesh(*pr.entropy_mesh)
        o.keig_init(n)
        for _ in range(2 if threads > 4 else 1):
            o.run_cycle(active=False, threads=threads)
        cycles, hist, dt = 0, 0, 0.0
        while dt < target_s and cycles < 200:
            t0 = time.perf_counter()
            st = o.run_cycle(active=True, threads=threads)
            dt += time.perf_counter() - t0
            hist += n
            cycles += 1
        return {"value": hist / dt, "unit": "histories/s", "cores": threads,
                "kind": "port", "k_collision_last": st["k_collision"],
                "sample": "%d active cycles of %d histories of the multi-zone problem "
                          "(POD sampler), %.1f s" % (cycles, n, dt)}
    raise ValueError(cfg)


def _profiled_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary of `workload` (profiles/<workload>_hbm_traffic.md), or None."""
    path = os.path.join(ROOT, "profiles", "%s_hbm_traffic.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 3 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel:
            return float(f[2])
    return None


def _tail_pod(ev):
    """Thermal collisions the tail kernel sampled with POD itself.  In the
    POD tracker, all of the tail's; in the surrogate tracker every thermal
    collision is either a network row, a network fallback (POD outside the
    tail kernel) or one of these (the tail's parked collisions are network
    rows)."""
    if ev["mlp_rows"] == 0:
        return ev["tail_tsl"]
    return ev["tsl_collisions"] - ev["mlp_rows"] - ev["tsl_fallbacks"]


def _tail_roofline(cfg, kt, ev):
    """The history-based tail kernel against the HBM roofline: the
    algorithmic bytes of the events it ran (the event kernels' per-event
    bytes; a thermal collision it parks for the network costs the tail
    kernel only its collision) over its HIP-event time."""
    if "tail" not in kt or kt["tail"][1] == 0:
        return None
    ms, launches = kt["tail"]
    b = BYTES[cfg]
    nbytes = (ev["tail_advances"] * b["advance"] + ev["tail_crossings"] * b["cross"] +
              ev["tail_collisions"] * b["collide"] + _tail_pod(ev) * b["tsl"])
    achieved = nbytes / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "tail", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "events": dict({k: ev["tail_" + k] for k in ("advances", "crossings",
                                                         "collisions", "tsl")},
                           tsl_pod=_tail_pod(ev)),
            "algorithmic_bytes_per_launch": nbytes / max(launches, 1),
            "avg_launch_ms": ms / max(launches, 1)}


def _roofline(cfg, kt, counts):
    """Roofline of the event kernel with the largest total time (the
    history-based tail kernel is reported beside it: _tail_roofline)."""
    dom = max((k for k in kt if k != "tail"), key=lambda k: kt[k][0])
    ms, launches = kt[dom]
    if dom == "mlp":
        achieved = counts["mlp"] * MLP_FLOP_PER_SAMPLE / (ms * 1e-3) / 1e12
        return {"bound": "mfma", "kernel": dom, "achieved": achieved,
                "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None,
                "flop_per_sample": MLP_FLOP_PER_SAMPLE,
                "avg_launch_ms": ms / max(launches, 1)}
    nbytes = counts[dom] * BYTES[cfg][dom]
    if dom == "advance":
        nbytes += counts["cross"] * BYTES[cfg]["cross"]
    # the advance kernel collides its particles itself when no collision
    # kernel ran
    fused = (dom == "advance" and counts.get("collide", 0) > 0 and
             kt.get("collide", (0.0, 0))[1] == 0)
    if fused:
        nbytes += counts["collide"] * BYTES[cfg]["collide_fused"]
    achieved = nbytes / (ms * 1e-3) / 1e9
    traffic = _profiled_traffic(dom, "c5_multizone") if cfg == "c5" else None
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x 2 + WRITE_SIZE, "
                            "profiles/c5_multizone_hbm_traffic.md)" if traffic else None,
            "algorithmic_bytes_per_launch": nbytes / max(launches, 1),
            "bytes_per_event": BYTES[cfg][dom],
            "collisions_in_kernel": fused,
            "bytes_per_collision": BYTES[cfg]["collide_fused"] if fused else None,
            "avg_launch_ms": ms / max(launches, 1)}


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
    counts = {"advance": ev["advances"], "collide": ev["collisions"],
              "cross": ev["crossings"], "tsl": ev["tsl_collisions"], "mlp": 0}
    kt = {k: v for k, v in kt.items() if v[1] > 0}
    return {
        "value": n / dt, "unit": "histories/s", "ms_per_step": dt * 1e3 / args.steps,
        "dtype": "f64",
        "config": {"workload": "c2: continuous-energy fixed-source slab, 60k-point "
                   "synthetic XS, 2 MeV point source, 1e6 histories per step",
                   "histories_per_step": batch},
        "fom": 1.0 / (rel_var * dt), "fom_tally": "total leakage",
        "leakage_probability": p, "collisions_per_s": ev["collisions"] / dt,
        "events": counts, "kernel_ms": {k: v[0] for k, v in kt.items()},
        "roofline": _roofline("c2", kt, counts),
        "tail_roofline": _tail_roofline("c2", kt, {k: c1[k] - c0[k] for k in c1}),
    }


FOM_BATCH = 5  # cycles per batch mean (cycles correlate through the source)


def batch_means(ks, batch=FOM_BATCH):
    nb = len(ks) // batch
    return np.asarray(ks[:nb * batch]).reshape(nb, batch).mean(1)


def fom_of(bm, seconds):
    """1/(sigma_rel^2 T) with sigma from batch means bm of the cycle k."""
    sig = bm.std(ddof=1) / np.sqrt(len(bm))
    return 1.0 / ((sig / bm.mean()) ** 2 * seconds)


def fom_ratio_interval(ks_a, t_a, ks_b, t_b, reps=4000, seed=1):
    """16th, 50th and 84th percentiles of FOM_a / FOM_b under a bootstrap of
    each run's batch means (block bootstrap over batches of cycles)."""
    rng = np.random.default_rng(seed)
    a, b = batch_means(ks_a), batch_means(ks_b)
    if len(a) < 3 or len(b) < 3:
        return None
    r = np.empty(reps)
    for i in range(reps):
        r[i] = fom_of(rng.choice(a, len(a)), t_a) / fom_of(rng.choice(b, len(b)), t_b)
    return [float(v) for v in np.percentile(r, [16, 50, 84])]


def entropy_check(h_inactive, h_active):
    """Source convergence: the Shannon entropy of the last inactive cycles
    against the active cycles' mean and spread, and the drift of the active
    cycles' entropy (least-squares slope per cycle, with its standard error)."""
    ha = np.asarray(h_active)
    tail = np.asarray(h_inactive[-5:]) if len(h_inactive) else ha[:1]
    sd = ha.std(ddof=1) if len(ha) > 1 else 0.0
    x = np.arange(len(ha), dtype=float)
    slope, se = 0.0, 0.0
    if len(ha) > 2:
        A = np.stack([x, np.ones_like(x)], 1)
        coef, res, *_ = np.linalg.lstsq(A, ha, rcond=None)
        slope = float(coef[0])
        resid = ha - A @ coef
        se = float(np.sqrt(resid @ resid / (len(ha) - 2) / ((x - x.mean()) ** 2).sum()))
    gap = float(tail.mean() - ha.mean())
    return {"inactive_last5_mean": float(tail.mean()), "active_mean": float(ha.mean()),
            "active_sd": float(sd), "gap": gap, "active_slope_per_cycle": slope,
            "active_slope_se": se,
            "converged": bool(abs(gap) <= 3 * sd / np.sqrt(len(tail)) + 1e-12 and
                              abs(slope) <= 3 * se + 1e-12)}


def _keig_leg(pr, tracker, warmup, steps):
    """Inactive cycles (untimed), then `steps` timed active cycles."""
    t = tracker
    t.set_entropy_mesh(*pr.entropy_mesh)
    t.keig_init(pr.histories)
    h_inactive = []
    for _ in range(warmup):
        h_inactive.append(t.run_cycle(active=False)["entropy"])
    c0 = t.counters()
    t.set_profiling(True)
    ks, ents = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        st = t.run_cycle(active=True)   # returns after a device synchronisation
        ks.append(st["k_collision"])
        ents.append(st["entropy"])
    dt = time.perf_counter() - t0
    kt = t.kernel_times()
    t.set_profiling(False)
    c1 = t.counters()
    ev = {k: c1[k] - c0[k] for k in c1}
    ks = np.array(ks)
    k = float(ks.mean())
    sig = float(ks.std(ddof=1) / np.sqrt(len(ks))) if len(ks) > 1 else float("nan")
    rel = sig / k
    bm = batch_means(ks)
    return {"seconds": dt, "k": k, "sigma_k": sig, "fom": 1.0 / (rel * rel * dt),
            "sigma_k_batch_means": float(bm.std(ddof=1) / np.sqrt(len(bm)))
            if len(bm) > 1 else None,
            "fom_batch_means": float(fom_of(bm, dt)) if len(bm) > 2 else None,
            "histories_per_s": steps * pr.histories / dt,
            "collisions_per_s": ev["collisions"] / dt,
            "entropy_last": ents[-1], "entropy": entropy_check(h_inactive, ents),
            "k_cycles": [float(v) for v in ks], "events": ev,
            "kernel_ms": {n: v[0] for n, v in kt.items() if v[1] > 0},
            "kernel_launches": {n: v[1] for n, v in kt.items() if v[1] > 0},
            "_kt": kt}


def tsl_batch_rates(tracker, n=1 << 22, reps=5, seed=4):
    """Batch thermal-scattering sampling (2^22 collisions, E log-uniform on
    1e-5..4 eV and T uniform on 300..1200 K): POD sampler vs learned sampler,
    samples/s over `reps` synchronised calls each."""
    rng = np.random.default_rng(seed)
    E = tracker.device_array(np.exp(rng.uniform(np.log(1e-11), np.log(3.99e-6), n)))
    T = tracker.device_array(rng.uniform(300.0, 1200.0, n))
    r = tracker.device_array(rng.integers(1, 2**31 - 1, n).astype(np.uint32))
    a = tracker.device_array(shape=n, dtype=np.float64)
    b = tracker.device_array(shape=n, dtype=np.float64)
    out = {}
    for name, fn in (("pod", tracker.tsl_sample_pod), ("surrogate", tracker.tsl_sample_surrogate)):
        fn(E, T, r, a, b)
        t0 = time.perf_counter()
        for _ in range(reps):
            fn(E, T, r, a, b)
        out[name] = n * reps / (time.perf_counter() - t0)
    for x in (E, T, r, a, b):
        x.free()
    return {"batch": n, "pod_samples_per_s": out["pod"],
            "surrogate_samples_per_s": out["surrogate"],
            "speedup": out["surrogate"] / out["pod"]}


def _progress(msg):
    """A progress line on stderr (stdout carries only the result line)."""
    print("[bench] " + msg, file=sys.stderr, flush=True)


def run_c5(args):
    from minimc_amd import abi, configs, surrogate
    from minimc_amd.tracker import Tracker
    from minimc_amd import autotune
    pr = configs.c5_multizone()
    n = pr.histories
    # launch parameters from the autotuner (short untimed runs of this
    # problem), unless given on the command line
    t0 = time.perf_counter()
    if args.tail_threshold is not None and args.graph_iters is not None:
        opts, table = {"tail_threshold": args.tail_threshold,
                       "graph_iters": args.graph_iters,
                "compact_interval": args.compact_interval}, []
    else:
        opts, table = autotune.tune(pr, n=n, cycles=3)
    t_tune = time.perf_counter() - t0
    _progress("launch options %s (%.1f s)" % (opts, t_tune))
    # plain POD tracker
    tp = Tracker(pr, tsl_mode=abi.MMC_TSL_POD, max_in_flight=n, **opts)
    pod = _keig_leg(pr, tp, args.warmup, args.steps)
    tp.close()
    _progress("POD tracker: %.4g histories/s" % pod["histories_per_s"])
    # surrogate tracker: the shipped network (minimc_amd/data), or one trained
    # here outside every timed region when that is missing
    t0 = time.perf_counter()
    if args.tail_network is not None:
        sopts, stable = dict(opts, tail_network=args.tail_network), []
    else:
        sopts, stable = autotune.tune_tail_network(
            pr, opts, n=n, cycles=3, prepare=lambda t: surrogate.cached_network(t))
    t_tune += time.perf_counter() - t0
    ts = Tracker(pr, tsl_mode=abi.MMC_TSL_SURROGATE, max_in_flight=n, **sopts)
    t0 = time.perf_counter()
    net_path = surrogate.SHIPPED
    _, trained_here = surrogate.cached_network(ts, path=net_path, epochs=args.train_epochs,
                                               batch=32768, lr=3e-3)
    t_net = time.perf_counter() - t0
    _progress("surrogate launch options %s" % sopts)
    sur = _keig_leg(pr, ts, args.warmup, args.steps)
    _progress("surrogate tracker: %.4g histories/s" % sur["histories_per_s"])
    batch = tsl_batch_rates(ts)
    ts.close()
    ev = sur["events"]
    kt = sur.pop("_kt")
    pod_kt = pod.pop("_kt")
    # the surrogate run's event-kernel rows: the network sampled mlp_rows
    # thermal collisions (of the event iterations and of the tail's rounds);
    # the POD sampler the network's fallbacks (in the thermal-scattering
    # kind) and the thermal collisions of the tail kernel's last launch
    counts = {"advance": ev["advances"] - ev["tail_advances"],
              "collide": ev["collisions"] - ev["tail_collisions"],
              "cross": ev["crossings"] - ev["tail_crossings"],
              "tsl": ev["tsl_fallbacks"], "mlp": ev["mlp_rows"]}
    kt = {k: v for k, v in kt.items() if v[1] > 0}
    pod_kt = {k: v for k, v in pod_kt.items() if v[1] > 0}
    dk = abs(sur["k"] - pod["k"]) / np.hypot(sur["sigma_k"], pod["sigma_k"])
    ks_s, ks_p = sur.pop("k_cycles"), pod.pop("k_cycles")
    ratio_ci = fom_ratio_interval(ks_s, sur["seconds"], ks_p, pod["seconds"])
    return {
        "value": sur["histories_per_s"], "unit": "histories/s",
        "ms_per_step": sur["seconds"] * 1e3 / args.steps, "dtype": "f64",
        "config": {"workload": "c5: multi-zone continuous-temperature k-eigenvalue, "
                   "60 cm box of 27 cubes (fuel / bound-H moderator / structure), "
                   "per-cube T 300-1200 K, surrogate thermal scattering in the "
                   "moderator; 1e6 histories per cycle, one step = one active cycle",
                   "histories_per_cycle": n, "inactive_cycles": args.warmup,
                   "active_cycles": args.steps},
        "fom_ratio": sur["fom"] / pod["fom"],
        "fom_ratio_batch_means": (sur["fom_batch_means"] / pod["fom_batch_means"]
                                  if sur["fom_batch_means"] and pod["fom_batch_means"]
                                  else None),
        "fom_ratio_68pct": ratio_ci,
        "fom_note": "FOM = 1/(sigma_rel(k)^2 T) over the active cycles; sigma from "
                    "batch means of %d cycles; the interval is a bootstrap over "
                    "batches of both runs" % FOM_BATCH,
        "thermal_sampling": {"network_rows": ev["mlp_rows"],
                             "pod_fallbacks": ev["tsl_fallbacks"],
                             "pod_in_tail_kernel": _tail_pod(ev),
                             "network_fraction": ev["mlp_rows"] / max(ev["tsl_collisions"], 1)},
        "histories_per_s_ratio": sur["histories_per_s"] / pod["histories_per_s"],
        "k_diff_sigma": float(dk),
        "surrogate": sur, "pod": pod, "tsl_batch": batch,
        "autotune": {"best": opts, "best_surrogate": sopts, "seconds": t_tune,
                     "ms_per_cycle": [[o, s * 1e3] for o, s in table],
                     "surrogate_ms_per_cycle": [[o, s * 1e3] for o, s in stable]},
        "network": {"trained_here": trained_here, "seconds": t_net,
                    "path": os.path.relpath(net_path, ROOT)},
        "roofline": _roofline("c5", kt, counts),
        "tail_roofline": _tail_roofline("c5", kt, ev),
        "pod_tail_roofline": _tail_roofline("c5", pod_kt, pod["events"]),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--config", default="c5", choices=("c5", "c2"))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--train-epochs", type=int, default=8)
    ap.add_argument("--tail-threshold", type=int, default=None,
                    help="with --graph-iters: skip the autotuner")
    ap.add_argument("--graph-iters", type=int, default=None)
    ap.add_argument("--tail-network", type=int, default=None,
                    help="surrogate tracker's tail_network option (default: tuned)")
    ap.add_argument("--compact-interval", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one MI355X; "
              "the multi-GPU design (history sharding + one all-reduce of FP64 "
              "tallies) is not run here.", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = 100 if args.config == "c5" else 10
    if args.warmup is None:
        args.warmup = 20 if args.config == "c5" else 1
    rank = int(os.environ.get("RANK", "0"))
    res = run_c5(args) if args.config == "c5" else run_c2(args)
    out = {"metric": "FOM 1/(sigma^2 T) + histories/s, surrogate-TSL vs POD-TSL "
                     "tracker, 1 MI355X",
           "value": res.pop("value"), "unit": res.pop("unit"), "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res.pop("ms_per_step"), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": res.pop("dtype"),
           "data": "synthetic"}
    out.update(res)
    if not args.no_cpu:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        _progress("CPU baseline on %d threads" % threads)
        out["cpu_baseline"] = cpu_baseline(args.config, threads)
    else:
        out["cpu_baseline"] = None
    if rank == 0:
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
