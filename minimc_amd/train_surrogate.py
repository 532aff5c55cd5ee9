"""Trains the thermal-scattering surrogate on the GPU and assesses it.

python -m minimc_amd.train_surrogate [--samples 26] [--epochs 3] [--batch 8192]
                                     [--lr 2e-3] [--rounds 1] [--gamma 1] [--lr-floor 1e-5]
                                     [--weighting local|bins] [--assess-n 2000000]
                                     [--out PATH] [--eval-only]

Training data: 2^samples POD transport-map targets generated on the device
from the synthetic S(alpha, beta, T) tables.  The assessment compares the
surrogate sampler with the POD sampler at a grid of (E, T): relative
Wasserstein-1 distance (W1 / std of the POD marginal) and the KS distance of
the alpha and beta marginals.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from minimc_amd import configs, surrogate  # noqa: E402
from minimc_amd.tracker import Tracker  # noqa: E402

ASSESS_POINTS = [(0.01, 300.0), (0.0253, 600.0), (0.1, 900.0), (0.3, 450.0),
                (0.6, 800.0), (1.5, 1000.0), (3.0, 1150.0), (3.9, 350.0)]


def w1_rel(a, b):
    q = np.linspace(0, 1, 4001)[1:-1]
    return float(np.mean(np.abs(np.quantile(a, q) - np.quantile(b, q))) / np.std(b))


def ks(a, b):
    a, b = np.sort(a), np.sort(b)
    grid = np.concatenate([a, b])
    fa = np.searchsorted(a, grid, side="right") / len(a)
    fb = np.searchsorted(b, grid, side="right") / len(b)
    return float(np.max(np.abs(fa - fb)))


def assess(t, n=400000, points=ASSESS_POINTS, seed=9):
    rows = []
    for E_eV, T in points:
        E = t.device_array(np.full(n, E_eV * 1e-6))
        Tt = t.device_array(np.full(n, T))
        r0 = np.random.default_rng(seed).integers(1, 2**31 - 1, n).astype(np.uint32)
        r1, r2 = t.device_array(r0), t.device_array(r0 ^ 0x5bd1e995 | 1)
        a1 = t.device_array(shape=n, dtype=np.float64)
        b1 = t.device_array(shape=n, dtype=np.float64)
        a2 = t.device_array(shape=n, dtype=np.float64)
        b2 = t.device_array(shape=n, dtype=np.float64)
        t.tsl_sample_pod(E, Tt, r1, a1, b1)
        t.tsl_sample_surrogate(E, Tt, r2, a2, b2)
        ap, bp, am, bm = a1.numpy(), b1.numpy(), a2.numpy(), b2.numpy()
        rows.append(dict(E_eV=E_eV, T=T, w1_beta=w1_rel(bm, bp), w1_alpha=w1_rel(am, ap),
                         ks_beta=ks(bm, bp), ks_alpha=ks(am, ap)))
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=26)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--rounds", type=int, default=1,
                    help="training rounds; between rounds the loss weight of each "
                         "(ln E, T) bin follows its relative error")
    ap.add_argument("--gamma", type=float, default=1.0)
    ap.add_argument("--lr-floor", type=float, default=1e-5)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--weighting", choices=("local", "bins"), default="local")
    ap.add_argument("--e-uniform", type=float, default=surrogate.E_UNIFORM_FRACTION,
                    help="fraction of training energies drawn uniformly in E")
    ap.add_argument("--assess-n", type=int, default=2000000,
                    help="samples per assessment point and sampler")
    args = ap.parse_args()
    pr = configs.tsl_box(configs.tsl_tables(), T=600.0)
    t = Tracker(pr, max_in_flight=1 << 16)
    s = surrogate.Surrogate(t)
    out = args.out or surrogate.SHIPPED
    info = {}
    if args.eval_only:
        s.load(out)
    else:
        t0 = time.perf_counter()
        losses = s.train(n_samples=1 << args.samples, batch=args.batch, epochs=args.epochs,
                         lr=args.lr, log=lambda m: print(m, flush=True), rounds=args.rounds,
                         gamma=args.gamma, lr_floor=args.lr_floor, seed=args.seed,
                         weighting=args.weighting, e_uniform=args.e_uniform)
        dt = time.perf_counter() - t0
        steps = args.epochs * ((1 << args.samples) // args.batch)
        info = dict(train_seconds=dt, steps=steps, samples=1 << args.samples,
                    epochs=args.epochs, batch=args.batch, lr=args.lr, rounds=args.rounds,
                    gamma=args.gamma, lr_floor=args.lr_floor, seed=args.seed,
                    weighting=args.weighting, e_uniform=args.e_uniform,
                    final_loss=losses[-1],
                    train_tflops=steps * args.batch * 3 * 2 * surrogate.n_params() / dt / 1e12)
        s.save(out, info=info)
    rows = assess(t, n=args.assess_n)
    for r in rows:
        print("E=%6.3f eV T=%6.1f  W1 beta %.4f alpha %.4f  KS beta %.4f alpha %.4f" % (
            r["E_eV"], r["T"], r["w1_beta"], r["w1_alpha"], r["ks_beta"], r["ks_alpha"]),
            flush=True)
    info["assessment"] = rows
    info["max_w1"] = max(max(r["w1_beta"], r["w1_alpha"]) for r in rows)
    print(json.dumps(info))


if __name__ == "__main__":
    main()
