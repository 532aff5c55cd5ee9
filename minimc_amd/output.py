"""Tally results in the reference's output format.

`minimc.cpp:17-24` writes `<input>.out`: the batch size on its own line,
then `EstimatorSet::to_string()` (`Estimator.cpp:207-213`), which is, for
every estimator, a blank line, `Estimator::to_string` (`Estimator.cpp:48-57`)
and another newline:

    name
    ====                       (one '=' per character of the name)
    <blank>
    cosine                     (ParticleBins::to_string, Bins.cpp:206-210)
    ------
    <cosine bin edges, or "none">
    <blank>
    energy
    ------
    <energy bin edges, or "none">
    <blank>
    mean                       (Scorable::GetScoreAsString, Scorable.cpp:51-70)
    ----
    S_b / W, ... each followed by ", "
    <blank>
    std dev
    -------
    sqrt(S2_b - S_b^2 / W) / W, ...
    <blank>

Every number is written as `std::ostream << std::scientific` writes it with
the default precision of 6, which is C's "%.6e" (so a NaN with its sign bit
set, as glibc's sqrt of a negative number returns, prints as "-nan").
Bin edges follow LinspaceBins / LogspaceBins / BoundaryBins::to_string
(Bins.cpp:84-90, :125-131, :164-170).  W is the total weight: the number of
histories for a fixed-source run.  For a k-eigenvalue run the tracker folds
each active cycle's per-history mean into S and S2, so W is the number of
active cycles and the "std dev" line is the standard error of the cycle
means.

A sensitivity (an entry of Problem.estimators with a "parent") is written
under its parent estimator, after the parent's scores, as
Sensitivity::to_string (Sensitivity.cpp:25-28) does: its name
("<estimator>::<perturbation>"), a '=' rule, a blank line, its score block
and a newline.

Estimators that the reference does not have (flux mesh, alpha and beta
histograms) are written in the same layout: the alpha or beta bins take the
energy block, and a mesh writes its shape in place of the bin edges.
"""
import math

import numpy as np

from . import abi


def scientific(v):
    """One number as libstdc++'s std::scientific (precision 6) prints it."""
    v = float(v)
    if math.isnan(v):
        return "-nan" if math.copysign(1.0, v) < 0 else "nan"
    return "%.6e" % v


def bins_to_string(b):
    """Bins::to_string of one bins spec (a dict as in Problem.estimators)."""
    if not b:
        return "none"  # NoBins::to_string (Bins.cpp:51)
    kind = b["kind"]
    out = []
    if kind == abi.MMC_BINS_BOUNDARIES:
        for x in b["bounds"]:
            out.append(scientific(x) + ", ")
        return "".join(out)
    n_bins = int(b["bins"]) + 2
    lo, hi = float(b["min"]), float(b["max"])
    width = (hi - lo) / (n_bins - 2)
    for i in range(n_bins - 1):
        x = lo + float(i) * width
        if kind == abi.MMC_BINS_LOGSPACE:
            x = math.pow(float(b.get("base", 10.0)), x)
        out.append(scientific(x) + ", ")
    return "".join(out)


def scores_to_string(S, S2, total_weight):
    """Scorable::GetScoreAsString (Scorable.cpp:51-70)."""
    W = float(total_weight)
    S = np.asarray(S, dtype=np.float64)
    S2 = np.asarray(S2, dtype=np.float64)
    mean = "".join(scientific(s / W) + ", " for s in S)
    with np.errstate(invalid="ignore"):
        sd = np.sqrt(S2 - S * S / W) / W
    dev = "".join(scientific(x) + ", " for x in sd)
    return "mean\n----\n" + mean + "\n\n" + "std dev\n-------\n" + dev + "\n\n"


def _bins_block(e):
    kind = e.get("kind", abi.MMC_EST_CURRENT)
    if kind == abi.MMC_EST_FLUX_MESH:
        m = e["mesh"]
        shape = ", ".join(str(int(n)) for n in m["n"])
        box = "".join("%s, %s, " % (scientific(lo), scientific(hi))
                      for lo, hi in zip(m["lo"], m["hi"]))
        return "mesh\n----\n" + shape + "\n" + box + "\n\n"
    cos = e.get("cosine")
    cosine = bins_to_string(cos["bins"]) if cos else "none"
    return ("cosine\n------\n" + cosine + "\n\n" +
            "energy\n------\n" + bins_to_string(e.get("energy")) + "\n\n")


def estimator_to_string(e, S, S2, total_weight):
    """Estimator::to_string (Estimator.cpp:48-57), without its sensitivities."""
    name = e.get("name") or ""
    return (name + "\n" + "=" * len(name) + "\n\n" + _bins_block(e) +
            scores_to_string(S, S2, total_weight))


def sensitivity_to_string(e, S, S2, total_weight):
    """Sensitivity::to_string (Sensitivity.cpp:25-28)."""
    name = e.get("name") or ""
    return (name + "\n" + "=" * len(name) + "\n\n" +
            scores_to_string(S, S2, total_weight) + "\n")


def to_string(problem, tallies, total_weight):
    """EstimatorSet::to_string (Estimator.cpp:207-213).  `tallies[i]` is
    (S, S2) of problem.estimators[i], as Tracker.tallies(i) returns them."""
    if len(tallies) != len(problem.estimators):
        raise ValueError("one (S, S2) pair per estimator")
    ests = problem.estimators
    out = []
    for i, (e, (S, S2)) in enumerate(zip(ests, tallies)):
        if e.get("parent") is not None:
            continue
        text = estimator_to_string(e, S, S2, total_weight)
        for k, c in enumerate(ests):  # Estimator::to_string, Estimator.cpp:53-55
            if c.get("parent") == i:
                text += sensitivity_to_string(c, tallies[k][0], tallies[k][1], total_weight)
        out.append("\n" + text + "\n")
    return "".join(out)


def out_text(problem, tallies, batchsize, total_weight=None):
    """The whole `.out` file: batch size, newline, the estimator set."""
    W = batchsize if total_weight is None else total_weight
    return "%d\n" % int(batchsize) + to_string(problem, tallies, W)


def write_out(path, problem, tallies, batchsize, total_weight=None):
    with open(path, "w") as f:
        f.write(out_text(problem, tallies, batchsize, total_weight))
