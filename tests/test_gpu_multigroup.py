"""GPU tracker vs the CPU oracle on multigroup problems (identical seeds).

Single events: integer outputs (cells, surfaces, events, groups, RNG states)
bit-exact, floats within 1e-12 relative.  Whole runs: identical per-particle
random streams make most trajectories identical, so tallies agree to a few
histories; the statistical checks then pin them to closed forms.
"""
import math
import os

import numpy as np
import pytest

from minimc_amd import abi, configs, load_xml
from oracle.oracle import Oracle, SoA
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

REL = 1e-12


@pytest.fixture(scope="module")
def mg_pair(golden):
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = load_xml(os.path.join(golden, "multigroup.xml"))
    return pr, Oracle(pr), Tracker(pr, max_in_flight=1 << 16)


def _close(a, b, rel=REL):
    a, b = np.asarray(a), np.asarray(b)
    same = (a == b) | (np.isinf(a) & np.isinf(b))
    ok = same | (np.abs(a - b) <= rel * np.maximum(np.abs(a), np.abs(b)))
    return bool(np.all(ok))


def _random_states(n, seed, cells=(0, 1, 2), rmax=2.9):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = rng.uniform(0, rmax, n)
    p = rng.normal(size=(n, 3))
    p = p / np.linalg.norm(p, axis=1)[:, None] * r[:, None]
    return SoA(n, x=p[:, 0], y=p[:, 1], z=p[:, 2], u=d[:, 0], v=d[:, 1], w=d[:, 2],
               E=rng.integers(1, 3, n).astype(float), rng=rng.integers(1, 2**31 - 1, n))


def test_source_sampling_bit_exact(mg_pair):
    pr, o, t = mg_pair
    n = 4096
    a = o.sample_source(1, n)
    b = t.sample_source(1, n, SoA(n))
    assert np.array_equal(a.rng, b.rng)
    assert np.array_equal(a.cell, b.cell)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.E, b.E)
    for f in ("u", "v", "w"):
        assert _close(getattr(a, f), getattr(b, f), 1e-14)
    # the direction x component is a pure canonical: bit-exact
    assert np.array_equal(a.u, b.u)


def test_find_cell_and_nearest_surface(mg_pair):
    pr, o, t = mg_pair
    s = _random_states(20000, 1)
    a = o.eval_event(abi.MMC_EVAL_FIND_CELL, s.copy())
    b = t.eval_event(abi.MMC_EVAL_FIND_CELL, s.copy())
    assert np.array_equal(a.cell, b.cell)
    s.cell = a.cell.copy()
    a = o.eval_event(abi.MMC_EVAL_NEAREST_SURFACE, s.copy())
    b = t.eval_event(abi.MMC_EVAL_NEAREST_SURFACE, s.copy())
    assert np.array_equal(a.iout, b.iout)
    assert _close(a.out0, b.out0)


def test_xs_distance_collide(mg_pair):
    pr, o, t = mg_pair
    s = _random_states(20000, 2)
    s.cell = o.eval_event(abi.MMC_EVAL_FIND_CELL, s.copy()).cell.copy()
    s = _subset(s, s.cell < 3)
    for kind in (abi.MMC_EVAL_XS_TOTAL, abi.MMC_EVAL_COLLISION_DISTANCE):
        a = o.eval_event(kind, s.copy())
        b = t.eval_event(kind, s.copy())
        assert _close(a.out0, b.out0), kind
        assert np.array_equal(a.rng, b.rng)
    a = o.eval_event(abi.MMC_EVAL_COLLIDE, s.copy())
    b = t.eval_event(abi.MMC_EVAL_COLLIDE, s.copy())
    assert np.array_equal(a.event, b.event)
    assert np.array_equal(a.E, b.E)
    assert np.array_equal(a.rng, b.rng)
    assert np.array_equal(a.iout, b.iout)
    for f in ("u", "v", "w"):
        assert _close(getattr(a, f), getattr(b, f), 1e-14)


def _subset(s, mask):
    out = SoA(int(mask.sum()))
    for f in SoA.DFIELDS + SoA.IFIELDS + ("rng",):
        setattr(out, f, np.ascontiguousarray(getattr(s, f)[mask]))
    return out


def test_scatter_rotation(mg_pair):
    pr, o, t = mg_pair
    s = _random_states(20000, 3)
    s.out0 = np.random.default_rng(4).uniform(-1, 1, s.n)
    s.out1[:] = 1.0
    a = o.eval_event(abi.MMC_EVAL_SCATTER_ROTATE, s.copy())
    b = t.eval_event(abi.MMC_EVAL_SCATTER_ROTATE, s.copy())
    assert np.array_equal(a.rng, b.rng)
    for f in ("u", "v", "w"):
        assert _close(getattr(a, f), getattr(b, f), 1e-13)


def test_fixed_source_leakage_matches_oracle():
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = configs.point_absorber_sphere()
    n = 100000
    o = Oracle(pr)
    o.run_fixed_source(n, threads=8)
    t = Tracker(pr, max_in_flight=1 << 15)  # several chunks
    t.run_fixed_source(n)
    So, S2o = o.tallies(0)
    St, S2t = t.tallies(0)
    assert abs(So[0] - St[0]) <= 3          # identical streams: nearly identical
    assert S2t[0] == St[0]                  # 0/1 scores per history
    p = math.exp(-1)
    assert abs(St[0] / n - p) < 3 * math.sqrt(p * (1 - p) / n)
    c = t.counters()
    assert c["histories"] == n
    assert c["crossings"] == int(St[0])


def test_reference_multigroup_fixed_source(golden):
    """The reference's multigroup.xml (3 shells, fission-free, 2 groups) with a
    current estimator on the outer shell: GPU vs oracle per bin."""
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = load_xml(os.path.join(golden, "multigroup.xml"))
    pr.estimators.append(dict(kind=abi.MMC_EST_CURRENT, name="out", surface=2,
                              cosine=dict(direction=(1, 0, 0),
                                          bins=dict(kind=abi.MMC_BINS_LINSPACE, min=-1,
                                                    max=1, bins=4)),
                              energy=dict(kind=abi.MMC_BINS_BOUNDARIES, bounds=[1.5])))
    pr.estimators.append(dict(kind=abi.MMC_EST_CURRENT, name="mid", surface=1))
    n = 20000
    o = Oracle(pr)
    o.run_fixed_source(n, threads=8)
    t = Tracker(pr, max_in_flight=1 << 14)
    t.run_fixed_source(n)
    for e in range(2):
        So, S2o = o.tallies(e)
        St, S2t = t.tallies(e)
        assert np.abs(So - St).max() <= 0.002 * n + 3, (e, So, St)
        assert np.abs(S2o - S2t).max() <= 0.01 * max(S2o.max(), 1) + 10


def test_keigenvalue_slab_matches_oracle():
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = configs.c1_slab(histories=4000)
    o = Oracle(pr)
    t = Tracker(pr, max_in_flight=1 << 14)
    for x in (o, t):
        x.set_entropy_mesh(*pr.entropy_mesh)
        x.keig_init(pr.histories)
    ko, kt = [], []
    same_first = None
    for c in range(30):
        a = o.run_cycle(active=c >= 10, threads=8)
        b = t.run_cycle(active=c >= 10)
        if c == 0:
            same_first = (a["n_sites"], b["n_sites"])
        if c >= 10:
            ko.append(a["k_collision"])
            kt.append(b["k_collision"])
    assert abs(same_first[0] - same_first[1]) <= 2
    ko, kt = np.array(ko), np.array(kt)
    se = math.sqrt(ko.var(ddof=1) / len(ko) + kt.var(ddof=1) / len(kt))
    assert abs(ko.mean() - kt.mean()) < 3 * se + 1e-9
    So, _ = o.tallies(0)
    St, _ = t.tallies(0)
    assert St.sum() > 0
    assert abs(St.sum() - So.sum()) / So.sum() < 0.05


def test_out_file_of_the_reference_input(golden, tmp_path):
    """The reference's point-source leakage input through the GPU tracker,
    written as the reference's .out file: the same text as the oracle's run
    apart from the last digits of the numbers, and means equal to S / N."""
    require_gpu()
    from minimc_amd import output
    from minimc_amd.tracker import Tracker
    pr = load_xml(os.path.join(golden, "point_source_leakage.xml"))
    o = Oracle(pr)
    o.run_fixed_source(pr.histories, threads=8)
    t = Tracker(pr, max_in_flight=1 << 12)
    t.run_fixed_source(pr.histories)
    n_est = len(pr.estimators)
    to = output.out_text(pr, [o.tallies(e) for e in range(n_est)], pr.histories)
    tg = output.out_text(pr, [t.tallies(e) for e in range(n_est)], pr.histories)
    lo, lg = to.split("\n"), tg.split("\n")
    assert len(lo) == len(lg)
    for a, b in zip(lo, lg):
        if "e+" in a or "e-" in a:
            va = [float(x) for x in a.split(",") if x.strip()]
            vb = [float(x) for x in b.split(",") if x.strip()]
            assert len(va) == len(vb)
            assert np.allclose(va, vb, rtol=0.05, atol=3.0 / pr.histories)
        else:
            assert a == b
    path = tmp_path / "point_source_leakage.out"
    output.write_out(str(path), pr, [t.tallies(e) for e in range(n_est)], pr.histories)
    assert path.read_text() == tg


def test_sensitivity_point_absorber_closed_form():
    """Indirect effects on the GPU: every leaking history of the point
    absorber streams once from the centre, so the sensitivity tally is
    (1 / sigma - (R + nudge)) times the leakage tally."""
    require_gpu()
    from minimc_amd.tracker import Tracker
    from tests.test_oracle_perturbation import NUDGE, absorber_with_sensitivity
    pr = absorber_with_sensitivity(radius=2.5, sigma=0.4)
    n = 50000
    t = Tracker(pr, max_in_flight=1 << 14)
    t.run_fixed_source(n)
    S, S2 = t.tallies(0)
    Ss, S2s = t.tallies(1)
    ie = 1 / 0.4 - (2.5 + NUDGE)
    assert S[0] > 0.1 * n * np.exp(-1.0)
    np.testing.assert_allclose(Ss[0], S[0] * ie, rtol=1e-12)
    np.testing.assert_allclose(S2s[0], S2[0] * ie * ie, rtol=1e-12)


@pytest.mark.parametrize("tracking", [abi.MMC_SURFACE_TRACKING, abi.MMC_CELL_DELTA_TRACKING])
def test_sensitivity_ce_slab_matches_oracle(tracking):
    """Total cross-section perturbation of the slab's nuclide, sensitivities
    of both leakage currents: many streams per history, the tail kernel and
    compaction on the GPU side."""
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = configs.c2_slab(n_points=60000, n_res=200)
    pr.tracking = tracking
    pr.perturbations.append(dict(name="total", nuclide=0))
    for i in range(2):
        e = pr.estimators[i]
        pr.estimators.append(dict(e, name=e["name"] + "::total", sensitivity=1, parent=i))
    n = 20000
    o = Oracle(pr)
    o.run_fixed_source(n, threads=8)
    t = Tracker(pr, max_in_flight=1 << 13)
    t.run_fixed_source(n)
    for e in (2, 3):
        So, S2o = o.tallies(e)
        St, S2t = t.tallies(e)
        # identical streams except after rare libm branch flips
        assert abs(So.sum() - St.sum()) <= 0.01 * np.abs(So).sum() + 1e-9, e
        assert abs(S2o.sum() - S2t.sum()) <= 0.02 * S2o.sum() + 1e-9, e
        assert np.all(np.isfinite(St))
    # compaction by the advance kernel after every iteration moves the
    # indirect effects with the particles: the same scores, and the tallies
    # are reduced in a fixed order, so S and S2 are bit-identical
    t1 = Tracker(pr, max_in_flight=1 << 13, compact_interval=1)
    t1.run_fixed_source(n)
    for e in range(4):
        assert all(np.array_equal(a, b) for a, b in zip(t.tallies(e), t1.tallies(e))), e
    t1.close()


def sliced_slab(n_slices=400, width=20.0):
    """The C1 slab cut into n_slices cells by PlaneX surfaces, fixed source
    at its centre, leakage currents on both faces.  Its descriptor block
    (over 24 KiB) is too large to stage, so the tracker runs the kernels that
    read every table from HBM.  The nuclide is fissile: with a small bank the
    fixed-source driver must rerun chunks whose secondaries overflow it."""
    pr = configs.c1_slab(width=width)
    pr.cells = []
    planes = [pr.add_surface("x%d" % i, abi.MMC_PLANEX, a=width * i / n_slices)
              for i in range(1, n_slices)]
    edges = [0] + planes + [1]  # surfaces 0 and 1 are the slab's faces
    for i in range(n_slices):
        pr.add_cell("s%d" % i, 0, [(edges[i], False), (edges[i + 1], True)])
    pr.add_cell("", -1, [(0, True)])
    pr.add_cell("", -1, [(1, False)])
    pr.problem = "fixedsource"
    pr.estimators = [e for e in pr.estimators if e["kind"] == abi.MMC_EST_CURRENT]
    # off every plane: a source point on a surface is an ill-posed start (its
    # first crossing has distance 0)
    pr.source = dict(pr.source, position=(width / 2 + 0.013, 0.0, 0.0))
    return pr


def test_unstaged_kernels_match_oracle():
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = sliced_slab()
    n = 10000
    o = Oracle(pr)
    o.run_fixed_source(n, threads=8)
    t = Tracker(pr, max_in_flight=1 << 13)
    t.run_fixed_source(n)
    for e in range(len(pr.estimators)):
        So, _ = o.tallies(e)
        St, S2t = t.tallies(e)
        # identical streams: nearly identical counts
        assert np.abs(So - St).max() <= 0.002 * n + 3, (e, So, St)
    c = t.counters()
    assert c["histories"] == n and c["crossings"] > 10 * n
