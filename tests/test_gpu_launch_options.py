"""Launch options change the schedule, not the physics: the history-based
tail kernel, bank compaction and the readback interval must leave every
k-eigenvalue cycle bit-identical (same events per particle in the same
order, so the same random numbers; k sums and the fission bank are reduced
in a fixed order).  The FP64 tallies (S, S2 of every bin: flux, currents,
sensitivities) are reduced in a fixed order too (per-history sums, then a
fixed tree per bin), so they are bit-identical across schedules as well."""
import numpy as np
import pytest

from minimc_amd import abi, configs
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

SCHEDULES = [
    dict(tail_threshold=-1, compact_interval=-1, graph_iters=8),   # events only
    dict(tail_threshold=1 << 30, compact_interval=-1, graph_iters=4),  # tail early
    dict(tail_threshold=2048, compact_interval=1, graph_iters=8),  # compaction in KAdvCol
    # separate advance and collision kernels (KAdvance, KCollide)
    dict(tail_threshold=2048, compact_interval=1, graph_iters=8, split_collide=1),
    dict(tail_threshold=-1, compact_interval=2, graph_iters=8, split_collide=1),
    dict(tail_threshold=2048, compact_interval=2, graph_iters=8),  # KCompact
    dict(),                                                         # defaults
    # batches replayed as captured graphs (k-eigenvalue only)
    dict(tail_threshold=2048, compact_interval=1, graph_iters=8, graphs=1),
    dict(tail_threshold=-1, compact_interval=0, graph_iters=4, graphs=1),
]
KEYS = ("k_analog", "k_collision", "k_track", "k_absorption", "entropy", "n_sites")


def _cycles(pr, opts, n_cycles, cap):
    from minimc_amd.tracker import Tracker
    t = Tracker(pr, tsl_mode=abi.MMC_TSL_POD, max_in_flight=cap, **opts)
    t.set_entropy_mesh(*pr.entropy_mesh)
    t.keig_init(pr.histories)
    out = [t.run_cycle(active=c >= 1) for c in range(n_cycles)]
    c = t.counters()
    c["tallies"] = [t.tallies(e) for e in range(len(pr.estimators))]
    t.close()
    return out, c


def _multizone_delta_linear_mixture(histories):
    """The multi-zone problem with cell delta tracking (virtual collisions), a
    linear temperature field in every cube (+-45 K about the cube's value, so a
    collision's temperature differs from its flight's; inside the thermal-
    scattering tables' 300-1200 K) and fuel mixed with
    bound hydrogen (the nuclide a collision samples is not always the last one
    summed into the material total)."""
    import itertools
    pr = configs.c5_multizone(histories=histories)
    e = configs.C5_EDGES
    g = (2.0, -1.5, 1.0)  # K/cm; 10 cm half-width: +-45 K
    for n, (i, j, k) in enumerate(itertools.product(range(3), repeat=3)):
        ctr = [(e[a] + e[a + 1]) / 2 for a in (i, j, k)]
        c = min(max(pr.fields[n + 1]["c"], 345.0), 1155.0)
        pr.fields[n + 1] = dict(kind=abi.MMC_FIELD_LINEAR, c=c, g=g,
                                b=c - sum(gg * x for gg, x in zip(g, ctr)),
                                lower=c - 45.0, upper=c + 45.0)
    pr.materials[0]["comps"] = [(0, 0.8), (1, 0.2)]  # fuel + bound hydrogen
    pr.tracking = abi.MMC_CELL_DELTA_TRACKING
    return pr


def _same_tallies(a, b, what):
    for e, ((S, S2), (T, T2)) in enumerate(zip(a, b)):
        assert np.array_equal(S, T), (what, e, np.flatnonzero(S != T)[:5])
        assert np.array_equal(S2, T2), (what, e, np.flatnonzero(S2 != T2)[:5])


@pytest.mark.parametrize("case", ["multigroup_slab", "multizone", "multizone_delta_linear"])
def test_schedules_give_identical_cycles(case):
    """Among the schedules, the joined advance-and-collision kernel and the
    tail kernel hand the advance's lookups (thermal-scattering total, nuclide
    total, nu Sigma_f) to the collision; the separate kernels (split_collide)
    look everything up again.  The third case makes the collision's
    temperature and sampled nuclide differ from the advance's."""
    require_gpu()
    if case == "multigroup_slab":
        pr, n_cycles, cap = configs.c1_slab(histories=4000), 6, 1 << 14
    elif case == "multizone":
        pr, n_cycles, cap = configs.c5_multizone(histories=20000), 4, 1 << 16
    else:
        pr, n_cycles, cap = _multizone_delta_linear_mixture(20000), 4, 1 << 16
    runs = [_cycles(pr, o, n_cycles, cap) for o in SCHEDULES]
    ref, ref_c = runs[0]
    for (st, c), opts in zip(runs[1:], SCHEDULES[1:]):
        for a, b in zip(ref, st):
            for k in KEYS:
                assert a[k] == b[k], (case, opts, k, a[k], b[k])
        for k in ("collisions", "crossings", "advances", "tsl_collisions", "fission_sites"):
            assert c[k] == ref_c[k], (case, opts, k)
        _same_tallies(ref_c["tallies"], c["tallies"], (case, opts))
    if case == "multizone_delta_linear":
        assert ref_c["virtual_collisions"] > 0 and ref_c["tsl_collisions"] > 0
    # the early-tail schedule really ran histories in the tail kernel
    assert runs[1][1]["tail_particles"] > 0
    assert ref_c["tail_particles"] == 0
    # the graph schedules replayed captured batches, from a few captures; with
    # a tail threshold the small slab can reach the tail before its first
    # captured batch, without the tail it cannot
    for (_, c), opts in zip(runs, SCHEDULES):
        if opts.get("graphs"):
            assert c["graph_launches"] >= c["graph_captures"], (case, opts, c)
            if opts["tail_threshold"] < 0 or case != "multigroup_slab":
                assert c["graph_launches"] > c["graph_captures"] > 0, (case, opts, c)
        else:
            assert c["graph_launches"] == 0


def test_surrogate_schedules_give_identical_cycles():
    """The surrogate step parks the collisions it cannot serve until the end
    of a batch of iterations; compaction in between must move them too."""
    require_gpu()
    from minimc_amd import surrogate
    from minimc_amd.tracker import Tracker
    pr = configs.c5_multizone(histories=20000)
    res = []
    for cp, gr in ((-1, 0), (1, 0), (2, 0), (1, 1)):
        t = Tracker(pr, tsl_mode=abi.MMC_TSL_SURROGATE, max_in_flight=1 << 16,
                    compact_interval=cp, tail_threshold=2048, graphs=gr)
        s, _ = surrogate.cached_network(t)  # the shipped network
        t.set_entropy_mesh(*pr.entropy_mesh)
        t.keig_init(pr.histories)
        res.append(([t.run_cycle(active=c >= 1) for c in range(4)], t.counters()))
        t.close()
    (ref, rc) = res[0]
    assert rc["tsl_fallbacks"] > 0  # the parked path was exercised
    for st, c in res[1:]:
        for a, b in zip(ref, st):
            for k in KEYS:
                assert a[k] == b[k], (k, a[k], b[k])
        assert c["tsl_fallbacks"] == rc["tsl_fallbacks"]


def test_surrogate_tail_rounds_match_event_iterations():
    """With the network serving every thermal collision of the tail
    (tail_network=1: tail rounds until no particle is left), the tail gives
    the same cycles as event iterations alone: the tail kernel runs the same
    event functions, and the network kernel the same rows.  Rounds down to
    4096 live particles (then POD) and the default (POD throughout the tail)
    are checked for their counters."""
    require_gpu()
    from minimc_amd import surrogate
    from minimc_amd.tracker import Tracker
    pr = configs.c5_multizone(histories=20000)
    runs = {}
    for name, opts in (("events", dict(tail_threshold=-1)),
                       ("rounds", dict(tail_threshold=8192, tail_network=1)),
                       ("rounds_early", dict(tail_threshold=1 << 30, tail_network=1,
                                             graph_iters=4)),
                       ("pod_tail", dict(tail_threshold=1 << 30)),
                       ("to_4096", dict(tail_threshold=1 << 30, tail_network=4096))):
        t = Tracker(pr, tsl_mode=abi.MMC_TSL_SURROGATE, max_in_flight=1 << 16,
                    compact_interval=1, **opts)
        surrogate.cached_network(t)  # the shipped network
        t.set_entropy_mesh(*pr.entropy_mesh)
        t.keig_init(pr.histories)
        runs[name] = ([t.run_cycle(active=c >= 1) for c in range(4)], t.counters())
        t.close()

    def pod_in_tail(c):
        return c["tsl_collisions"] - c["mlp_rows"] - c["tsl_fallbacks"]

    ref, rc = runs["events"]
    assert rc["tail_particles"] == 0 and pod_in_tail(rc) == 0
    for name in ("rounds", "rounds_early"):
        st, c = runs[name]
        for a, b in zip(ref, st):
            for k in KEYS:
                assert a[k] == b[k], (name, k, a[k], b[k])
        for k in ("collisions", "advances", "tsl_collisions", "mlp_rows", "tsl_fallbacks"):
            assert c[k] == rc[k], (name, k)
        assert c["tail_tsl"] > 0 and pod_in_tail(c) == 0, (name, c)
    # POD throughout the tail: the tail kernel samples its thermal collisions
    _, c = runs["pod_tail"]
    assert pod_in_tail(c) == c["tail_tsl"] > 0, c
    # rounds down to 4096 live particles, then POD
    _, c = runs["to_4096"]
    assert 0 < pod_in_tail(c) < c["tail_tsl"], c
    assert c["mlp_rows"] > runs["pod_tail"][1]["mlp_rows"], c


def test_record_overflow_rerun_gives_identical_cycles():
    """A k-eigenvalue cycle whose score records overflow their buffer (4 per
    bank slot to start with) grows the buffer and runs again from the same
    source.  With a bank of exactly one cycle's histories, the C1 slab's flux
    mesh (a score per collision) overflows in the first active cycle; a bank
    16x larger does not.  Cycles and tallies must be bit-identical."""
    require_gpu()
    pr = configs.c1_slab(histories=8192)
    small, c_small = _cycles(pr, dict(), 5, 8192)
    large, c_large = _cycles(pr, dict(), 5, 8192 * 16)
    # collisions per active cycle against the small bank's first buffer
    per_cycle = c_small["collisions"] / 5
    assert per_cycle > 4 * 8192 + 1024, per_cycle
    assert per_cycle < 4 * 8192 * 16, per_cycle
    for a, b in zip(small, large):
        for k in KEYS:
            assert a[k] == b[k], (k, a[k], b[k])
    for k in ("collisions", "crossings", "advances", "fission_sites"):
        assert c_small[k] == c_large[k], k
    _same_tallies(c_small["tallies"], c_large["tallies"], "record overflow")


def _ce_slab_with_sensitivity_and_flux():
    """The continuous-energy slab (reduced grid) with a total cross-section
    perturbation, the sensitivity of both leakage tallies to it, and a flux
    mesh: every kind of fixed-source score."""
    pr = configs.c2_slab(n_points=6000, n_res=20)
    pr.perturbations.append(dict(name="total", nuclide=0))
    for e in (0, 1):
        pr.estimators.append(dict(pr.estimators[e], name="sens%d" % e, sensitivity=1,
                                  parent=e))
    pr.estimators.append(dict(kind=abi.MMC_EST_FLUX_MESH, name="flux",
                              mesh=dict(n=[40, 1, 1], lo=[-20.0, -1e30, -1e30],
                                        hi=[20.0, 1e30, 1e30])))
    return pr


@pytest.mark.parametrize("tracking", [abi.MMC_SURFACE_TRACKING, abi.MMC_CELL_DELTA_TRACKING])
def test_fixed_source_tallies_bit_identical_across_schedules(tracking):
    """Fixed source: the same histories under different schedules (tail,
    compaction, readback interval, bank size and hence chunking of the
    event loop within one chunk) give bit-identical S and S2."""
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = _ce_slab_with_sensitivity_and_flux()
    n = 6000
    res = []
    for opts in SCHEDULES:
        t = Tracker(pr, tracking=tracking, max_in_flight=1 << 13, **opts)
        t.run_fixed_source(n)
        res.append(([t.tallies(e) for e in range(len(pr.estimators))], t.counters()))
        t.close()
    ref, rc = res[0]
    assert rc["histories"] == n
    assert sum(S.sum() for S, _ in ref[:2]) > 0.05 * n  # some histories leak
    assert ref[4][0].sum() > 0                           # flux was scored
    assert np.any(ref[2][0] != 0)                         # sensitivities were scored
    for (tal, c), opts in zip(res[1:], SCHEDULES[1:]):
        _same_tallies(ref, tal, opts)
        for k in ("collisions", "crossings", "advances"):
            assert c[k] == rc[k], (opts, k)
