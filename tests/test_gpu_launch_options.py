"""Launch options change the schedule, not the physics: the history-based
tail kernel, bank compaction and the readback interval must leave every
k-eigenvalue cycle bit-identical (same events per particle in the same
order, so the same random numbers; k sums and the fission bank are reduced
in a fixed order)."""
import numpy as np
import pytest

from minimc_amd import abi, configs
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu

SCHEDULES = [
    dict(tail_threshold=-1, compact_interval=-1, graph_iters=8),   # events only
    dict(tail_threshold=1 << 30, compact_interval=-1, graph_iters=4),  # tail early
    dict(tail_threshold=2048, compact_interval=1, graph_iters=8),  # compaction in KAdvance
    dict(tail_threshold=2048, compact_interval=2, graph_iters=8),  # KCompact
    dict(),                                                         # defaults
]
KEYS = ("k_analog", "k_collision", "k_track", "k_absorption", "entropy", "n_sites")


def _cycles(pr, opts, n_cycles, cap):
    from minimc_amd.tracker import Tracker
    t = Tracker(pr, tsl_mode=abi.MMC_TSL_POD, max_in_flight=cap, **opts)
    t.set_entropy_mesh(*pr.entropy_mesh)
    t.keig_init(pr.histories)
    out = [t.run_cycle(active=c >= 1) for c in range(n_cycles)]
    c = t.counters()
    t.close()
    return out, c


@pytest.mark.parametrize("case", ["multigroup_slab", "multizone"])
def test_schedules_give_identical_cycles(case):
    require_gpu()
    if case == "multigroup_slab":
        pr, n_cycles, cap = configs.c1_slab(histories=4000), 6, 1 << 14
    else:
        pr, n_cycles, cap = configs.c5_multizone(histories=20000), 4, 1 << 16
    runs = [_cycles(pr, o, n_cycles, cap) for o in SCHEDULES]
    ref, ref_c = runs[0]
    for (st, c), opts in zip(runs[1:], SCHEDULES[1:]):
        for a, b in zip(ref, st):
            for k in KEYS:
                assert a[k] == b[k], (case, opts, k, a[k], b[k])
        for k in ("collisions", "crossings", "advances", "tsl_collisions", "fission_sites"):
            assert c[k] == ref_c[k], (case, opts, k)
    # the early-tail schedule really ran histories in the tail kernel
    assert runs[1][1]["tail_particles"] > 0
    assert ref_c["tail_particles"] == 0


def test_surrogate_schedules_give_identical_cycles():
    """The surrogate step parks the collisions it cannot serve until the end
    of a batch of iterations; compaction in between must move them too."""
    require_gpu()
    from minimc_amd import surrogate
    from minimc_amd.tracker import Tracker
    pr = configs.c5_multizone(histories=20000)
    res = []
    for cp in (-1, 1, 2):
        t = Tracker(pr, tsl_mode=abi.MMC_TSL_SURROGATE, max_in_flight=1 << 16,
                    compact_interval=cp, tail_threshold=2048)
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
