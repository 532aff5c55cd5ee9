"""Autotuner host logic (no GPU): problem signatures and the search."""
from minimc_amd import autotune, configs


def test_signature_is_stable_and_specific():
    a = autotune._signature(configs.c1_slab(), 0, 1000)
    assert a == autotune._signature(configs.c1_slab(), 0, 1000)
    assert a != autotune._signature(configs.c1_slab(width=10.0), 0, 1000)
    assert a != autotune._signature(configs.c1_slab(), 1, 1000)
    assert a != autotune._signature(configs.c1_slab(), 0, 2000)


def test_search_picks_the_fastest(monkeypatch, tmp_path):
    # synthetic cost: best tail 16384, best readback interval 16, compaction
    # every 2 iterations
    def cost(problem, opts, n, cycles, tsl_mode, prepare):
        t = opts["tail_threshold"]
        c = opts.get("compact_interval", 0)
        return ((abs(t - 16384) if t > 0 else 1e6) + abs(opts["graph_iters"] - 16)
                + (0 if c == 2 else 0.5))
    monkeypatch.setattr(autotune, "_time_run", cost)
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "at.json"))
    best, table = autotune.tune(configs.c1_slab(), n=1000)
    assert best == {"tail_threshold": 16384, "graph_iters": 16, "compact_interval": 2}
    # (no captured-graph candidate: compaction every 2 iterations; separate
    # advance and collision kernels (a k-eigenvalue problem); the tail block
    # sizes, and no thermal-scattering block sizes: the slab has no
    # thermal-scattering law)
    assert len(table) == (len(autotune.TAIL_CANDIDATES) + len(autotune.ITER_CANDIDATES) - 1
                          + len(autotune.COMPACT_CANDIDATES) + 1
                          + len(autotune.TAIL_BLOCK_CANDIDATES))
    assert [o for o, _ in table[-2:]] == [dict(best, split_collide=1),
                                          dict(best, tail_block=512)]
    # the cached result is returned without timing again
    monkeypatch.setattr(autotune, "_time_run", None)
    best2, _ = autotune.tune(configs.c1_slab(), n=1000)
    assert best2 == best
