"""GPU vs oracle on continuous-energy physics (identical particle states).

Grid search + lin-lin interpolation use no libm: bit-exact.  Free-gas
physics goes through log/sqrt/cos/erf/exp: floats within 1e-12 relative and
RNG states equal wherever both sides took the same rejection path.
"""
import math

import numpy as np
import pytest

from minimc_amd import abi, configs, synth
from oracle.oracle import Oracle, SoA
from tests.gpu_util import require_gpu
from tests.test_oracle_continuous import one_cell_problem

pytestmark = pytest.mark.gpu
K = 8.617333262145e-11


@pytest.fixture(scope="module")
def data():
    return synth.ce_nuclide()  # the full 60,000-point grid


def _tracker(pr, **kw):
    require_gpu()
    from minimc_amd.tracker import Tracker
    return Tracker(pr, **kw)


def test_xs_lookup_bit_exact(data):
    pr = one_cell_problem(configs.ce_tables(data))
    o, t = Oracle(pr), _tracker(pr, max_in_flight=1 << 12)
    E = data["E"]
    rng = np.random.default_rng(3)
    q = np.concatenate([E[::7], E[1::997] * (1 + 1e-16),
                        np.exp(rng.uniform(np.log(E[0]) - 1, np.log(E[-1]) + 1, 200000))])
    s = SoA(len(q), E=q, cell=0)
    a = o.eval_event(abi.MMC_EVAL_XS_TOTAL, s.copy())
    b = t.eval_event(abi.MMC_EVAL_XS_TOTAL, s.copy())
    assert np.array_equal(a.out0, b.out0)
    assert np.array_equal(a.out1, b.out1)


def test_xs_free_gas_adjusted(data):
    pr = one_cell_problem(configs.ce_tables(data, with_total=False, awr=0.99), T=1100.0)
    o, t = Oracle(pr), _tracker(pr, max_in_flight=1 << 12)
    q = np.exp(np.random.default_rng(4).uniform(np.log(1e-11), np.log(20.0), 100000))
    s = SoA(len(q), E=q, cell=0)
    a = o.eval_event(abi.MMC_EVAL_XS_TOTAL, s.copy())
    b = t.eval_event(abi.MMC_EVAL_XS_TOTAL, s.copy())
    np.testing.assert_allclose(b.out0, a.out0, rtol=1e-12)


def test_collide_free_gas(data):
    pr = one_cell_problem(configs.ce_tables(data), T=293.6)
    o, t = Oracle(pr), _tracker(pr, max_in_flight=1 << 12)
    n = 50000
    rng = np.random.default_rng(5)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    s = SoA(n, E=np.exp(rng.uniform(np.log(1e-10), np.log(2.0), n)), u=d[:, 0],
            v=d[:, 1], w=d[:, 2], cell=0, rng=rng.integers(1, 2**31 - 1, n))
    a = o.eval_event(abi.MMC_EVAL_COLLIDE, s.copy())
    b = t.eval_event(abi.MMC_EVAL_COLLIDE, s.copy())
    same = a.rng == b.rng
    assert same.mean() > 0.999
    assert np.array_equal(a.event[same], b.event[same])
    for f in ("E", "u", "v", "w"):
        x, y = getattr(a, f)[same], getattr(b, f)[same]
        assert np.all(np.abs(x - y) <= 1e-12 * np.maximum(np.abs(x), 1e-300) + 1e-15), f


def test_free_gas_equilibrium_gpu():
    E = synth.log_grid(200, 1e-11, 20.0)
    nuc = dict(awr=1.0, reactions=[dict(kind=abi.MMC_RX_SCATTER,
                                        xs=(E, np.full_like(E, 20.0)), temperature=293.6)])
    t = _tracker(one_cell_problem(nuc, T=293.6), max_in_flight=1 << 12)
    n = 100000
    rng = np.random.default_rng(1)
    s = SoA(n, E=1e-6, u=1.0, cell=0, rng=rng.integers(1, 2**31 - 1, n))
    means = []
    for k in range(120):
        s = t.eval_event(abi.MMC_EVAL_COLLIDE, s)
        if k >= 60:
            means.append(s.E.mean())
    assert abs(np.mean(means) / (K * 293.6) - 1.75) < 0.015


def test_ce_slab_leakage_matches_oracle():
    pr = configs.c2_slab(n_points=60000, n_res=200)
    n = 20000
    o = Oracle(pr)
    o.run_fixed_source(n, threads=8)
    t = _tracker(pr, max_in_flight=1 << 13)
    t.run_fixed_source(n)
    tot_o = sum(o.tallies(e)[0].sum() for e in range(2))
    tot_t = sum(t.tallies(e)[0].sum() for e in range(2))
    # identical streams: trajectories agree except after rare libm branch flips
    assert abs(tot_o - tot_t) <= 0.01 * n
    for e in range(2):
        So, _ = o.tallies(e)
        St, S2t = t.tallies(e)
        # coarse spectrum (10 groups of 30 bins) within 4 sigma of each other
        go, gt = So[1:301].reshape(10, 30).sum(1), St[1:301].reshape(10, 30).sum(1)
        assert np.all(np.abs(go - gt) <= 4 * np.sqrt(go + gt + 1) + 2)
        assert np.all(S2t >= St - 1e-9)
    c = t.counters()
    assert c["histories"] == n and c["collisions"] > 10 * n


def test_two_grid_slab_matches_oracle():
    """Two nuclides on different energy grids, each in half of a slab: the
    event kernels stage only the first grid's bucket hash in LDS, so the
    second nuclide's lookups must use its own hash."""
    pr = configs.c2_slab(n_points=60000, n_res=200)
    other = synth.ce_nuclide(n=24000, n_res=80, seed=11, e_min=2e-11, e_max=19.0)
    pr.nuclides.append(configs.ce_tables(other, awr=25.0))
    s_mid = pr.add_surface("mid", abi.MMC_PLANEX, a=0.5)
    # split the slab cell at x = 0.5: left keeps nuclide 0, right gets nuclide 1
    m1 = pr.add_material("other", pr.materials[0]["aden"], [(1, 1.0)])
    s0, s1 = 0, 1  # the slab's left and right planes
    pr.cells[0] = dict(pr.cells[0], faces=[(s0, False), (s_mid, True)])
    pr.add_cell("right", m1, [(s_mid, False), (s1, True)])
    n = 20000
    o = Oracle(pr)
    o.run_fixed_source(n, threads=8)
    t = _tracker(pr, max_in_flight=1 << 13)
    t.run_fixed_source(n)
    tot_o = sum(o.tallies(e)[0].sum() for e in range(2))
    tot_t = sum(t.tallies(e)[0].sum() for e in range(2))
    assert abs(tot_o - tot_t) <= 0.01 * n
    for e in range(2):
        So, _ = o.tallies(e)
        St, _ = t.tallies(e)
        go, gt = So[1:301].reshape(10, 30).sum(1), St[1:301].reshape(10, 30).sum(1)
        assert np.all(np.abs(go - gt) <= 4 * np.sqrt(go + gt + 1) + 2)
