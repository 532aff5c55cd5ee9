"""GPU thermal scattering (POD) vs the oracle on identical states."""
import numpy as np
import pytest

from minimc_amd import abi, configs
from oracle.oracle import Oracle, SoA
from tests.gpu_util import require_gpu

pytestmark = pytest.mark.gpu
K = 8.617333262145e-11


@pytest.fixture(scope="module")
def pair():
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = configs.tsl_box(configs.tsl_tables(), T=600.0)
    return pr, Oracle(pr), Tracker(pr, max_in_flight=1 << 16)


def _states(n, seed):
    rng = np.random.default_rng(seed)
    E = np.exp(rng.uniform(np.log(1e-11), np.log(3.99e-6), n))
    T = rng.uniform(300, 1200, n)
    return SoA(n, E=E, out0=T, cell=0, u=1.0, rng=rng.integers(1, 2**31 - 1, n))


def test_tsl_sample_event_parity(pair):
    pr, o, t = pair
    s = _states(100000, 0)
    a = o.eval_event(abi.MMC_EVAL_TSL_SAMPLE, s.copy())
    b = t.eval_event(abi.MMC_EVAL_TSL_SAMPLE, s.copy())
    same = a.rng == b.rng
    assert same.mean() > 0.999
    for x, y in ((a.out0, b.out0), (a.out1, b.out1)):
        x, y = x[same], y[same]
        assert np.all(np.abs(x - y) <= 1e-11 * np.maximum(np.abs(x), 1e-6)), \
            np.max(np.abs(x - y) / np.maximum(np.abs(x), 1e-6))


def test_tsl_batch_sampler_matches_event(pair):
    pr, o, t = pair
    s = _states(200000, 1)
    E = t.device_array(s.E)
    T = t.device_array(s.out0)
    rng = t.device_array(s.rng)
    alpha = t.device_array(shape=s.n, dtype=np.float64)
    beta = t.device_array(shape=s.n, dtype=np.float64)
    t.tsl_sample_pod(E, T, rng, alpha, beta)
    a, b, r = o.tsl_sample(s.E, s.out0, s.rng, threads=8)
    same = r == rng.numpy()
    assert same.mean() > 0.999
    np.testing.assert_allclose(alpha.numpy()[same], a[same], rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(beta.numpy()[same], b[same], rtol=1e-11, atol=1e-12)


def test_tsl_total_and_collide(pair):
    pr, o, t = pair
    n = 50000
    rng = np.random.default_rng(2)
    E = np.exp(rng.uniform(np.log(1e-11), np.log(20.0), n))
    s = SoA(n, E=E, cell=0, rng=rng.integers(1, 2**31 - 1, n), u=1.0)
    a = o.eval_event(abi.MMC_EVAL_XS_TOTAL, s.copy())
    b = t.eval_event(abi.MMC_EVAL_XS_TOTAL, s.copy())
    np.testing.assert_allclose(b.out0, a.out0, rtol=1e-13)
    np.testing.assert_allclose(b.out1, a.out1, rtol=1e-13)
    a = o.eval_event(abi.MMC_EVAL_COLLIDE, s.copy())
    b = t.eval_event(abi.MMC_EVAL_COLLIDE, s.copy())
    same = a.rng == b.rng
    assert same.mean() > 0.999
    assert np.array_equal(a.event[same], b.event[same])
    np.testing.assert_allclose(b.E[same], a.E[same], rtol=1e-11)


def test_broomstick_histograms_match_oracle():
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = configs.c3_broomstick()
    n = 50000
    o = Oracle(pr)
    o.run_fixed_source(n, threads=8)
    t = Tracker(pr, max_in_flight=1 << 14)
    t.run_fixed_source(n)
    for e in range(2):   # alpha, beta histograms
        So, _ = o.tallies(e)
        St, _ = t.tallies(e)
        assert abs(So.sum() - St.sum()) <= 0.01 * So.sum() + 5
        go = So.reshape(-1, 2)[:100].sum(1) if len(So) % 2 == 0 else So
        gt = St.reshape(-1, 2)[:100].sum(1) if len(St) % 2 == 0 else St
        assert np.all(np.abs(go - gt) <= 4 * np.sqrt(go + gt + 1) + 3)
    c = t.counters()
    assert c["tsl_collisions"] > 0.5 * n


def test_tsl_index_out_of_range_is_an_error(pair):
    """A thermal-scattering table index beyond the World's tables is refused
    before any kernel reads the table descriptors."""
    from minimc_amd.tracker import MmcError
    pr, _, t = pair
    n = 16
    E, T = t.device_array(np.full(n, 1e-7)), t.device_array(np.full(n, 600.0))
    r = t.device_array(np.arange(1, n + 1, dtype=np.uint32))
    a = t.device_array(shape=n, dtype=np.float64)
    b = t.device_array(shape=n, dtype=np.float64)
    for bad in (-1, len(pr.tsl)):
        with pytest.raises(MmcError, match="no such tsl table"):
            t.tsl_sample_pod(E, T, r, a, b, tsl=bad)


def test_partition_with_an_empty_axis_is_refused():
    """A thermal-scattering partition with no CDF points is refused when the
    World is built: the samplers evaluate both ends of a CDF interval at
    in-range indices of the partition's axes."""
    require_gpu()
    import copy
    from minimc_amd.tracker import MmcError, Tracker
    t = copy.deepcopy(configs.tsl_tables())
    p = t["beta_partitions"][0]
    p["F"] = np.zeros(0)
    p["cdf_modes"] = np.zeros((0, len(p["S"])))
    with pytest.raises(MmcError, match="bad tsl partition"):
        Tracker(configs.tsl_box(t, T=600.0), max_in_flight=1 << 10)


@pytest.mark.parametrize("rank", [3, 5])
def test_sampler_and_totals_at_ranks_other_than_eight(rank):
    """The POD sampler and the thermal-scattering totals against the oracle on
    the reduced tables at ranks 3 and 5: the evaluations that read the
    coefficient rows in memory and sum over a mode count known only at run
    time (rank 8 takes the register path, tested on the full tables above).
    (At ranks 6 and 8 some of these states exceed the reference's resample
    limit on the reduced tables, an error in the oracle as on the GPU.)"""
    require_gpu()
    from minimc_amd.tracker import Tracker
    pr = configs.tsl_box(configs.tsl_tables(reduced=True, rank=rank), T=600.0)
    o, t = Oracle(pr), Tracker(pr, max_in_flight=1 << 16)
    s = _states(50000, 10 + rank)
    a = o.eval_event(abi.MMC_EVAL_TSL_SAMPLE, s.copy())
    b = t.eval_event(abi.MMC_EVAL_TSL_SAMPLE, s.copy())
    same = a.rng == b.rng
    assert same.mean() > 0.999
    for x, y in ((a.out0, b.out0), (a.out1, b.out1)):
        x, y = x[same], y[same]
        assert np.all(np.abs(x - y) <= 1e-11 * np.maximum(np.abs(x), 1e-6)), \
            np.max(np.abs(x - y) / np.maximum(np.abs(x), 1e-6))
    rng = np.random.default_rng(20 + rank)
    n = 20000
    E = np.exp(rng.uniform(np.log(1e-11), np.log(20.0), n))
    s = SoA(n, E=E, cell=0, rng=rng.integers(1, 2**31 - 1, n), u=1.0)
    a = o.eval_event(abi.MMC_EVAL_XS_TOTAL, s.copy())
    b = t.eval_event(abi.MMC_EVAL_XS_TOTAL, s.copy())
    np.testing.assert_allclose(b.out0, a.out0, rtol=1e-13)
    np.testing.assert_allclose(b.out1, a.out1, rtol=1e-13)
