"""The FP64 MLP oracle (oracle/mlp_ref.py) against central finite
differences, and the surrogate's host-side helpers."""
import numpy as np

from minimc_amd import surrogate
from oracle import mlp_ref


def test_gradient_matches_finite_differences():
    widths = (4, 7, 5, 2)
    rng = np.random.default_rng(0)
    n_p = surrogate.n_params(widths)
    p = rng.normal(size=n_p) * 0.5
    x = rng.normal(size=(13, 4))
    y = rng.normal(size=(13, 2))
    loss, g = mlp_ref.loss_and_grad(p, widths, x, y)
    eps = 1e-6
    for i in rng.choice(n_p, 40, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (mlp_ref.loss_and_grad(pp, widths, x, y)[0] -
              mlp_ref.loss_and_grad(pm, widths, x, y)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd)
    # loss definition: sum of squares / (2 n n_out)
    out = mlp_ref.forward(p, widths, x)
    assert loss == np.sum((out - y) ** 2) / (2 * 13 * 2)


def test_parameter_count_and_layout():
    assert surrogate.n_params() == 133378   # 4-256-256-256-2
    p = np.arange(surrogate.n_params(), dtype=np.float64)
    layers = mlp_ref.unpack(p, surrogate.WIDTHS)
    assert layers[0][0].shape == (256, 4) and layers[0][0][0, 1] == 1.0
    assert layers[0][1][0] == 1024.0
    assert layers[-1][0].shape == (2, 256)


def test_input_sampling_box():
    x = surrogate.sample_inputs(100000, 3)
    assert x[:, 0].min() >= np.log(1e-11) - 1e-5 and x[:, 0].max() <= np.log(4e-6) + 1e-5
    assert x[:, 1].min() >= 300 and x[:, 1].max() <= 1200
    assert 0 <= x[:, 2].min() and x[:, 3].max() < 1


def test_loss_weights_equalise_bin_variance():
    """Training weights w = var_global / var_bin (clipped, mean 1): on a
    target whose spread differs 4x between temperature halves, w * var_bin
    is the same in every bin."""
    from minimc_amd import surrogate
    x = surrogate.sample_inputs(400000, 3)
    rng = np.random.default_rng(4)
    scale = np.where(x[:, 1] > 750.0, 4.0, 1.0)
    y = np.stack([rng.normal(0, 1, len(x)) * scale, rng.normal(0, 1, len(x))], 1)
    tab = surrogate.weight_table(x, y.astype(np.float32))
    b = surrogate.weight_bins(x)
    assert np.allclose(tab[b].mean(0), 1.0)
    cnt = np.bincount(b)
    var0 = np.bincount(b, y[:, 0] ** 2) / cnt
    prod = tab[:, 0] * var0
    assert prod.max() / prod.min() < 1.1  # equalised (sampling noise only)
    assert tab[:, 1].max() / tab[:, 1].min() < 1.35  # nothing to equalise (noise only)


def test_alpha_range_kinematics():
    """alpha_min, alpha_max = (sqrt E -+ sqrt E')^2 / (A kT), E' = E + beta kT:
    their product is (beta / A)^2 and their sum 2 (E + E') / (A kT)."""
    rng = np.random.default_rng(2)
    E = np.exp(rng.uniform(np.log(1e-11), np.log(4e-6), 1000))
    T = rng.uniform(300, 1200, 1000)
    kT = surrogate.K_B * T
    beta = rng.uniform(-E / kT, 20.0)
    A = 0.99916733
    y1 = beta * kT / (E + kT)
    lo, hi = surrogate.alpha_range(E, T, y1, A)
    Ep = E + beta * kT
    assert np.allclose(lo * hi, (beta / A) ** 2, rtol=1e-9, atol=1e-30)
    assert np.allclose(lo + hi, 2 * (E + Ep) / (A * kT), rtol=1e-9)
    assert np.all(lo <= hi)


def test_local_spread_interpolation_is_bilinear():
    """local_spread_at interpolates node tables bilinearly in (ln E, T): a
    table linear in both node indices is reproduced exactly between nodes,
    and the nodes themselves are hit."""
    nE, nT = 6, 5
    i, j = np.meshgrid(np.arange(nE), np.arange(nT), indexing="ij")
    t1 = 2.0 + 0.5 * i + 0.25 * j
    t2 = 1.0 + i * 0.1 + j * 0.3
    rng = np.random.default_rng(3)
    x = np.zeros((500, 4), dtype=np.float64)
    fe = rng.uniform(0, nE - 1, 500)
    ft = rng.uniform(0, nT - 1, 500)
    lo, hi = np.log(surrogate.E_LO), np.log(surrogate.E_HI)
    x[:, 0] = lo + fe / (nE - 1) * (hi - lo)
    x[:, 1] = surrogate.T_LO + ft / (nT - 1) * (surrogate.T_HI - surrogate.T_LO)
    s1, s2 = surrogate.local_spread_at((t1, t2), x)
    assert np.allclose(s1, 2.0 + 0.5 * fe + 0.25 * ft, rtol=1e-9)
    assert np.allclose(s2, 1.0 + 0.1 * fe + 0.3 * ft, rtol=1e-9)
    xn = np.zeros((2, 4))
    xn[:, 0] = [lo, hi]
    xn[:, 1] = [surrogate.T_LO, surrogate.T_HI]
    s1, _ = surrogate.local_spread_at((t1, t2), xn)
    assert np.allclose(s1, [t1[0, 0], t1[-1, -1]])
