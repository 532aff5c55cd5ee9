"""Synthetic continuous-temperature thermal scattering law and its POD form.

The reference consumes S(alpha, beta, T) only through proper-orthogonal-
decomposition tables (src/ThermalScattering.cpp, after Pavlou & Ji,
doi:10.1016/j.anucene.2014.04.028); the generator that produced them
(pyminimc) is not part of the reference.  This module writes such tables
from an analytic law:

  S(a, b) = exp(-b/2) S_sym(a, b)        (detailed balance)
  S_sym   = exp(-(a^2 + b^2) / 4a) / sqrt(4 pi a)              free gas
          + c a exp(-d a) [N(b; b0, w) + N(b; -b0, w)]
            / (2 sinh(b0 / 2))                                 bound mode
  with b0 = eps / kT (eps, c, d, w seeded), N a unit Gaussian density (an
  Einstein-oscillator-like one-phonon line broadened to width w).

Conventions (ThermalScattering.cpp:167-169): beta = (E' - E) / kT,
alpha = (E + E' - 2 mu sqrt(E E')) / (A kT).

Pipeline:
 1. tabulate S on 200 |beta| x 300 alpha x 24 T points;
 2. beta inverse CDFs: for each incident E and T, p(b | E) ~ integral of S
    over alpha in [a-(b), a+(b)], on beta in [-E/kT, beta_max], inverted at
    fixed CDF levels F_j;
 3. alpha inverse CDFs: for each tabulated |beta| and T, p(a | b) ~ S_sym,
    inverted at fixed CDF levels;
 4. inelastic cross section sigma(E, T) = sigma_b A kT / 4E * double
    integral of S over the kinematic domain;
 5. rank-r SVD of each table ("partitions" split the E and beta axes):
    value(F_j; x_i, T_k) = sum_r S_r U[j, r] V[(i, k), r].
Energies in MeV.
"""
import numpy as np

K_B = 8.617333262145e-11  # MeV / K
EV = 1e-6


def trapz(y, x, axis=-1):
    return np.trapezoid(y, x, axis=axis)


class Law:
    """The analytic S(alpha, beta, T)."""

    def __init__(self, seed=20261202, awr=0.99916733):
        rng = np.random.default_rng(seed)
        self.awr = awr
        self.eps = rng.uniform(0.15, 0.25) * EV   # bound-mode energy (MeV)
        self.c = rng.uniform(0.5, 1.5)
        self.d = rng.uniform(0.3, 0.8)
        self.w_rel = rng.uniform(0.08, 0.15)

    def s_sym(self, a, b, T):
        a = np.maximum(a, 1e-300)
        fg = np.exp(-(a * a + b * b) / (4 * a)) / np.sqrt(4 * np.pi * a)
        b0 = self.eps / (K_B * T)
        w = self.w_rel * b0
        g = (np.exp(-0.5 * ((b - b0) / w) ** 2) + np.exp(-0.5 * ((b + b0) / w) ** 2)) \
            / (w * np.sqrt(2 * np.pi))
        return fg + self.c * a * np.exp(-self.d * a) * g / (2 * np.sinh(0.5 * b0))

    def s(self, a, b, T):
        return np.exp(-b / 2) * self.s_sym(a, b, T)

    def alpha_bounds(self, E, b, T):
        Ep = np.maximum(E + b * K_B * T, 0.0)
        akT = self.awr * K_B * T
        return (np.sqrt(E) - np.sqrt(Ep)) ** 2 / akT, (np.sqrt(E) + np.sqrt(Ep)) ** 2 / akT


def _grids(n_beta=200, n_alpha=300, n_T=24, beta_max=160.0, alpha_max=2000.0,
           T_lo=300.0, T_hi=1200.0):
    betas = beta_max * ((np.arange(n_beta) + 1.0) / n_beta) ** 2  # > 0, ascending
    alphas = alpha_max * (np.arange(n_alpha) / (n_alpha - 1)) ** 2
    Ts = np.linspace(T_lo, T_hi, n_T)
    return betas, alphas, Ts


def _alpha_integral(law, E, b, T, n=48):
    """integral of S(a, b) da over [a-(b), a+(b)] (Gauss-Legendre in sqrt a)."""
    lo, hi = law.alpha_bounds(E, b, T)
    x, w = np.polynomial.legendre.leggauss(n)
    s_lo, s_hi = np.sqrt(lo), np.sqrt(hi)
    # a = s^2, da = 2 s ds
    s = 0.5 * (s_hi - s_lo)[..., None] * x + 0.5 * (s_hi + s_lo)[..., None]
    a = s * s
    f = law.s(a, b[..., None], T) * 2 * s
    return 0.5 * (s_hi - s_lo) * (f * w).sum(-1)


def _inverse_cdf(x, pdf, F):
    """Piecewise-linear pdf on x -> CDF by trapezoid -> x at levels F."""
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(x))])
    cdf /= cdf[-1]
    # strictly increasing for interpolation
    keep = np.concatenate([[True], np.diff(cdf) > 0])
    return np.interp(F, cdf[keep], x[keep])


def beta_tables(law, Es, Ts, F, beta_cutoff, n_beta_fine=800):
    """beta(F_j; E_i, T_k), shape (n_F, n_E, n_T)."""
    out = np.zeros((len(F), len(Es), len(Ts)))
    for k, T in enumerate(Ts):
        for i, E in enumerate(Es):
            b_min = -E / (K_B * T)
            # fine beta grid, dense near 0 and near b_min
            u = np.linspace(0.0, 1.0, n_beta_fine)
            neg = b_min * (1 - u ** 2) if b_min < 0 else np.array([])
            pos = beta_cutoff * u ** 2
            bg = np.unique(np.concatenate([neg, pos]))
            pdf = _alpha_integral(law, E, bg, T)
            out[:, i, k] = _inverse_cdf(bg, pdf, F)
    return out


def alpha_tables(law, betas, Ts, F, alpha_cutoff, n_alpha_fine=4000):
    """alpha(F_j; |beta|_i, T_k), shape (n_F, n_beta, n_T); p(a | b) ~ S_sym."""
    out = np.zeros((len(F), len(betas), len(Ts)))
    ag = alpha_cutoff * np.linspace(0.0, 1.0, n_alpha_fine) ** 2
    ag[0] = 1e-12
    for k, T in enumerate(Ts):
        for i, b in enumerate(betas):
            pdf = law.s_sym(ag, b, T)
            out[:, i, k] = _inverse_cdf(ag, pdf, F)
    return out


def inelastic_xs(law, Es, Ts, sigma_b=20.0, beta_cutoff=20.0, n_beta=800):
    """sigma(E, T) = sigma_b A kT / (4E) * int int S da db, shape (n_E, n_T)."""
    out = np.zeros((len(Es), len(Ts)))
    for k, T in enumerate(Ts):
        for i, E in enumerate(Es):
            b_min = -E / (K_B * T)
            u = np.linspace(0.0, 1.0, n_beta)
            bg = np.unique(np.concatenate([b_min * (1 - u ** 2), beta_cutoff * u ** 2]))
            inner = _alpha_integral(law, E, bg, T)
            out[i, k] = sigma_b * law.awr * K_B * T / (4 * E) * trapz(inner, bg)
    return out


def pod(table, rank):
    """table (n_F, n_x, n_T) -> (U (n_F, r), S (r), V (n_x, n_T, r))."""
    nF, nx, nT = table.shape
    U, S, Vt = np.linalg.svd(table.reshape(nF, nx * nT), full_matrices=False)
    r = min(rank, len(S))
    return (np.ascontiguousarray(U[:, :r]), np.ascontiguousarray(S[:r]),
            np.ascontiguousarray(Vt[:r].T.reshape(nx, nT, r)))


def build(seed=20261202, rank=8, n_E=96, E_min=1e-5 * EV, E_max=4.0 * EV,
          n_F_beta=200, n_F_alpha=300, n_beta=200, n_alpha=300, n_T=24,
          beta_cutoff=20.0, alpha_cutoff=2000.0, E_split=0.1 * EV, beta_split=20.0,
          sigma_b=20.0, awr=0.99916733, reduced=False):
    """Returns the TSL dict in the layout of mmc_tsl (Problem.tsl entries)
    plus the raw tables under 'raw'.  `reduced` shrinks every axis for tests."""
    if reduced:
        n_E, n_F_beta, n_F_alpha, n_beta, n_T = 24, 40, 60, 40, 6
    law = Law(seed, awr)
    betas, alphas, Ts = _grids(n_beta, n_alpha, n_T, alpha_max=alpha_cutoff)
    Es = np.exp(np.linspace(np.log(E_min), np.log(E_max), n_E))
    # CDF levels strictly inside (0, 1), denser in the tails
    def levels(n):
        u = (np.arange(n) + 0.5) / n
        return 0.5 - 0.5 * np.cos(np.pi * u)
    Fb, Fa = levels(n_F_beta), levels(n_F_alpha)
    braw = beta_tables(law, Es, Ts, Fb, beta_cutoff)
    araw = alpha_tables(law, betas, Ts, Fa, alpha_cutoff)
    xs = inelastic_xs(law, Es, Ts, sigma_b=sigma_b, beta_cutoff=beta_cutoff)
    # partitions
    bparts, aparts = [], []
    for sel in (Es < E_split, Es >= E_split):
        U, S, V = pod(braw[:, sel, :], rank)
        bparts.append(dict(F=Fb, cdf_modes=U, S=S, x=Es[sel], T=Ts, x_T_modes=V))
    for sel in (betas < beta_split, betas >= beta_split):
        U, S, V = pod(araw[:, sel, :], rank)
        aparts.append(dict(F=Fa, cdf_modes=U, S=S, x=betas[sel], T=Ts, x_T_modes=V))
    Ux, Sx, Vx = np.linalg.svd(xs, full_matrices=False)
    r = min(rank, len(Sx))
    majorant = 1.05 * xs.max(axis=1)
    return dict(
        majorant=(Es, majorant),
        total_E_axis=Es, total_E=np.ascontiguousarray(Ux[:, :r]),
        total_T_axis=Ts, total_T=np.ascontiguousarray(Vx[:r].T), total_S=Sx[:r].copy(),
        beta_partitions=bparts, alpha_partitions=aparts,
        beta_cutoff=beta_cutoff, alpha_cutoff=alpha_cutoff, awr=awr,
        raw=dict(beta=braw, alpha=araw, xs=xs, Es=Es, Ts=Ts, betas=betas, Fb=Fb, Fa=Fa,
                 alphas=alphas),
        law=law)


def save_tsl_npz(path, t):
    flat = {}
    for k in ("total_E_axis", "total_E", "total_T_axis", "total_T", "total_S"):
        flat[k] = t[k]
    flat["majorant_x"], flat["majorant_y"] = t["majorant"]
    flat["cutoffs"] = np.array([t["beta_cutoff"], t["alpha_cutoff"]])
    for name in ("beta_partitions", "alpha_partitions"):
        flat[name + "_n"] = np.array([len(t[name])])
        for i, p in enumerate(t[name]):
            for k, v in p.items():
                flat["%s_%d_%s" % (name, i, k)] = v
    np.savez(path, **flat)


def load_tsl_npz(path):
    z = np.load(path)
    t = {k: z[k] for k in ("total_E_axis", "total_E", "total_T_axis", "total_T", "total_S")}
    t["majorant"] = (z["majorant_x"], z["majorant_y"])
    t["beta_cutoff"], t["alpha_cutoff"] = (float(v) for v in z["cutoffs"])
    for name in ("beta_partitions", "alpha_partitions"):
        t[name] = []
        for i in range(int(z[name + "_n"][0])):
            t[name].append({k: z["%s_%d_%s" % (name, i, k)]
                            for k in ("F", "cdf_modes", "S", "x", "T", "x_T_modes")})
    return t
