"""Synthetic nuclear data for the benchmark problems.

The reference's own data are not available (pointwise ENDF tables and POD
coefficients ship as git-lfs pointers), so every continuous-energy problem
here runs on fictitious nuclides generated from fixed seeds:

- `ce_nuclide`: pointwise cross sections on a log-uniform energy grid,
  1/v capture + constant potential scattering + single-level Breit-Wigner
  resonances (capture and elastic channels, no interference).

Energies in MeV, cross sections in barns.
"""
import numpy as np

EV = 1e-6  # MeV


def log_grid(n=60000, e_min=1e-5 * EV, e_max=20.0):
    return np.exp(np.linspace(np.log(e_min), np.log(e_max), n))


def ce_nuclide(n=60000, n_res=200, seed=20261202, sigma_c0=0.5, sigma_pot=4.0,
               res_lo=1.0 * EV, res_hi=1e5 * EV, e_min=1e-5 * EV, e_max=20.0,
               log_peak=(0.5, 2.5), capture_share=(0.001, 0.01)):
    """Returns dict(E, capture, elastic, total) on one shared grid.

    capture = sigma_c0 * sqrt(0.0253 eV / E) + sum_r capture resonances
    elastic = sigma_pot + sum_r elastic resonances
    resonance r: peak sigma0_r at E_r, width Gamma_r, Lorentzian shape
      psi_r(E) = (Gamma_r/2)^2 / ((E - E_r)^2 + (Gamma_r/2)^2),
      capture share g_r, capture channel ~ sqrt(E_r / E) psi_r.
    """
    rng = np.random.default_rng(seed)
    E = log_grid(n, e_min, e_max)
    e_r = np.exp(rng.uniform(np.log(res_lo), np.log(res_hi), n_res))
    gamma = e_r * 10 ** rng.uniform(-3.0, -1.5, n_res)   # relative widths
    sigma0 = 10 ** rng.uniform(*log_peak, n_res)          # barns at peak
    g = rng.uniform(*capture_share, n_res)                # capture share
    capture = sigma_c0 * np.sqrt(0.0253 * EV / E)
    elastic = np.full_like(E, sigma_pot)
    for k in range(n_res):
        h = 0.5 * gamma[k]
        psi = h * h / ((E - e_r[k]) ** 2 + h * h)
        capture += sigma0[k] * g[k] * np.sqrt(e_r[k] / E) * psi
        elastic += sigma0[k] * (1 - g[k]) * psi
    total = capture + elastic
    return dict(E=E, capture=capture, elastic=elastic, total=total,
                resonances=dict(E=e_r, gamma=gamma, sigma0=sigma0, g=g))


def fissile_nuclide(base, nu=2.4, fission_fraction=0.5, seed=7):
    """Adds a fission channel: sigma_f = fission_fraction * capture(E) of
    `base` (1/v plus resonances), nubar constant."""
    E = base["E"]
    fission = fission_fraction * base["capture"]
    capture = (1 - fission_fraction) * base["capture"]
    out = dict(base)
    out.update(capture=capture, fission=fission,
               total=capture + base["elastic"] + fission,
               nubar=(np.array([E[0], E[-1]]), np.array([nu, nu])))
    return out
