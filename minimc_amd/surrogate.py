"""Training and use of the thermal-scattering surrogate (the learned sampler).

The network maps (ln E, T, u1, u2) to (y1, y2) = (beta kT / (E + kT),
position of alpha in [alpha_min(beta), alpha_max(beta)]).  Its training
targets come from the POD tables themselves: `mmc_tsl_pod_map` evaluates the
POD inverse CDFs of beta (level u1) and alpha (level u2) on the device.  The
forward and backward passes and the Adam steps are the library's own FP32
kernels; inference is two BF16-MFMA kernels.  Outside the trained box
(E outside [e_min, e_max] or T outside [t_lo, t_hi]) and where the network's
outgoing energy leaves the box (E' < e_min, including E' <= 0), the tracker
falls back to POD.
"""
import ctypes as C
import hashlib
import os

import numpy as np

from . import abi
from .tracker import lib

WIDTHS = (4, 256, 256, 256, 2)
E_LO, E_HI = 1e-11, 4e-6        # MeV: 1e-5 eV .. 4 eV
T_LO, T_HI = 300.0, 1200.0      # K


def n_params(widths=WIDTHS):
    return sum(widths[i + 1] * widths[i] + widths[i + 1] for i in range(len(widths) - 1))


def make_desc(out_shift=(0.0, 0.5), out_scale=(1.0, 0.3), widths=WIDTHS,
              e_min=E_LO, e_max=E_HI, t_lo=T_LO, t_hi=T_HI):
    d = abi.MlpDesc()
    d.n_layers = len(widths) - 1
    for i, w in enumerate(widths):
        d.widths[i] = w
    lo = np.array([np.log(E_LO), T_LO, 0.0, 0.0])
    hi = np.array([np.log(E_HI), T_HI, 1.0, 1.0])
    for k in range(4):
        d.in_shift[k] = 0.5 * (lo[k] + hi[k])
        d.in_scale[k] = 2.0 / (hi[k] - lo[k])
    for k in range(2):
        d.out_shift[k] = out_shift[k]
        d.out_scale[k] = out_scale[k]
    d.e_min, d.e_max, d.t_lo, d.t_hi = e_min, e_max, t_lo, t_hi
    return d


# Loss weights: the sampler's accuracy is assessed per (E, T) (Wasserstein distance
# relative to the POD marginal's spread there), and that spread varies about
# 3x over the box (the beta target is narrow above a few kT).  Training
# therefore weights each sample by var_global / var_bin of its target, with
# variances per bin of a WEIGHT_BINS (ln E x T) grid measured on the first
# chunk of targets, clipped to WEIGHT_CLIP and normalised to mean 1.
WEIGHT_BINS = (32, 8)
WEIGHT_CLIP = (0.2, 20.0)


def weight_bins(x):
    nE, nT = WEIGHT_BINS
    fe = (x[:, 0] - np.log(E_LO)) / (np.log(E_HI) - np.log(E_LO))
    ft = (x[:, 1] - T_LO) / (T_HI - T_LO)
    ie = np.clip((fe * nE).astype(np.int64), 0, nE - 1)
    it = np.clip((ft * nT).astype(np.int64), 0, nT - 1)
    return ie * nT + it


def weight_table(x, y):
    """[bins x 2] weights var_global / var_bin from raw inputs and targets."""
    b = weight_bins(x)
    nb = WEIGHT_BINS[0] * WEIGHT_BINS[1]
    cnt = np.maximum(np.bincount(b, minlength=nb), 1)
    tab = np.ones((nb, 2))
    for k in range(2):
        m1 = np.bincount(b, y[:, k], minlength=nb) / cnt
        m2 = np.bincount(b, y[:, k].astype(np.float64) ** 2, minlength=nb) / cnt
        var = np.maximum(m2 - m1 * m1, 1e-12)
        tab[:, k] = np.clip(y[:, k].var() / var, *WEIGHT_CLIP)
        tab[:, k] /= (tab[b, k]).mean()
    return tab


# Local loss weights (the default): the assessment measures, at one (E, T),
# the Wasserstein-1 distance of the beta and alpha marginals relative to
# their standard deviations there.  To first order that is the mean
# |error| of the map in beta (alpha) units over that local spread, so each
# sample's squared error is weighted by 1 / (local spread)^2 of what it
# predicts: for y1 (affine in beta at fixed E, T) the spread of y1 itself;
# for y2 (alpha = alpha_min + y2 (alpha_max - alpha_min)) the factor
# (alpha_max - alpha_min)^2 / var(alpha).  The spreads are measured with the
# POD map on a LOCAL_GRID (ln E x T) grid of nodes, LOCAL_LEVELS (u1 x u2)
# stratified levels per node, and interpolated bilinearly.
LOCAL_GRID = (96, 32)
LOCAL_LEVELS = (64, 16)
LOCAL_CLIP = (0.02, 50.0)
K_B = 8.617333262145e-11  # MeV / K


def tsl_awr(problem, tsl):
    """The atomic weight ratio of the nuclide whose scattering uses thermal-
    scattering table `tsl` (the kinematics of alpha)."""
    for n in problem.nuclides:
        for r in n.get("reactions", []):
            if r.get("tsl", -1) == tsl:
                return float(n["awr"])
    raise ValueError("no nuclide scatters with thermal-scattering table %d" % tsl)


def alpha_range(E, T, y1, awr):
    """Kinematic alpha bounds for the beta that y1 encodes (E in MeV)."""
    kT = K_B * T
    beta = y1 * (E + kT) / kT
    root = np.sqrt(np.maximum(E + beta * kT, 0.0))
    se = np.sqrt(E)
    akT = awr * kT
    return (se - root) ** 2 / akT, (se + root) ** 2 / akT


def _bilinear(tab, fe, ft):
    """tab [nE x nT] at fractional node coordinates fe, ft (clipped)."""
    nE, nT = tab.shape
    fe = np.clip(fe, 0, nE - 1 - 1e-9)
    ft = np.clip(ft, 0, nT - 1 - 1e-9)
    i, j = fe.astype(np.int64), ft.astype(np.int64)
    a, b = fe - i, ft - j
    return ((1 - a) * (1 - b) * tab[i, j] + a * (1 - b) * tab[i + 1, j] +
            (1 - a) * b * tab[i, j + 1] + a * b * tab[i + 1, j + 1])


def local_spread_at(tabs, x):
    """Local standard deviations (of y1 and of alpha) at raw inputs x."""
    sd1, sda = tabs
    nE, nT = sd1.shape
    fe = (x[:, 0].astype(np.float64) - np.log(E_LO)) / (np.log(E_HI) - np.log(E_LO)) * (nE - 1)
    ft = (x[:, 1].astype(np.float64) - T_LO) / (T_HI - T_LO) * (nT - 1)
    return _bilinear(sd1, fe, ft), _bilinear(sda, fe, ft)


# Fraction of training energies drawn uniformly in E (the rest log-uniform):
# the log-uniform draw alone leaves few samples in the upper decade, where
# the beta law is widest.
E_UNIFORM_FRACTION = 0.3


def sample_inputs(n, seed, e_uniform=E_UNIFORM_FRACTION):
    """Raw inputs: E from a mixture of log-uniform (1 - e_uniform) and
    uniform (e_uniform) on [1e-5, 4) eV, T uniform on [300, 1200] K, u1, u2
    uniform on [0, 1)."""
    rng = np.random.default_rng(seed)
    x = np.empty((n, 4), dtype=np.float32)
    lnE = rng.uniform(np.log(E_LO), np.log(E_HI), n)
    lin = rng.random(n) < e_uniform
    lnE[lin] = np.log(rng.uniform(E_LO, E_HI, int(lin.sum())))
    x[:, 0] = lnE
    x[:, 1] = rng.uniform(T_LO, T_HI, n)
    x[:, 2] = rng.uniform(0.0, 1.0, n)
    x[:, 3] = rng.uniform(0.0, 1.0, n)
    return x


class Surrogate:
    """Owns the network on a Tracker's context."""

    def __init__(self, tracker, tsl=0):
        self.t = tracker
        self.tsl = tsl
        self.desc = None

    def _check(self, rc):
        self.t._check(rc)

    def pod_map(self, x_dev, n):
        y = self.t.device_array(shape=(n, 2), dtype=np.float32)
        self._check(lib().mmc_tsl_pod_map(self.t.h, self.tsl, n, x_dev.data_ptr(),
                                          y.data_ptr()))
        return y

    def set(self, params=None, desc=None):
        self.desc = desc if desc is not None else make_desc()
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        self._check(lib().mmc_mlp_set(self.t.h, self.tsl, C.byref(self.desc),
                                      None if p is None else p.ctypes.data_as(abi.c_float_p),
                                      n_params() if p is None else p.size))

    def params(self):
        p = np.zeros(n_params(), dtype=np.float32)
        self._check(lib().mmc_mlp_get(self.t.h, p.ctypes.data_as(abi.c_float_p), p.size))
        return p

    def grad(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = C.c_float()
        g = np.zeros(n_params(), dtype=np.float32)
        self._check(lib().mmc_mlp_grad(self.t.h, len(x), x.ctypes.data_as(abi.c_float_p),
                                       y.ctypes.data_as(abi.c_float_p), C.byref(loss),
                                       g.ctypes.data_as(abi.c_float_p)))
        return loss.value, g

    def infer(self, x_dev, n):
        y = self.t.device_array(shape=(n, 2), dtype=np.float32)
        self._check(lib().mmc_mlp_infer(self.t.h, n, x_dev.data_ptr(), y.data_ptr()))
        return y

    def awr(self):
        return tsl_awr(self.t.problem, self.tsl)

    def local_spread(self, grid=LOCAL_GRID, levels=LOCAL_LEVELS):
        """[nE x nT] tables of the standard deviation of y1 and of alpha under
        the POD map, at the nodes of a grid over ln E and T."""
        nE, nT = grid
        n1, n2 = levels
        lnE = np.linspace(np.log(E_LO), np.log(E_HI), nE)
        T = np.linspace(T_LO, T_HI, nT)
        u1 = (np.arange(n1) + 0.5) / n1
        u2 = (np.arange(n2) + 0.5) / n2
        g = np.stack(np.meshgrid(lnE, T, u1, u2, indexing="ij"), -1).reshape(-1, 4)
        x = np.ascontiguousarray(g, dtype=np.float32)
        xd = self.t.device_array(x)
        y = self.pod_map(xd, len(x))
        yh = y.numpy().astype(np.float64)
        xd.free()
        y.free()
        E = np.exp(x[:, 0].astype(np.float64))
        Tx = x[:, 1].astype(np.float64)
        amin, amax = alpha_range(E, Tx, yh[:, 0], self.awr())
        alpha = amin + yh[:, 1] * (amax - amin)
        sd1 = yh[:, 0].reshape(nE, nT, -1).std(-1)
        sda = alpha.reshape(nE, nT, -1).std(-1)
        return np.maximum(sd1, 1e-12), np.maximum(sda, 1e-30)

    def local_weights(self, tabs, x, y):
        """Unnormalised per-sample loss weights [n x 2] (see LOCAL_GRID)."""
        s1, sa = local_spread_at(tabs, x)
        E = np.exp(x[:, 0].astype(np.float64))
        amin, amax = alpha_range(E, x[:, 1].astype(np.float64), y[:, 0].astype(np.float64),
                                 self.awr())
        return np.stack([1.0 / s1 ** 2, ((amax - amin) / sa) ** 2], -1)

    def bin_errors(self, n=1 << 20, seed=7919, tabs=None):
        """Per (ln E, T) weight bin and output: the mean absolute error of the
        network (BF16 inference) against the POD map on n fresh inputs,
        relative to the spread of the target in that bin.  For a network
        monotone in u this is the bin's Wasserstein-1 error of the marginal,
        relative to its standard deviation.  With local spread tables `tabs`
        the errors are relative to the local spreads instead (alpha units for
        the second output)."""
        x = sample_inputs(n, seed)
        xd = self.t.device_array(x)
        y = self.pod_map(xd, n)
        p = self.infer(xd, n)
        yh, ph = y.numpy().astype(np.float64), p.numpy().astype(np.float64)
        for a in (xd, y, p):
            a.free()
        b = weight_bins(x)
        nb = WEIGHT_BINS[0] * WEIGHT_BINS[1]
        cnt = np.maximum(np.bincount(b, minlength=nb), 1)
        err = np.zeros((nb, 2))
        if tabs is not None:
            rel = np.abs(ph - yh) * np.sqrt(self.local_weights(tabs, x, yh))
            for k in range(2):
                err[:, k] = np.bincount(b, rel[:, k], minlength=nb) / cnt
            return err
        for k in range(2):
            m1 = np.bincount(b, yh[:, k], minlength=nb) / cnt
            m2 = np.bincount(b, yh[:, k] ** 2, minlength=nb) / cnt
            sd = np.sqrt(np.maximum(m2 - m1 * m1, 1e-24))
            err[:, k] = np.bincount(b, np.abs(ph[:, k] - yh[:, k]), minlength=nb) / cnt / sd
        return err

    def train(self, n_samples=1 << 26, batch=8192, epochs=3, lr=2e-3, seed=11,
              chunk=1 << 24, log=None, weighted=True, rounds=1, gamma=1.0, lr_floor=1e-5,
              weighting="local", e_uniform=E_UNIFORM_FRACTION):
        """Generates n_samples POD targets on the device and trains.  Loss
        weights (unless weighted is False): weighting "local" weights each
        sample by the inverse local variance of what it predicts (LOCAL_GRID),
        "bins" by var_global / var_bin over WEIGHT_BINS.  With rounds > 1 the
        epochs are split into rounds, and after each round but the last the
        weights of every (ln E, T) bin are multiplied by (its relative error
        / the mean relative error)^gamma (bin_errors): the training effort
        moves to where the network is worst, which is what a bound on the
        largest error needs.  e_uniform: the fraction of training energies
        drawn uniformly in E (sample_inputs)."""
        n = n_samples
        x = self.t.device_array(shape=(n, 4), dtype=np.float32)
        y = self.t.device_array(shape=(n, 2), dtype=np.float32)
        w = self.t.device_array(shape=(n, 2), dtype=np.float32) if weighted else None
        tabs = self.local_spread() if weighted and weighting == "local" else None
        wloc = np.empty((n, 2), dtype=np.float32) if tabs is not None else None
        nb = WEIGHT_BINS[0] * WEIGHT_BINS[1]
        mult = np.ones((nb, 2))

        def upload(c0, wc):
            wc = np.ascontiguousarray(wc, dtype=np.float32)
            lib().mmc_memcpy_h2d(self.t.h, w.ptr + c0 * 8, wc.ctypes.data, wc.nbytes)

        # inputs in chunks (host RNG), targets from the POD map on the device
        stats, wtab, b0, norm, scale = [], None, None, None, None
        for c0 in range(0, n, chunk):
            m = min(chunk, n - c0)
            xc = sample_inputs(m, seed + c0, e_uniform)
            lib().mmc_memcpy_h2d(self.t.h, x.ptr + c0 * 16, xc.ctypes.data, xc.nbytes)
            self._check(lib().mmc_tsl_pod_map(self.t.h, self.tsl, m, x.ptr + c0 * 16,
                                              y.ptr + c0 * 8))
            if c0 == 0 or tabs is not None:
                yc = np.empty((m, 2), dtype=np.float32)
                lib().mmc_memcpy_d2h(self.t.h, yc.ctypes.data, y.ptr + c0 * 8, yc.nbytes)
            if c0 == 0:
                stats = (yc.mean(0), yc.std(0) + 1e-6)
                b0 = weight_bins(xc)
                if weighted and tabs is None:
                    wtab = weight_table(xc, yc)
            if tabs is not None:
                wl = self.local_weights(tabs, xc, yc)
                if norm is None:
                    norm = np.median(wl, axis=0)
                wl = np.clip(wl / norm, *LOCAL_CLIP)
                if scale is None:
                    scale = 1.0 / wl.mean(0)
                wl = wl * scale
                wloc[c0:c0 + m] = wl
                upload(c0, wl)
            elif weighted:
                upload(c0, wtab[weight_bins(xc)])
        desc = make_desc(out_shift=tuple(float(v) for v in stats[0]),
                         out_scale=tuple(float(v) for v in stats[1]))
        self.set(None, desc)
        self._check(lib().mmc_mlp_normalize(self.t.h, n, x.data_ptr(), y.data_ptr()))
        steps_per_epoch = n // batch
        losses = []
        per_round = max(1, epochs // max(rounds, 1))
        for ep in range(epochs):
            if weighted and rounds > 1 and ep > 0 and ep % per_round == 0 and \
                    ep // per_round < rounds:
                err = self.bin_errors(tabs=tabs)
                if log:
                    worst = np.argsort(-err.max(1))[:3]
                    nT = WEIGHT_BINS[1]
                    log("round %d: largest bin error beta %.4f alpha %.4f; worst bins %s" % (
                        ep // per_round, err[:, 0].max(), err[:, 1].max(),
                        ", ".join("E %.3g eV T %.0f K: %.4f/%.4f" % (
                            1e6 * np.exp(np.log(E_LO) + (b // nT + 0.5) / WEIGHT_BINS[0] *
                                         (np.log(E_HI) - np.log(E_LO))),
                            T_LO + (b % nT + 0.5) / nT * (T_HI - T_LO), err[b, 0], err[b, 1])
                                  for b in worst)))
                tab = mult if tabs is not None else wtab
                for k in range(2):
                    tab[:, k] *= (err[:, k] / err[:, k].mean()) ** gamma
                    tab[:, k] = np.clip(tab[:, k], 0.05, 50.0)
                    tab[:, k] /= tab[b0, k].mean()
                for c0 in range(0, n, chunk):
                    m = min(chunk, n - c0)
                    bins = weight_bins(sample_inputs(m, seed + c0, e_uniform))
                    if tabs is not None:
                        upload(c0, wloc[c0:c0 + m] * mult[bins])
                    else:
                        upload(c0, wtab[bins])
            # cosine-decayed learning rate in 8 chunks per epoch
            for part in range(8):
                frac = (ep * 8 + part) / (epochs * 8)
                o = abi.TrainOpts()
                o.batch = batch
                o.steps = steps_per_epoch // 8
                o.lr = float(lr * 0.5 * (1 + np.cos(np.pi * frac)) + lr_floor)
                o.beta1, o.beta2, o.eps = 0.9, 0.999, 1e-8
                o.offset = (part * (steps_per_epoch // 8)) * batch
                o.weights = w.ptr if weighted else None
                loss = C.c_float()
                self._check(lib().mmc_mlp_train(self.t.h, n, x.data_ptr(), y.data_ptr(),
                                                C.byref(o), C.byref(loss)))
                losses.append(loss.value)
                if log:
                    log("epoch %d part %d lr %.2e loss %.3e" % (ep, part, o.lr, loss.value))
        x.free()
        y.free()
        if w is not None:
            w.free()
        return losses

    def save(self, path, info=None):
        """Writes the network and the digest of the thermal-scattering data
        it was trained on (plus free-form training information)."""
        d = self.desc
        np.savez(path, params=self.params(), in_shift=np.array(d.in_shift),
                 in_scale=np.array(d.in_scale), out_shift=np.array(d.out_shift),
                 out_scale=np.array(d.out_scale),
                 tables_digest=np.array(tables_digest(self.t.problem.tsl[self.tsl])),
                 info=np.array(repr(info or {})))

    def load(self, path):
        """Loads a saved network; raises ValueError if it was trained on other
        thermal-scattering data than this tracker's."""
        z = np.load(path, allow_pickle=False)
        want = tables_digest(self.t.problem.tsl[self.tsl])
        if "tables_digest" not in z or str(z["tables_digest"]) != want:
            raise ValueError("%s was trained on other thermal-scattering data" % path)
        desc = make_desc(out_shift=tuple(z["out_shift"]), out_scale=tuple(z["out_scale"]))
        self.set(z["params"], desc)


def _feed(h, v):
    if isinstance(v, dict):
        for k in sorted(v, key=str):
            h.update(str(k).encode())
            _feed(h, v[k])
    elif isinstance(v, (list, tuple)):
        for a in v:
            _feed(h, a)
    elif isinstance(v, np.ndarray):
        h.update(np.ascontiguousarray(v).tobytes())
    else:
        h.update(repr(v).encode())


def tables_digest(t):
    """SHA-1 of a thermal-scattering table set (nested dicts/lists of arrays
    and numbers)."""
    h = hashlib.sha1()
    _feed(h, t)
    return h.hexdigest()


# The network trained for the synthetic thermal-scattering tables of the
# benchmark configurations (configs.tsl_tables()), shipped with the package.
SHIPPED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "tsl_surrogate.npz")


def cached_network(tracker, tsl=0, path=None, log=None, **train_kw):
    """Loads the network at `path` (default: the shipped one), training it
    on the device first when it is missing or was trained on other tables.
    Returns (Surrogate, trained_now)."""
    path = path or SHIPPED
    s = Surrogate(tracker, tsl)
    try:
        s.load(path)
        return s, False
    except (OSError, ValueError):
        pass
    s.train(log=log, **train_kw)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    s.save(path, info=dict(train_kw))
    return s, True
