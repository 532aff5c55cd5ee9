"""Training and use of the thermal-scattering surrogate (the learned sampler).

The network maps (ln E, T, u1, u2) to (y1, y2) = (beta kT / (E + kT),
position of alpha in [alpha_min(beta), alpha_max(beta)]).  Its training
targets come from the POD tables themselves: `mmc_tsl_pod_map` evaluates the
POD inverse CDFs of beta (level u1) and alpha (level u2) on the device.  The
forward and backward passes and the Adam steps are the library's own FP32
kernels; inference is the fused BF16-MFMA kernel.  Outside the trained box
(E outside [e_min, e_max] or T outside [t_lo, t_hi]) and where the network's
energy transfer leaves the kinematic support (E' <= 0), the tracker falls back
to POD.
"""
import ctypes as C
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


def sample_inputs(n, seed):
    """Raw inputs: E log-uniform on [1e-5, 4) eV, T uniform on [300, 1200] K,
    u1, u2 uniform on [0, 1)."""
    rng = np.random.default_rng(seed)
    x = np.empty((n, 4), dtype=np.float32)
    x[:, 0] = rng.uniform(np.log(E_LO), np.log(E_HI), n)
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

    def train(self, n_samples=1 << 26, batch=8192, epochs=3, lr=2e-3, seed=11,
              chunk=1 << 24, log=None):
        """Generates n_samples POD targets on the device and trains."""
        n = n_samples
        x = self.t.device_array(shape=(n, 4), dtype=np.float32)
        y = self.t.device_array(shape=(n, 2), dtype=np.float32)
        # inputs in chunks (host RNG), targets from the POD map on the device
        stats = []
        for c0 in range(0, n, chunk):
            m = min(chunk, n - c0)
            xc = sample_inputs(m, seed + c0)
            lib().mmc_memcpy_h2d(self.t.h, x.ptr + c0 * 16, xc.ctypes.data, xc.nbytes)
            self._check(lib().mmc_tsl_pod_map(self.t.h, self.tsl, m, x.ptr + c0 * 16,
                                              y.ptr + c0 * 8))
            if c0 == 0:
                yc = np.empty((m, 2), dtype=np.float32)
                lib().mmc_memcpy_d2h(self.t.h, yc.ctypes.data, y.ptr, yc.nbytes)
                stats = (yc.mean(0), yc.std(0) + 1e-6)
        desc = make_desc(out_shift=tuple(float(v) for v in stats[0]),
                         out_scale=tuple(float(v) for v in stats[1]))
        self.set(None, desc)
        self._check(lib().mmc_mlp_normalize(self.t.h, n, x.data_ptr(), y.data_ptr()))
        steps_per_epoch = n // batch
        losses = []
        for ep in range(epochs):
            # cosine-decayed learning rate in 8 chunks per epoch
            for part in range(8):
                frac = (ep * 8 + part) / (epochs * 8)
                o = abi.TrainOpts()
                o.batch = batch
                o.steps = steps_per_epoch // 8
                o.lr = float(lr * 0.5 * (1 + np.cos(np.pi * frac)) + 1e-5)
                o.beta1, o.beta2, o.eps = 0.9, 0.999, 1e-8
                o.offset = (part * (steps_per_epoch // 8)) * batch
                loss = C.c_float()
                self._check(lib().mmc_mlp_train(self.t.h, n, x.data_ptr(), y.data_ptr(),
                                                C.byref(o), C.byref(loss)))
                losses.append(loss.value)
                if log:
                    log("epoch %d part %d lr %.2e loss %.3e" % (ep, part, o.lr, loss.value))
        x.free()
        y.free()
        return losses

    def save(self, path):
        d = self.desc
        np.savez(path, params=self.params(), in_shift=np.array(d.in_shift),
                 in_scale=np.array(d.in_scale), out_shift=np.array(d.out_shift),
                 out_scale=np.array(d.out_scale))

    def load(self, path):
        z = np.load(path)
        desc = make_desc(out_shift=tuple(z["out_shift"]), out_scale=tuple(z["out_scale"]))
        self.set(z["params"], desc)


def cached_network(tracker, tsl=0, path=None, log=None, **train_kw):
    """Loads the trained network from minimc_amd/_build, training it first
    when it is not there."""
    path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build",
                                "tsl_surrogate.npz")
    s = Surrogate(tracker, tsl)
    if os.path.exists(path):
        s.load(path)
    else:
        s.train(log=log, **train_kw)
        s.save(path)
    return s
