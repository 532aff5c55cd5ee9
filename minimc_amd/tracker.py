"""Python binding of libminimc_amd.so (the HIP tracker's C ABI).

`Tracker(problem)` mirrors the reference drivers: `run_fixed_source` is
FixedSource::Solve's history loop (src/FixedSource.cpp:22-77), `keig_init`
and `run_cycle` the k-eigenvalue cycle body the reference leaves
unimplemented (src/KEigenvalue.cpp:36-62).  There is no CPU fallback: if the
library is missing or no device is present, construction raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libminimc_amd.so")
_lib = None

_SYMBOLS = {
    "mmc_abi_version": ([], C.c_int),
    "mmc_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "mmc_create": ([C.POINTER(abi.WorldDesc), C.POINTER(abi.TallyDesc),
                    C.POINTER(abi.Options), C.POINTER(C.c_void_p)], C.c_int),
    "mmc_destroy": ([C.c_void_p], None),
    "mmc_last_error": ([C.c_void_p], C.c_char_p),
    "mmc_run_fixed_source": ([C.c_void_p, C.POINTER(abi.SourceDesc), C.c_uint64,
                              C.c_uint64, C.c_uint64], C.c_int),
    "mmc_keig_init": ([C.c_void_p, C.POINTER(abi.SourceDesc), C.c_uint64], C.c_int),
    "mmc_run_cycle": ([C.c_void_p, C.c_int, C.POINTER(abi.CycleStats)], C.c_int),
    "mmc_get_source": ([C.c_void_p, C.POINTER(abi.Site), C.c_uint64,
                        C.POINTER(C.c_uint64)], C.c_int),
    "mmc_set_source": ([C.c_void_p, C.POINTER(abi.Site), C.c_uint64], C.c_int),
    "mmc_set_entropy_mesh": ([C.c_void_p, C.POINTER(C.c_int32), abi.c_double_p,
                              abi.c_double_p], C.c_int),
    "mmc_tally_size": ([C.c_void_p, C.c_int, C.POINTER(C.c_uint64)], C.c_int),
    "mmc_get_tallies": ([C.c_void_p, C.c_int, abi.c_double_p, abi.c_double_p,
                         C.c_uint64], C.c_int),
    "mmc_reset_tallies": ([C.c_void_p], C.c_int),
    "mmc_get_counters": ([C.c_void_p, C.POINTER(abi.Counters)], C.c_int),
    "mmc_set_profiling": ([C.c_void_p, C.c_int], C.c_int),
    "mmc_get_kernel_times": ([C.c_void_p, abi.c_double_p, C.POINTER(C.c_uint64)], C.c_int),
    "mmc_eval_event": ([C.c_void_p, C.c_int, C.POINTER(abi.ParticleSoA), C.c_uint64],
                       C.c_int),
    "mmc_sample_source": ([C.c_void_p, C.POINTER(abi.SourceDesc), C.c_uint64, C.c_uint64,
                           C.POINTER(abi.ParticleSoA)], C.c_int),
    "mmc_device_malloc": ([C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)], C.c_int),
    "mmc_device_free": ([C.c_void_p, C.c_void_p], C.c_int),
    "mmc_memcpy_h2d": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64], C.c_int),
    "mmc_memcpy_d2h": ([C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64], C.c_int),
    "mmc_tsl_sample_pod": ([C.c_void_p, C.c_int, C.c_uint64, C.c_void_p, C.c_void_p,
                            C.c_void_p, C.c_void_p, C.c_void_p], C.c_int),
    "mmc_mlp_set": ([C.c_void_p, C.c_int, C.POINTER(abi.MlpDesc), abi.c_float_p,
                     C.c_uint64], C.c_int),
    "mmc_mlp_get": ([C.c_void_p, abi.c_float_p, C.c_uint64], C.c_int),
    "mmc_mlp_infer": ([C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p], C.c_int),
    "mmc_tsl_sample_surrogate": ([C.c_void_p, C.c_int, C.c_uint64, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p],
                                 C.c_int),
    "mmc_tsl_pod_map": ([C.c_void_p, C.c_int, C.c_uint64, C.c_void_p, C.c_void_p], C.c_int),
    "mmc_mlp_normalize": ([C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p], C.c_int),
    "mmc_mlp_train": ([C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p,
                       C.POINTER(abi.TrainOpts), C.POINTER(C.c_float)], C.c_int),
    "mmc_mlp_grad": ([C.c_void_p, C.c_uint64, abi.c_float_p, abi.c_float_p,
                      C.POINTER(C.c_float), abi.c_float_p], C.c_int),
}


def lib():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libminimc_amd.so not built: run `python -m minimc_amd.build` "
                "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in _SYMBOLS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def device_count():
    n = C.c_int(0)
    lib().mmc_device_count(C.byref(n))
    return n.value


class DeviceArray:
    """A device buffer owned by a Tracker's context (the library's own HIP
    runtime; no second runtime is loaded into the process)."""

    def __init__(self, tracker, shape, dtype):
        self.t = tracker
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        tracker._check(lib().mmc_device_malloc(tracker.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, tracker, a):
        a = np.ascontiguousarray(a)
        d = cls(tracker, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        self.t._check(lib().mmc_memcpy_h2d(self.t.h, self.ptr, a.ctypes.data, self.nbytes))

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        self.t._check(lib().mmc_memcpy_d2h(self.t.h, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def numel(self):
        return int(np.prod(self.shape))

    def data_ptr(self):
        return self.ptr

    def free(self):
        if self.ptr and self.t.h:
            lib().mmc_device_free(self.t.h, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("minimc_amd error %d: %s" % (code, msg))
        self.code = code


class Tracker:
    """GPU tracker for one Problem (one context on one device)."""

    def __init__(self, problem, tracking=None, tsl_mode=abi.MMC_TSL_POD, device=0,
                 max_in_flight=0, graph_iters=0, max_steps=0, tail_threshold=0,
                 compact_interval=0, graphs=0, tail_block=0, tsl_block=0, tail_network=0,
                 split_collide=0):
        L = lib()
        self.problem = problem
        self._w, self._t, self._keep = problem.descriptors()
        opt = abi.Options()
        opt.tracking = problem.tracking if tracking is None else tracking
        opt.tsl_mode = tsl_mode
        opt.device = device
        opt.max_in_flight = max_in_flight
        opt.graph_iters = graph_iters
        opt.max_steps = max_steps
        opt.tail_threshold = tail_threshold
        opt.compact_interval = compact_interval
        opt.graphs = graphs
        opt.tail_block = tail_block
        opt.tsl_block = tsl_block
        opt.tail_network = tail_network
        opt.split_collide = split_collide
        self.options = opt
        h = C.c_void_p()
        rc = L.mmc_create(C.byref(self._w), C.byref(self._t), C.byref(opt), C.byref(h))
        self.h = h if h.value else None
        if rc:
            msg = L.mmc_last_error(self.h).decode() if self.h else "no device"
            self.close()
            raise MmcError(rc, msg)

    def close(self):
        if getattr(self, "h", None):
            lib().mmc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise MmcError(rc, lib().mmc_last_error(self.h).decode())

    # -- drivers ------------------------------------------------------------
    def run_fixed_source(self, n, seed0=None, first=0, src=None):
        sd = self.problem.source_desc(src)
        seed0 = self.problem.seed if seed0 is None else seed0
        self._check(lib().mmc_run_fixed_source(self.h, C.byref(sd), seed0, first, n))

    def set_entropy_mesh(self, n, lo, hi):
        na = (C.c_int32 * 3)(*n)
        lo = np.asarray(lo, dtype=np.float64)
        hi = np.asarray(hi, dtype=np.float64)
        self._check(lib().mmc_set_entropy_mesh(self.h, na, abi.dptr(lo), abi.dptr(hi)))

    def keig_init(self, n, src=None):
        sd = self.problem.source_desc(src)
        self._check(lib().mmc_keig_init(self.h, C.byref(sd), n))

    def run_cycle(self, active):
        st = abi.CycleStats()
        self._check(lib().mmc_run_cycle(self.h, 1 if active else 0, C.byref(st)))
        return st.as_dict()

    def get_source(self):
        n = C.c_uint64()
        self._check(lib().mmc_get_source(self.h, None, 0, C.byref(n)))
        arr = (abi.Site * n.value)()
        self._check(lib().mmc_get_source(self.h, arr, n.value, C.byref(n)))
        return arr

    def set_source(self, sites):
        self._check(lib().mmc_set_source(self.h, sites, len(sites)))

    def tallies(self, e):
        n = C.c_uint64()
        self._check(lib().mmc_tally_size(self.h, e, C.byref(n)))
        S, S2 = np.zeros(n.value), np.zeros(n.value)
        self._check(lib().mmc_get_tallies(self.h, e, abi.dptr(S), abi.dptr(S2), n.value))
        return S, S2

    def reset_tallies(self):
        self._check(lib().mmc_reset_tallies(self.h))

    def counters(self):
        c = abi.Counters()
        self._check(lib().mmc_get_counters(self.h, C.byref(c)))
        return c.as_dict()

    def set_profiling(self, on=True):
        self._check(lib().mmc_set_profiling(self.h, 1 if on else 0))

    KERNELS = ("advance", "cross", "collide", "tsl", "mlp", "tail")

    def kernel_times(self):
        ms = np.zeros(len(self.KERNELS))
        n = (C.c_uint64 * len(self.KERNELS))()
        self._check(lib().mmc_get_kernel_times(self.h, abi.dptr(ms), n))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(self.KERNELS)}

    # -- single events ------------------------------------------------------
    def eval_event(self, kind, soa):
        s = soa.struct()
        self._check(lib().mmc_eval_event(self.h, kind, C.byref(s), soa.n))
        return soa

    def sample_source(self, seed0, n, soa, src=None):
        sd = self.problem.source_desc(src)
        s = soa.struct()
        self._check(lib().mmc_sample_source(self.h, C.byref(sd), seed0, n, C.byref(s)))
        return soa

    # -- device batches (DeviceArray or any object with numel/data_ptr) -------
    def device_array(self, a=None, shape=None, dtype=None):
        if a is not None:
            return DeviceArray.from_numpy(self, a)
        return DeviceArray(self, shape, dtype)

    def tsl_sample_pod(self, E, T, rng, alpha, beta, tsl=0):
        self._check(lib().mmc_tsl_sample_pod(
            self.h, tsl, E.numel(), E.data_ptr(), T.data_ptr(), rng.data_ptr(),
            alpha.data_ptr(), beta.data_ptr()))

    def tsl_sample_surrogate(self, E, T, rng, alpha, beta, tsl=0):
        self._check(lib().mmc_tsl_sample_surrogate(
            self.h, tsl, E.numel(), E.data_ptr(), T.data_ptr(), rng.data_ptr(),
            alpha.data_ptr(), beta.data_ptr()))
