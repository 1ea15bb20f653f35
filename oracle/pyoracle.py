"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never as the product.
See g4s_oracle.h for what each function restates from the reference
(Crypto-TII/GPU4SNN gpu4snn.cu) and how it is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


class OrcNet(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("n_exc", C.c_int64), ("nsyn", C.c_int32), ("seed", C.c_uint64),
        ("ge", C.c_float), ("gi", C.c_float),
        ("post", C.c_void_p), ("w", C.c_void_p),
        ("a", C.c_void_p), ("b", C.c_void_p), ("c", C.c_void_p), ("d", C.c_void_p),
    ]


def build() -> str:
    """Compile liboracle.so from g4s_oracle.c (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_u01.restype = C.c_float
        L.orc_normal.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_normal.restype = C.c_float
        L.orc_thalamic.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_thalamic.restype = C.c_float
        L.orc_logf_many.argtypes = [_f32p, _f32p, C.c_int64]
        L.orc_sqrtf_many.argtypes = [_f32p, _f32p, C.c_int64]
        L.orc_cos2pi_many.argtypes = [_u32p, _f32p, C.c_int64]
        L.orc_normal_many.argtypes = [_u32p, _u32p, _f32p, C.c_int64]
        L.orc_thalamic_many.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, _f32p, C.c_int64]
        L.orc_philox_many.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, _u32p, C.c_int64]
        L.orc_gains.argtypes = [C.c_int, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_init_params.argtypes = [C.c_uint64, C.c_int64, C.c_int64] + [_f32p] * 6
        L.orc_build_synapses.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_int32, _u32p, _f32p]
        L.orc_build_synapses.restype = C.c_int
        L.orc_simulate.argtypes = [C.POINTER(OrcNet), C.c_int32, C.c_int32, _f32p, _f32p, _f32p,
                                   _i32p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_simulate.restype = C.c_int
        L.orc_simulate_order.argtypes = list(L.orc_simulate.argtypes) + [C.c_int32]
        L.orc_simulate_order.restype = C.c_int
        L.orc_simulate_f64_forced.argtypes = [C.POINTER(OrcNet), C.c_int32, C.c_int32, _u8p, _i64p,
                                              C.c_int64, _f64p, _f64p, C.c_void_p, C.c_void_p]
        L.orc_simulate_f64_forced.restype = C.c_int
        L.orc_simulate_mt.argtypes = list(L.orc_simulate.argtypes) + [C.c_int32]
        L.orc_simulate_mt.restype = C.c_int
        L.orc_simulate_f64_forced_ranges.argtypes = [C.POINTER(OrcNet), C.c_int32, C.c_int32, C.c_void_p, C.c_int64,
                                                     _i64p, _i64p, C.c_int32, _f64p, _f64p, C.c_void_p, C.c_void_p]
        L.orc_simulate_f64_forced_ranges.restype = C.c_int
        _lib = L
    return _lib


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def philox_many(seed, a0, b, tag, n):
    out = np.zeros(4 * n, np.uint32)
    lib().orc_philox_many(seed, a0, b, tag, out, n)
    return out.reshape(n, 4)


def logf(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().orc_logf_many(x, out, x.size)
    return out


def sqrtf(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().orc_sqrtf_many(x, out, x.size)
    return out


def cos2pi(x):
    x = np.ascontiguousarray(x, np.uint32)
    out = np.empty(x.size, np.float32)
    lib().orc_cos2pi_many(x, out, x.size)
    return out


def normal(x0, x1):
    x0 = np.ascontiguousarray(x0, np.uint32)
    x1 = np.ascontiguousarray(x1, np.uint32)
    out = np.empty(x0.size, np.float32)
    lib().orc_normal_many(x0, x1, out, x0.size)
    return out


def thalamic(seed, t, n, i0=0):
    out = np.empty(n, np.float32)
    lib().orc_thalamic_many(seed, t, i0, out, n)
    return out


def gains(regime, scale=1.0):
    ge, gi = C.c_float(), C.c_float()
    lib().orc_gains(regime, scale, C.byref(ge), C.byref(gi))
    return ge.value, gi.value


def n_exc_of(n):
    """floor(4N/5): the reference's ceil(4*N/5) over an integer quotient (cu:744)."""
    return (4 * n) // 5


class Network:
    """One network as the oracle sees it: parameters, state, synapse table."""

    def __init__(self, n, nsyn, seed, regime=1, thal_scale=1.0, n_exc=None, table=None):
        """table: optional (post, w) arrays to use instead of building them
        (e.g. a table already checked equal to orc_build_synapses')."""
        self.n, self.nsyn, self.seed = int(n), int(nsyn), int(seed)
        self.n_exc = n_exc_of(self.n) if n_exc is None else int(n_exc)
        self.ge, self.gi = gains(regime, thal_scale)
        L = lib()
        self.a, self.b, self.c, self.d, self.v, self.u = (np.empty(self.n, np.float32) for _ in range(6))
        L.orc_init_params(self.seed, self.n, self.n_exc, self.a, self.b, self.c, self.d, self.v, self.u)
        if table is not None:
            self.post = np.ascontiguousarray(table[0], np.uint32)
            self.w = np.ascontiguousarray(table[1], np.float32)
            assert self.post.size == self.n * self.nsyn and self.w.size == self.post.size
        else:
            self.post = np.empty(self.n * self.nsyn, np.uint32)
            self.w = np.empty(self.n * self.nsyn, np.float32)
            rc = L.orc_build_synapses(self.seed, self.n, self.n_exc, self.nsyn, self.post, self.w)
            if rc != 0:
                raise ValueError(f"orc_build_synapses failed ({rc}) for N={n} Nsyn={nsyn}")
        self.isyn = np.zeros(self.n, np.float32)
        self.nospks = np.zeros(self.n, np.int32)
        self.t = 0

    def _cnet(self):
        return OrcNet(self.n, self.n_exc, self.nsyn, self.seed, self.ge, self.gi,
                      self.post.ctypes.data, self.w.ctypes.data, self.a.ctypes.data,
                      self.b.ctypes.data, self.c.ctypes.data, self.d.ctypes.data)

    def run(self, steps, raster=True, threads=1, descending=False):
        """Advance `steps` steps; returns (raster uint8[steps, n] or None, counts, lastI).
        threads != 1 runs the multithreaded form (same results bit for bit;
        <= 0: one thread per core).  descending=True adds each step's
        spiking rows in descending presynaptic order (single thread)."""
        r = np.zeros((steps, self.n), np.uint8) if raster else None
        counts = np.zeros(steps, np.int32)
        last = np.zeros(self.n, np.float32)
        cn = self._cnet()
        if descending:
            rc = lib().orc_simulate_order(C.byref(cn), self.t, steps, self.v, self.u, self.isyn, self.nospks,
                                          r.ctypes.data if raster else None, counts.ctypes.data, last.ctypes.data, 1)
            if rc != 0:
                raise RuntimeError(f"orc_simulate_order failed ({rc})")
            self.t += steps
            return r, counts, last
        if threads != 1:
            rc = lib().orc_simulate_mt(C.byref(cn), self.t, steps, self.v, self.u, self.isyn, self.nospks,
                                       r.ctypes.data if raster else None, counts.ctypes.data, last.ctypes.data,
                                       int(threads))
            if rc != 0:
                raise RuntimeError(f"orc_simulate_mt failed ({rc})")
            self.t += steps
            return r, counts, last
        rc = lib().orc_simulate(C.byref(cn), self.t, steps, self.v, self.u, self.isyn, self.nospks,
                                r.ctypes.data if raster else None, counts.ctypes.data, last.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"orc_simulate failed ({rc})")
        self.t += steps
        return r, counts, last

    def run_f64_forced_ranges(self, forced_words, ranges, t0, v0, u0):
        """fp64 teacher-forced integration of the neurons in `ranges` (list of
        ascending [lo, hi)) forced by a bit-packed raster [steps, words]."""
        fw = np.ascontiguousarray(forced_words, np.uint32)
        steps = fw.shape[0]
        lo = np.array([r[0] for r in ranges], np.int64)
        hi = np.array([r[1] for r in ranges], np.int64)
        nsub = int((hi - lo).sum())
        v = np.ascontiguousarray(v0, np.float64).copy()
        u = np.ascontiguousarray(u0, np.float64).copy()
        assert v.size == nsub and u.size == nsub
        tr = np.zeros((steps, nsub), np.float64)
        own = np.zeros((steps, nsub), np.uint8)
        rc = lib().orc_simulate_f64_forced_ranges(C.byref(self._cnet()), t0, steps, fw.ctypes.data, fw.shape[1],
                                                  lo, hi, len(ranges), v, u, tr.ctypes.data, own.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"orc_simulate_f64_forced_ranges failed ({rc})")
        return tr, own

    def run_f64_forced(self, forced_raster, sub, t0, v0, u0):
        """fp64 teacher-forced integration of neurons `sub` from state (v0, u0) at step t0."""
        steps = forced_raster.shape[0]
        sub = np.ascontiguousarray(sub, np.int64)
        v = np.ascontiguousarray(v0, np.float64).copy()
        u = np.ascontiguousarray(u0, np.float64).copy()
        tr = np.zeros((steps, sub.size), np.float64)
        own = np.zeros((steps, sub.size), np.uint8)
        cn = self._cnet()
        rc = lib().orc_simulate_f64_forced(C.byref(cn), t0, steps,
                                           np.ascontiguousarray(forced_raster, np.uint8), sub, sub.size,
                                           v, u, tr.ctypes.data, own.ctypes.data)
        if rc != 0:
            raise RuntimeError(f"orc_simulate_f64_forced failed ({rc})")
        return tr, own
