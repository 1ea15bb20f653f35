"""gpu4snn_amd -- MI355X (gfx950) spike-propagation hot path of GPU4SNN.

Python view of the C ABI in include/gpu4snn.h (libgpu4snn.so, built in-tree
by gpu4snn_amd/build.py).  The library is required: there is no CPU fallback,
and every call fails loudly if the shared object or the GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

__all__ = ["ALG", "ACCUM", "Network", "Batch", "Server", "lib", "G4SError", "words_of", "raster_to_dense",
           "n_exc_of", "footprint"]

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libgpu4snn.so")
CLI_PATH = os.path.join(_PKG, "bin", "gpu4snn")

ALG = {"AP": 0, "N": 1, "S": 2, "AB": 3, "SKL": 4}
ALG_NAMES = {v: k for k, v in ALG.items()}
ACCUM = {"atomic": 0, "det": 1}

_lib = None


class G4SError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gpu4snn error {code}: {msg}")
        self.code = code


class Spec(C.Structure):
    _fields_ = [("n", C.c_int64), ("nsyn", C.c_int32), ("regime", C.c_int32), ("seed", C.c_uint64),
                ("thal_scale", C.c_float), ("reserved", C.c_int32), ("n_exc", C.c_int64)]


class RunOpts(C.Structure):
    _fields_ = [("block", C.c_int32), ("slices", C.c_int32), ("chunk", C.c_int32), ("graph", C.c_int32),
                ("kernel_events", C.c_int32)]


class Plan(C.Structure):
    _fields_ = [("alg", C.c_int32), ("accum", C.c_int32), ("opts", RunOpts), ("us_per_step", C.c_float),
                ("rate_hz", C.c_float)]

    def as_dict(self):
        return {"alg": ALG_NAMES[self.alg], "accum": {0: "atomic", 1: "det"}[self.accum],
                "block": self.opts.block, "slices": self.opts.slices, "us_per_step": self.us_per_step,
                "rate_hz": self.rate_hz}


class Timing(C.Structure):
    _fields_ = [("device_ms", C.c_double), ("wall_ms", C.c_double), ("steps", C.c_int64),
                ("spikes", C.c_int64), ("launches", C.c_int32), ("grid", C.c_int32), ("kernel_ms", C.c_double),
                ("kernel_launches", C.c_int32), ("reserved", C.c_int32)]


class LaunchInfo(C.Structure):
    _fields_ = [("kernel", C.c_char * 64), ("grid", C.c_int32), ("block", C.c_int32), ("lds_bytes", C.c_int32),
                ("vgprs", C.c_int32), ("sgprs", C.c_int32), ("waves_per_simd", C.c_int32),
                ("resident_per_cu", C.c_int32), ("cus", C.c_int32), ("persistent", C.c_int32)]


class BatchOpts(C.Structure):
    _fields_ = [("chunk", C.c_int32), ("slice_len", C.c_int32), ("graph", C.c_int32), ("schedule", C.c_int32),
                ("budget", C.c_int64), ("block", C.c_int32), ("streams", C.c_int32)]


class BatchPlan(C.Structure):
    _fields_ = [("slice_len", C.c_int32), ("copies", C.c_int32), ("us_per_inst_step", C.c_float),
                ("rate_hz", C.c_float)]


class MixPlan(C.Structure):
    _fields_ = [("buckets", C.c_int32), ("changed", C.c_int32), ("trials", C.c_int32),
                ("us_per_step_default", C.c_float), ("us_per_step_tuned", C.c_float),
                ("us_per_step_first", C.c_float)]


class InstInfo(C.Structure):
    _fields_ = [("n", C.c_int64), ("nsyn", C.c_int32), ("slice_len", C.c_int32), ("t", C.c_int64),
                ("steps", C.c_int64), ("spikes", C.c_int64), ("bytes", C.c_int64), ("done", C.c_int32),
                ("slices", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("accum", C.c_int32), ("chunk", C.c_int32), ("slice_len", C.c_int32), ("reserved", C.c_int32),
                ("budget", C.c_int64)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("error", C.c_int32), ("n", C.c_int64), ("steps", C.c_int64),
                ("t", C.c_int64), ("spikes", C.c_int64), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("submitted", "completed", "failed", "pending", "active", "chunks", "steps",
                                          "sim_steps_done", "bytes_in_use", "max_bytes", "max_active")] + \
               [(k, C.c_double) for k in ("busy_ms", "device_ms", "now_s")]


_vp = C.c_void_p
CHUNK_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_int32))


def lib():
    """Load libgpu4snn.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise G4SError(-2, f"{LIB_PATH} not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
        L = C.CDLL(LIB_PATH)
        sigs = {
            "g4s_create": ([C.POINTER(Spec), C.POINTER(_vp)], C.c_int),
            "g4s_reset": ([_vp], C.c_int),
            "g4s_set_regime": ([_vp, C.c_int, C.c_float], C.c_int),
            "g4s_run": ([_vp, C.c_int, C.c_int, C.c_int32, _vp, _vp, C.POINTER(RunOpts)], C.c_int),
            "g4s_get_timing": ([_vp, C.POINTER(Timing)], C.c_int),
            "g4s_get_step_times": ([_vp, _vp, C.c_int32], C.c_int),
            "g4s_get_launch_info": ([_vp, C.POINTER(LaunchInfo)], C.c_int),
            "g4s_kernel_resources": ([C.c_char_p, C.c_char_p, _vp], C.c_int),
            "g4s_run_stream": ([_vp, C.c_int, C.c_int, C.c_int32, CHUNK_FN, _vp, C.POINTER(RunOpts)], C.c_int),
            "g4s_run_noalloc": ([_vp, C.c_int, C.c_int, C.c_int32, _vp, _vp, C.POINTER(RunOpts)], C.c_int),
            "g4s_current_step": ([_vp], C.c_int64),
            "g4s_get_state": ([_vp, _vp, _vp, _vp, _vp], C.c_int),
            "g4s_get_params": ([_vp, _vp, _vp, _vp, _vp], C.c_int),
            "g4s_set_state": ([_vp, _vp, _vp], C.c_int),
            "g4s_get_synapses": ([_vp, _vp, _vp], C.c_int),
            "g4s_destroy": ([_vp], C.c_int),
            "g4s_last_error": ([], C.c_char_p),
            "g4s_version": ([], C.c_int),
            "g4s_build_synapses_host": ([C.POINTER(Spec), _vp, _vp, C.c_int32], C.c_int),
            "g4s_build_synapses_device": ([C.POINTER(Spec), _vp, _vp], C.c_int),
            "g4s_device_thalamic": ([C.c_uint64, C.c_uint32, C.c_int64, _vp], C.c_int),
            "g4s_device_normal_words": ([_vp, _vp, C.c_int64, _vp], C.c_int),
            "g4s_autotune": ([_vp, C.c_int, C.c_int32, C.c_int, C.POINTER(Plan)], C.c_int),
            "g4s_batch_footprint": ([C.POINTER(Spec), C.c_int32, C.c_int32], C.c_int64),
            "g4s_batch_create": ([C.c_int, C.POINTER(BatchOpts), C.POINTER(_vp)], C.c_int),
            "g4s_batch_add": ([_vp, C.POINTER(Spec), C.c_int64, C.c_int, C.POINTER(C.c_int32)], C.c_int),
            "g4s_batch_step": ([_vp, C.POINTER(C.c_int32)], C.c_int),
            "g4s_batch_run": ([_vp], C.c_int),
            "g4s_batch_info": ([_vp, C.c_int32, C.POINTER(InstInfo)], C.c_int),
            "g4s_batch_counts": ([_vp, C.c_int32, _vp], C.c_int),
            "g4s_batch_raster": ([_vp, C.c_int32, _vp], C.c_int),
            "g4s_batch_state": ([_vp, C.c_int32, _vp, _vp, _vp, _vp], C.c_int),
            "g4s_batch_remove": ([_vp, C.c_int32], C.c_int),
            "g4s_batch_timing": ([_vp, C.POINTER(Timing)], C.c_int),
            "g4s_batch_destroy": ([_vp], C.c_int),
            "g4s_batch_tune": ([C.c_int, C.POINTER(Spec), C.c_int32, C.c_int, C.POINTER(BatchPlan)], C.c_int),
            "g4s_batch_tune_mix": ([C.c_int, C.POINTER(Spec), C.c_int32, C.c_int32, C.POINTER(BatchOpts), C.c_int32,
                                    C.POINTER(MixPlan)], C.c_int),
            "g4s_server_create": ([C.POINTER(ServerOpts), C.POINTER(_vp)], C.c_int),
            "g4s_server_submit": ([_vp, C.POINTER(Spec), C.c_int64, C.c_int, C.POINTER(C.c_int64)], C.c_int),
            "g4s_server_info": ([_vp, C.c_int64, C.POINTER(RequestInfo)], C.c_int),
            "g4s_server_wait": ([_vp, C.c_int64, C.c_double], C.c_int),
            "g4s_server_drain": ([_vp, C.c_double], C.c_int),
            "g4s_server_counts": ([_vp, C.c_int64, _vp], C.c_int),
            "g4s_server_raster": ([_vp, C.c_int64, _vp], C.c_int),
            "g4s_server_release": ([_vp, C.c_int64], C.c_int),
            "g4s_server_get_stats": ([_vp, C.POINTER(ServerStats)], C.c_int),
            "g4s_server_destroy": ([_vp], C.c_int),
        }
        for name, (args, res) in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise G4SError(rc, lib().g4s_last_error().decode())
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data


def n_exc_of(n):
    """floor(4N/5): the reference's ceil(4*N/5) of an integer quotient (gpu4snn.cu:744)."""
    return (4 * int(n)) // 5


def words_of(n):
    return (int(n) + 31) // 32


def raster_to_dense(words, n):
    """[steps, ceil(n/32)] uint32 bitmap rows -> [steps, n] uint8 0/1."""
    w = np.ascontiguousarray(words, np.uint32)
    bits = np.unpackbits(w.view(np.uint8).reshape(w.shape[0], -1), axis=1, bitorder="little")
    return bits[:, :n]


def make_spec(n, nsyn, seed, regime=1, thal_scale=1.0, n_exc=-1):
    return Spec(int(n), int(nsyn), int(regime), int(seed) & (2**64 - 1), float(thal_scale), 0, int(n_exc))


def build_synapses_host(n, nsyn, seed, regime=1, thal_scale=1.0, threads=0):
    """Connectivity table built on the host without a GPU (same as g4s_create uploads)."""
    spec = make_spec(n, nsyn, seed, regime, thal_scale)
    post = np.empty(int(n) * int(nsyn), np.uint32)
    w = np.empty(int(n) * int(nsyn), np.float32)
    _check(lib().g4s_build_synapses_host(C.byref(spec), _ptr(post), _ptr(w), threads))
    return post, w


def build_synapses_device(n, nsyn, seed, regime=1, thal_scale=1.0):
    """Connectivity table built by the device builder and copied back."""
    spec = make_spec(n, nsyn, seed, regime, thal_scale)
    post = np.empty(int(n) * int(nsyn), np.uint32)
    w = np.empty(int(n) * int(nsyn), np.float32)
    _check(lib().g4s_build_synapses_device(C.byref(spec), _ptr(post), _ptr(w)))
    return post, w


def device_thalamic(seed, t, n):
    out = np.empty(int(n), np.float32)
    _check(lib().g4s_device_thalamic(seed, t, n, _ptr(out)))
    return out


def device_normal_words(x0, x1):
    x0 = np.ascontiguousarray(x0, np.uint32)
    x1 = np.ascontiguousarray(x1, np.uint32)
    out = np.empty(x0.size, np.float32)
    _check(lib().g4s_device_normal_words(_ptr(x0), _ptr(x1), x0.size, _ptr(out)))
    return out


class Network:
    """One device-resident network (g4s_create .. g4s_destroy)."""

    def __init__(self, n, nsyn, seed, regime=1, thal_scale=1.0, n_exc=-1):
        self.n, self.nsyn, self.seed = int(n), int(nsyn), int(seed)
        self.spec = make_spec(n, nsyn, seed, regime, thal_scale, n_exc)
        self._h = _vp()
        _check(lib().g4s_create(C.byref(self.spec), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().g4s_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def t(self):
        return lib().g4s_current_step(self._h)

    def reset(self):
        _check(lib().g4s_reset(self._h))

    def set_regime(self, regime, thal_scale=1.0):
        """g4s_set_regime: new regime gains x thal_scale, connectivity kept."""
        _check(lib().g4s_set_regime(self._h, int(regime), float(thal_scale)))

    def run(self, steps, alg="SKL", accum="atomic", raster=True, block=0, chunk=0, graph=0, slices=0,
            kernel_events=False):
        """Simulate `steps` steps; returns (raster words [steps, words] or None, counts)."""
        a = ALG[alg] if isinstance(alg, str) else int(alg)
        m = ACCUM[accum] if isinstance(accum, str) else int(accum)
        r = np.zeros((steps, words_of(self.n)), np.uint32) if raster else None
        counts = np.zeros(steps, np.int32)
        opts = RunOpts(block, slices, chunk, graph, int(bool(kernel_events)))
        _check(lib().g4s_run(self._h, a, m, int(steps), _ptr(r), _ptr(counts), C.byref(opts)))
        return r, counts

    def run_stream(self, steps, on_chunk, alg="AB", accum="atomic", chunk=0, block=0, slices=0):
        """g4s_run_stream: on_chunk(t0, raster_words[rows, words], counts[rows])
        is called per chunk (copies of the library's buffers) while the next
        chunk runs; returning True stops the run."""
        w = words_of(self.n)

        def cb(_user, t0, rows, raster, counts):
            r = np.ctypeslib.as_array(raster, shape=(rows * w,)).reshape(rows, w).copy()
            c = np.ctypeslib.as_array(counts, shape=(rows,)).copy()
            return 1 if on_chunk(int(t0), r, c) else 0

        fn = CHUNK_FN(cb)
        a = ALG[alg] if isinstance(alg, str) else int(alg)
        opts = RunOpts(block, slices, chunk, 0, 0)
        _check(lib().g4s_run_stream(self._h, a, ACCUM[accum], int(steps), fn, None, C.byref(opts)))

    def autotune(self, accum="atomic", trial_steps=100, force=False):
        """Fastest scheme + tiling for this network (g4s_autotune); resets the network."""
        p = Plan()
        _check(lib().g4s_autotune(self._h, ACCUM[accum], int(trial_steps), int(bool(force)), C.byref(p)))
        return p

    def run_plan(self, plan, steps, raster=True, chunk=0):
        """Run `steps` steps with a plan returned by autotune()."""
        return self.run(steps, alg=plan.alg, accum=plan.accum, raster=raster, block=plan.opts.block,
                        chunk=chunk, graph=plan.opts.graph, slices=plan.opts.slices)

    def timing(self):
        t = Timing()
        _check(lib().g4s_get_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timing._fields_}

    def step_times(self):
        """Device time of each step of the last run, ms (g4s_get_step_times)."""
        t = self.timing()
        out = np.zeros(max(0, int(t["steps"])), np.float32)
        k = lib().g4s_get_step_times(self._h, _ptr(out), out.size)
        if k < 0:
            _check(k)
        return out[:k]

    def launch_info(self):
        """Launch shape and occupancy of the last run's dominant kernel."""
        li = LaunchInfo()
        _check(lib().g4s_get_launch_info(self._h, C.byref(li)))
        d = {k: getattr(li, k) for k, _ in LaunchInfo._fields_}
        d["kernel"] = li.kernel.decode()
        return d

    def state(self):
        v, u, I = (np.empty(self.n, np.float32) for _ in range(3))
        nos = np.empty(self.n, np.int32)
        _check(lib().g4s_get_state(self._h, _ptr(v), _ptr(u), _ptr(I), _ptr(nos)))
        return v, u, I, nos

    def params(self):
        a, b, c, d = (np.empty(self.n, np.float32) for _ in range(4))
        _check(lib().g4s_get_params(self._h, _ptr(a), _ptr(b), _ptr(c), _ptr(d)))
        return a, b, c, d

    def set_state(self, v=None, u=None):
        v = None if v is None else np.ascontiguousarray(v, np.float32)
        u = None if u is None else np.ascontiguousarray(u, np.float32)
        _check(lib().g4s_set_state(self._h, _ptr(v), _ptr(u)))

    def synapses(self):
        post = np.empty(self.n * self.nsyn, np.uint32)
        w = np.empty(self.n * self.nsyn, np.float32)
        _check(lib().g4s_get_synapses(self._h, _ptr(post), _ptr(w)))
        return post, w


def kernel_resources(name, targs=""):
    """The compiler's resource report of one kernel instantiation of this
    build: waves per SIMD, VGPRs, AGPRs, SGPRs, static LDS, scratch per lane."""
    out = np.zeros(6, np.int32)
    _check(lib().g4s_kernel_resources(name.encode(), targs.encode(), out.ctypes.data))
    return dict(zip(("waves_per_simd", "vgprs", "agprs", "sgprs", "lds", "scratch"), (int(x) for x in out)))


def footprint(n, nsyn, chunk=0, slice_len=0, seed=0, regime=1, thal_scale=1.0):
    """Device bytes of one batched instance (g4s_batch_footprint)."""
    v = lib().g4s_batch_footprint(C.byref(make_spec(n, nsyn, seed, regime, thal_scale)), chunk, slice_len)
    if v < 0:
        raise G4SError(-1, lib().g4s_last_error().decode())
    return v


def batch_tune_mix(specs, accum="atomic", trial_steps=100, max_trials=40, chunk=0, streams=0, block=0):
    """g4s_batch_tune_mix on a mix of networks given as (N, Nsyn, seed,
    thal_scale) tuples (balanced regime); returns the plan summary.  The
    tuned slice lengths are recorded per bucket and used by later Batch.add
    calls of batches with slice_len 0."""
    arr = (Spec * len(specs))(*[make_spec(n, ns, sd, 1, sc) for (n, ns, sd, sc) in specs])
    opts = BatchOpts(chunk, 0, 0, 0, 0, block, streams)
    p = MixPlan()
    _check(lib().g4s_batch_tune_mix(ACCUM[accum], arr, len(specs), int(trial_steps), C.byref(opts), int(max_trials),
                                    C.byref(p)))
    return {k: getattr(p, k) for k, _ in MixPlan._fields_}


def batch_tune(n, nsyn, seed=1, accum="det", thal_scale=1.0, trial_steps=100, force=False):
    """g4s_batch_tune: slice length for the bucket of (N, Nsyn); returns a dict."""
    p = BatchPlan()
    spec = make_spec(n, nsyn, seed, 1, thal_scale)
    _check(lib().g4s_batch_tune(ACCUM[accum], C.byref(spec), int(trial_steps), int(bool(force)), C.byref(p)))
    return {k: getattr(p, k) for k, _ in BatchPlan._fields_}


class Batch:
    """Many device-resident networks advanced together (g4s_batch_*)."""

    SCHEDULES = {"step": 0, "queue": 1}

    def __init__(self, accum="det", chunk=0, slice_len=0, graph=0, budget=0, schedule="step", block=0, streams=0):
        """schedule "step": one launch per step over all instances
        (graph-replayed unless graph < 0); "queue": one persistent launch per
        chunk fed by a device work queue."""
        self._h = _vp()
        self.accum = accum
        opts = BatchOpts(chunk, slice_len, graph, self.SCHEDULES[schedule], int(budget), block, streams)
        _check(lib().g4s_batch_create(ACCUM[accum], C.byref(opts), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().g4s_batch_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, n, nsyn, seed, steps, regime=1, thal_scale=1.0, keep_raster=True, n_exc=-1):
        slot = C.c_int32()
        spec = make_spec(n, nsyn, seed, regime, thal_scale, n_exc)
        _check(lib().g4s_batch_add(self._h, C.byref(spec), int(steps), int(bool(keep_raster)), C.byref(slot)))
        return slot.value

    def step(self):
        """One chunk; returns the number of instances advanced."""
        a = C.c_int32()
        _check(lib().g4s_batch_step(self._h, C.byref(a)))
        return a.value

    def run(self):
        _check(lib().g4s_batch_run(self._h))

    def info(self, slot):
        i = InstInfo()
        _check(lib().g4s_batch_info(self._h, slot, C.byref(i)))
        return {k: getattr(i, k) for k, _ in InstInfo._fields_}

    def counts(self, slot):
        t = self.info(slot)["t"]
        c = np.zeros(t, np.int32)
        _check(lib().g4s_batch_counts(self._h, slot, _ptr(c)))
        return c

    def raster(self, slot):
        i = self.info(slot)
        w = np.zeros((i["t"], words_of(i["n"])), np.uint32)
        _check(lib().g4s_batch_raster(self._h, slot, _ptr(w)))
        return w

    def state(self, slot):
        n = self.info(slot)["n"]
        v, u, I = (np.empty(n, np.float32) for _ in range(3))
        nos = np.empty(n, np.int32)
        _check(lib().g4s_batch_state(self._h, slot, _ptr(v), _ptr(u), _ptr(I), _ptr(nos)))
        return v, u, I, nos

    def remove(self, slot):
        _check(lib().g4s_batch_remove(self._h, slot))

    def timing(self):
        t = Timing()
        _check(lib().g4s_batch_timing(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in Timing._fields_}


REQ_STATES = {0: "pending", 1: "active", 2: "done", 3: "failed"}


class Server:
    """Serving loop over a batch (g4s_server_*): submit requests from any
    thread; a worker admits them by device footprint and runs them together."""

    def __init__(self, accum="det", chunk=0, budget=0, slice_len=0):
        self._h = _vp()
        opts = ServerOpts(ACCUM[accum], chunk, slice_len, 0, int(budget))
        _check(lib().g4s_server_create(C.byref(opts), C.byref(self._h)))

    def close(self):
        if self._h:
            lib().g4s_server_destroy(self._h)
            self._h = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, n, nsyn, seed, steps, regime=1, thal_scale=1.0, keep_raster=True):
        rid = C.c_int64()
        spec = make_spec(n, nsyn, seed, regime, thal_scale)
        _check(lib().g4s_server_submit(self._h, C.byref(spec), int(steps), int(bool(keep_raster)), C.byref(rid)))
        return rid.value

    def info(self, rid):
        i = RequestInfo()
        _check(lib().g4s_server_info(self._h, rid, C.byref(i)))
        d = {k: getattr(i, k) for k, _ in RequestInfo._fields_}
        d["state"] = REQ_STATES[d["state"]]
        return d

    def wait(self, rid, timeout=0.0):
        _check(lib().g4s_server_wait(self._h, rid, float(timeout)))

    def drain(self, timeout=0.0):
        _check(lib().g4s_server_drain(self._h, float(timeout)))

    def counts(self, rid):
        c = np.zeros(self.info(rid)["t"], np.int32)
        _check(lib().g4s_server_counts(self._h, rid, _ptr(c)))
        return c

    def raster(self, rid):
        i = self.info(rid)
        w = np.zeros((i["t"], words_of(i["n"])), np.uint32)
        _check(lib().g4s_server_raster(self._h, rid, _ptr(w)))
        return w

    def release(self, rid):
        _check(lib().g4s_server_release(self._h, rid))

    def stats(self):
        st = ServerStats()
        _check(lib().g4s_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}
