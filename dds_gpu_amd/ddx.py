"""Python mirror of the ddx C ABI (include/ddx.h), through ctypes.

This is glue for tests, the bench and scripts; the solver itself is libddx.so. Nothing here
falls back to Python or CPU code when the library is missing: every call raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libddx.so")
# Build variants of the same C ABI: "" the default exact TT entries, "sig" 12-byte signature
# entries (the reference's -DDD_TT_SIG build), "shared" one lock-free TT for all searches (its
# -DDD_SHARED_TT build).
LIB_PATHS = {"": LIB_PATH, "sig": os.path.join(_HERE, "lib", "libddx_sig.so"),
             "shared": os.path.join(_HERE, "lib", "libddx_shared.so")}

SEED_PLAIN, SEED_STEP, SEED_CERT, SEED_CERTP = 0, 1, 2, 3
SCHED_PRIORITY, SCHED_CHAIN, SCHED_SPLIT, SCHED_THREAD = 0, 1, 2, 3
E_ARG, E_STOPPED = -1, -5

_libs = {}


class Stats(ctypes.Structure):
    _fields_ = [("nodes", ctypes.c_uint64), ("tt_probes", ctypes.c_uint64),
                ("tt_hits", ctypes.c_uint64), ("jobs_done", ctypes.c_uint64),
                ("seconds", ctypes.c_double), ("tt_entry_reads", ctypes.c_uint64),
                ("tt_stores", ctypes.c_uint64), ("tt_line_cache_hits", ctypes.c_uint64),
                ("splits", ctypes.c_uint64), ("tasks_taken", ctypes.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GpuConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("lanes", ctypes.c_int64), ("tt_log2", ctypes.c_int),
                ("jobs_per_deal", ctypes.c_int), ("schedule", ctypes.c_int),
                ("seed_mode", ctypes.c_int), ("block", ctypes.c_int), ("profile", ctypes.c_int),
                ("team", ctypes.c_int), ("graph", ctypes.c_int), ("waves_per_simd", ctypes.c_int),
                ("lds_lines", ctypes.c_int), ("split_min_nodes", ctypes.c_int)]


class DdxError(RuntimeError):
    def __init__(self, rc, msg):
        super().__init__("ddx error %d: %s" % (rc, msg))
        self.rc = rc


def lib(variant=""):
    """Loads libddx.so, or a build variant of it (built by dds_gpu_amd.build); raises if it is
    missing."""
    if variant not in _libs:
        if variant not in LIB_PATHS:
            raise ValueError("unknown library build %r (one of %s)" % (variant, sorted(LIB_PATHS)))
        path = LIB_PATHS[variant]
        if not os.path.exists(path):
            raise RuntimeError("%s is not built: run `python -m dds_gpu_amd.build` "
                               "(expected %s)" % (os.path.basename(path), path))
        L = ctypes.CDLL(path)
        P, I64, U64, I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int
        L.ddx_last_error.restype = ctypes.c_char_p
        L.ddx_version.restype = ctypes.c_char_p
        L.ddx_random_deals.argtypes = [U64, I64, I64, I, P]
        L.ddx_pgx_read.argtypes = [ctypes.c_char_p, I64, I64, P, P, P]
        L.ddx_pgx_rows.argtypes = [ctypes.c_char_p, P]
        L.ddx_pgx_write.argtypes = [ctypes.c_char_p, P, P, I64]
        L.ddx_cpu_create.argtypes = [I, P]
        L.ddx_cpu_destroy.argtypes = [P]
        L.ddx_cpu_solve_strain.argtypes = [P, P, I, P]
        L.ddx_cpu_solve_strain_guess.argtypes = [P, P, I, P, P]
        L.ddx_cpu_ns_tricks.argtypes = [P, P, I, I, I, P]
        L.ddx_cpu_nodes.argtypes = [P]
        L.ddx_cpu_nodes.restype = U64
        L.ddx_cpu_solve_batch.argtypes = [P, I64, I, I, I, I, P, P, P, P]
        L.ddx_gpu_default_config.argtypes = [P]
        L.ddx_gpu_open.argtypes = [P, P]
        L.ddx_gpu_close.argtypes = [P]
        L.ddx_gpu_device_bytes.argtypes = [P]
        L.ddx_gpu_device_bytes.restype = U64
        L.ddx_gpu_solve.argtypes = [P, P, I64, P, P, P]
        L.ddx_gpu_solve_async.argtypes = [P, P, I64, P, P, P]
        L.ddx_gpu_wait.argtypes = [P, P]
        L.ddx_gpu_request_stop.argtypes = [P]
        L.ddx_gpu_last_kernel_ms.argtypes = [P]
        L.ddx_gpu_last_kernel_ms.restype = ctypes.c_double
        L.ddx_gpu_profile.argtypes = [P, P, I]
        L.ddx_gpu_progress.argtypes = [P, P]
        L.ddx_stream_open.argtypes = [P, I, P]
        L.ddx_stream_push.argtypes = [P, P, P, I64]
        L.ddx_stream_poll.argtypes = [P, I, P, P, P]
        L.ddx_stream_in_flight.argtypes = [P]
        L.ddx_stream_in_flight.restype = I64
        L.ddx_stream_finish.argtypes = [P]
        L.ddx_stream_close.argtypes = [P, P]
        L.ddx_mlp_create.argtypes = [I, U64, P]
        L.ddx_mlp_destroy.argtypes = [P]
        L.ddx_mlp_param_count.restype = I64
        L.ddx_mlp_params.argtypes = [P, P, I]
        L.ddx_mlp_grads.argtypes = [P, P, P, I64, P, P]
        L.ddx_mlp_train.argtypes = [P, P, P, I64, I, I, ctypes.c_float, U64, P]
        L.ddx_mlp_predict.argtypes = [P, P, I64, P, P]
        L.ddx_mlp_predict_host.argtypes = [P, P, I64, P]
        L.ddx_mlp_predict_noalloc.argtypes = [P, P, I64, P, P]
        L.ddx_mlp_save.argtypes = [P, ctypes.c_char_p]
        L.ddx_mlp_load.argtypes = [P, ctypes.c_char_p]
        _libs[variant] = L
    return _libs[variant]


def _check(rc, ok=(0,), L=None):
    if rc not in ok:
        raise DdxError(rc, (L or lib()).ddx_last_error().decode())
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def random_deals(seed, n, k=13, first=0):
    """Deals [first, first+n) of seeded stream `seed`, k cards per hand: (n, 4) uint64."""
    h = np.zeros((n, 4), np.uint64)
    _check(lib().ddx_random_deals(seed, first, n, k, _ptr(h)))
    return h


def pgx_read(path, start=0, n=-1):
    """Hands (n, 4) uint64 and stored tables (n, 20) uint8 of a Pgx .npy file."""
    rows = ctypes.c_int64()
    _check(lib().ddx_pgx_rows(path.encode(), ctypes.byref(rows)))
    count = rows.value - start if n < 0 else n
    hands = np.zeros((count, 4), np.uint64)
    tables = np.zeros((count, 20), np.uint8)
    got = ctypes.c_int64()
    _check(lib().ddx_pgx_read(path.encode(), start, count, _ptr(hands), _ptr(tables), ctypes.byref(got)))
    return hands, tables


def pgx_write(path, hands, tables=None):
    hands = np.ascontiguousarray(hands, np.uint64)
    t = None if tables is None else np.ascontiguousarray(tables, np.uint8)
    _check(lib().ddx_pgx_write(path.encode(), _ptr(hands), _ptr(t), len(hands)))


def cpu_solve(hands, threads=1, tt_log2=10, jobs_per_deal=5, seed_mode=SEED_PLAIN, seeds=None,
              job_nodes=False, variant=""):
    """Tables (n, 20) of the CPU solver; with job_nodes (threads == 1) also the nodes per job."""
    hands = np.ascontiguousarray(hands, np.uint64)
    n = len(hands)
    out = np.zeros((n, 20), np.uint8)
    nodes = np.zeros(n * jobs_per_deal, np.uint64) if job_nodes else None
    s = None if seeds is None else np.ascontiguousarray(seeds, np.uint8)
    st = Stats()
    L = lib(variant)
    _check(L.ddx_cpu_solve_batch(_ptr(hands), n, threads, tt_log2, jobs_per_deal, seed_mode,
                                 _ptr(s), _ptr(out), _ptr(nodes), ctypes.byref(st)), L=L)
    return (out, st, nodes) if job_nodes else (out, st)


class CpuContext:
    """One CPU solver context (reference DDCtx + its TT)."""

    def _ck(self, rc, ok=(0,)):
        return _check(rc, ok, self._L)

    def __init__(self, tt_log2=12, variant=""):
        self._L = lib(variant)
        self.h = ctypes.c_void_p()
        self._ck(self._L.ddx_cpu_create(tt_log2, ctypes.byref(self.h)))

    def solve_strain(self, hands, strain, ns4=None):
        hands = np.ascontiguousarray(hands, np.uint64)
        out = np.zeros(4, np.int32)
        if ns4 is None:
            self._ck(self._L.ddx_cpu_solve_strain(self.h, _ptr(hands), strain, _ptr(out)))
        else:
            g = np.ascontiguousarray(ns4, np.int32)
            self._ck(self._L.ddx_cpu_solve_strain_guess(self.h, _ptr(hands), strain, _ptr(g), _ptr(out)))
        return out

    def ns_tricks(self, hands, strain, leader, guess=7):
        hands = np.ascontiguousarray(hands, np.uint64)
        v = ctypes.c_int()
        self._ck(self._L.ddx_cpu_ns_tricks(self.h, _ptr(hands), strain, leader, guess, ctypes.byref(v)))
        return v.value

    @property
    def nodes(self):
        return self._L.ddx_cpu_nodes(self.h)

    def close(self):
        if self.h:
            self._L.ddx_cpu_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GpuSolver:
    """A set of solver lanes on one GPU (ddx_gpu_open / ddx_gpu_solve)."""

    def _ck(self, rc, ok=(0,)):
        return _check(rc, ok, self._L)

    def __init__(self, lanes=65536, tt_log2=10, jobs_per_deal=5, schedule=SCHED_CHAIN,
                 seed_mode=SEED_PLAIN, block=256, profile=False, device=0, team=1, graph=False,
                 waves_per_simd=0, lds_lines=0, split_min_nodes=0, variant=""):
        self._L = lib(variant)
        cfg = GpuConfig()
        self._L.ddx_gpu_default_config(ctypes.byref(cfg))
        cfg.device, cfg.lanes, cfg.tt_log2 = device, lanes, tt_log2
        cfg.jobs_per_deal, cfg.schedule, cfg.seed_mode = jobs_per_deal, schedule, seed_mode
        cfg.block, cfg.profile, cfg.team, cfg.graph = block, int(profile), team, int(graph)
        cfg.waves_per_simd = waves_per_simd
        cfg.lds_lines = lds_lines
        cfg.split_min_nodes = split_min_nodes
        self.cfg = cfg
        self.h = ctypes.c_void_p()
        self._ck(self._L.ddx_gpu_open(ctypes.byref(cfg), ctypes.byref(self.h)))

    def solve(self, hands, seeds=None, allow_stop=False):
        hands = np.ascontiguousarray(hands, np.uint64)
        n = len(hands)
        out = np.zeros((n, 20), np.uint8)
        s = None if seeds is None else np.ascontiguousarray(seeds, np.uint8)
        st = Stats()
        self._ck(self._L.ddx_gpu_solve(self.h, _ptr(hands), n, _ptr(s), _ptr(out), ctypes.byref(st)),
               ok=(0, E_STOPPED) if allow_stop else (0,))
        return out, st

    def solve_async(self, d_hands, n, d_out, d_seeds=0, stream=0):
        """Device-resident batch (raw device pointers as ints); then wait()."""
        self._ck(self._L.ddx_gpu_solve_async(self.h, ctypes.c_void_p(d_hands), n,
                                         ctypes.c_void_p(d_seeds) if d_seeds else None,
                                         ctypes.c_void_p(d_out), ctypes.c_void_p(stream) if stream else None))

    def wait(self, allow_stop=False):
        st = Stats()
        self._ck(self._L.ddx_gpu_wait(self.h, ctypes.byref(st)), ok=(0, E_STOPPED) if allow_stop else (0,))
        return st

    def request_stop(self):
        self._L.ddx_gpu_request_stop(self.h)

    @property
    def last_kernel_ms(self):
        return self._L.ddx_gpu_last_kernel_ms(self.h)

    @property
    def device_bytes(self):
        return self._L.ddx_gpu_device_bytes(self.h)

    def progress(self):
        v = ctypes.c_uint64()
        self._ck(self._L.ddx_gpu_progress(self.h, ctypes.byref(v)))
        return v.value

    def progress_nodes(self):
        """(jobs finished, nodes spent up to each search's last finished job) of a running solve."""
        v, w = ctypes.c_uint64(), ctypes.c_uint64()
        self._ck(self._L.ddx_gpu_progress_nodes(self.h, ctypes.byref(v), ctypes.byref(w)))
        return v.value, w.value

    def profile(self, reset=True):
        v = np.zeros(32, np.uint64)
        self._ck(self._L.ddx_gpu_profile(self.h, _ptr(v), int(reset)))
        return v

    def close(self):
        if self.h:
            self._L.ddx_gpu_close(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ns_tricks_of(tables):
    """NS tricks per (strain, declarer) from declarer tables (n, 20): E/W entries are 13 - x."""
    t = np.asarray(tables).reshape(-1, 5, 4).astype(np.int16)
    ns = t.copy()
    ns[:, :, 1::2] = 13 - t[:, :, 1::2]
    return ns.reshape(-1, 20)


E_FULL = 1


class Stream:
    """The persistent streaming solver (ddx_stream_*): push deals, poll finished tables."""

    def _ck(self, rc, ok=(0,)):
        return _check(rc, ok, self._L)

    def __init__(self, lanes=65536, tt_log2=10, jobs_per_deal=20, seed_mode=SEED_PLAIN,
                 ring_slots=4096, device=0, team=1, waves_per_simd=0, variant=""):
        self._L = lib(variant)
        cfg = GpuConfig()
        self._L.ddx_gpu_default_config(ctypes.byref(cfg))
        cfg.device, cfg.lanes, cfg.tt_log2, cfg.team = device, lanes, tt_log2, team
        cfg.waves_per_simd = waves_per_simd
        cfg.jobs_per_deal, cfg.seed_mode = jobs_per_deal, seed_mode
        self.h = ctypes.c_void_p()
        self._ck(self._L.ddx_stream_open(ctypes.byref(cfg), ring_slots, ctypes.byref(self.h)))
        self._ids = np.zeros(4096, np.int64)
        self._tab = np.zeros((4096, 20), np.uint8)

    def push(self, hands4, deal_id, seeds20=None):
        """True if queued, False if the ring is full."""
        h = np.ascontiguousarray(hands4, np.uint64)
        s = None if seeds20 is None else np.ascontiguousarray(seeds20, np.uint8)
        rc = self._ck(self._L.ddx_stream_push(self.h, _ptr(h), _ptr(s), deal_id), ok=(0, E_FULL))
        return rc == 0

    def poll(self):
        got = ctypes.c_int()
        self._ck(self._L.ddx_stream_poll(self.h, len(self._ids), _ptr(self._ids), _ptr(self._tab),
                                     ctypes.byref(got)))
        n = got.value
        return self._ids[:n].copy(), self._tab[:n].copy()

    def in_flight(self):
        return self._L.ddx_stream_in_flight(self.h)

    def finish(self):
        self._ck(self._L.ddx_stream_finish(self.h))

    def close(self):
        st = Stats()
        if self.h:
            h, self.h = self.h, ctypes.c_void_p()
            self._ck(self._L.ddx_stream_close(h, ctypes.byref(st)))
        return st

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TrickNet:
    """The trick-count MLP on the GPU (ddx_mlp_*)."""

    def __init__(self, seed=1, device=0, path=None):
        self.h = ctypes.c_void_p()
        _check(lib().ddx_mlp_create(device, seed, ctypes.byref(self.h)))
        if path:
            self.load(path)

    @staticmethod
    def param_count():
        return lib().ddx_mlp_param_count()

    def params(self):
        p = np.zeros(self.param_count(), np.float32)
        _check(lib().ddx_mlp_params(self.h, _ptr(p), 1))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib().ddx_mlp_params(self.h, _ptr(p), 0))

    def grads(self, hands, tables):
        hands = np.ascontiguousarray(hands, np.uint64)
        tables = np.ascontiguousarray(tables, np.uint8)
        g = np.zeros(self.param_count(), np.float32)
        la = np.zeros(2, np.float32)
        _check(lib().ddx_mlp_grads(self.h, _ptr(hands), _ptr(tables), len(hands), _ptr(g), _ptr(la)))
        return g, float(la[0]), float(la[1])

    def train(self, hands, tables, epochs=1, batch_deals=1024, lr=1e-3, shuffle_seed=0):
        hands = np.ascontiguousarray(hands, np.uint64)
        tables = np.ascontiguousarray(tables, np.uint8)
        la = np.zeros(2, np.float32)
        _check(lib().ddx_mlp_train(self.h, _ptr(hands), _ptr(tables), len(hands), epochs, batch_deals,
                                   lr, shuffle_seed, _ptr(la)))
        return float(la[0]), float(la[1])

    def predict(self, hands):
        """NS-trick seeds (n, 20) [strain][declarer]."""
        hands = np.ascontiguousarray(hands, np.uint64)
        out = np.zeros((len(hands), 20), np.uint8)
        _check(lib().ddx_mlp_predict_host(self.h, _ptr(hands), len(hands), _ptr(out)))
        return out

    def predict_device(self, d_hands, n, d_seeds, stream=0):
        _check(lib().ddx_mlp_predict(self.h, ctypes.c_void_p(d_hands), n, ctypes.c_void_p(d_seeds),
                                     ctypes.c_void_p(stream) if stream else None))

    def predict_device_noalloc(self, d_hands, n, d_seeds, stream=0):
        """ddx_mlp_predict_noalloc: at most 256 deals, activations in the library's fixed scratch."""
        _check(lib().ddx_mlp_predict_noalloc(self.h, ctypes.c_void_p(d_hands), n, ctypes.c_void_p(d_seeds),
                                             ctypes.c_void_p(stream) if stream else None))

    def save(self, path):
        _check(lib().ddx_mlp_save(self.h, path.encode()))

    def load(self, path):
        _check(lib().ddx_mlp_load(self.h, path.encode()))

    def close(self):
        if self.h:
            lib().ddx_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
