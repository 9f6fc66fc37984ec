"""Python bindings (ctypes) for librapidfuzz_amd.so, the gfx950 scoring engine.

Thin mirror of the C ABI in include/rapidfuzz_amd.h, used by the tests and
bench.py.  There is no CPU fallback: importing this module without the built
library raises, and every call runs on the GPU.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "librapidfuzz_amd.so")

(LEVENSHTEIN, INDEL, LCSSEQ, OSA, JARO, JARO_WINKLER, RATIO, HAMMING, PREFIX, POSTFIX, PARTIAL_RATIO, TOKEN_SORT_RATIO,
 PARTIAL_TOKEN_SORT_RATIO, QRATIO, TOKEN_SET_RATIO, TOKEN_RATIO, PARTIAL_TOKEN_SET_RATIO,
 PARTIAL_TOKEN_RATIO, DAMERAU_LEVENSHTEIN, WRATIO) = range(20)
DISTANCE, SIMILARITY, NORMALIZED_DISTANCE, NORMALIZED_SIMILARITY = range(4)
CACHED, MULTI = 0, 1
SCHEDULE_BUCKETED, SCHEDULE_PLAIN = 0, 1
OPTION_PAIRS_GRAPH, OPTION_PAIRS_CHAIN, OPTION_CDIST_BAND, OPTION_COUNT_WORD_STEPS = 1, 2, 3, 4
COUNTER_BAND_WORD_STEPS, COUNTER_BAND_PAIRS = 1, 2
SCORER_SIGNED_UNITS = 1
INT64_MAX = (1 << 63) - 1

_STATUS = {0: "RFA_OK", 1: "RFA_ERR_INVALID_ARGUMENT", 2: "RFA_ERR_UNSUPPORTED", 3: "RFA_ERR_HIP",
           4: "RFA_ERR_OUT_OF_MEMORY", 5: "RFA_ERR_NO_DEVICE"}

# every symbol include/rapidfuzz_amd.h declares
EXPORTED = ["rfa_abi_version", "rfa_last_error", "rfa_context_create", "rfa_context_destroy",
            "rfa_context_synchronize", "rfa_context_stream", "rfa_context_set_option", "rfa_context_counter", "rfa_strings_create", "rfa_strings_destroy",
            "rfa_strings_count", "rfa_scorer_create", "rfa_scorer_destroy", "rfa_scorer_set_option", "rfa_cdist", "rfa_cdist_cutoffs",
            "rfa_pairs", "rfa_pairs_noalloc", "rfa_pairs_autotune", "rfa_cdist_autotune", "rfa_topk", "rfa_editops",
            "rfa_partial_ratio_alignment", "rfa_server_create", "rfa_server_destroy", "rfa_server_submit", "rfa_server_wait", "rfa_server_replay",
            "rfa_server_get_stats"]


class RfaError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class InvalidArgument(RfaError, ValueError):
    pass


class Weights(ctypes.Structure):
    _fields_ = [("insert_cost", ctypes.c_uint64), ("delete_cost", ctypes.c_uint64),
                ("replace_cost", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("metric", ctypes.c_int), ("kind", ctypes.c_int), ("score_cutoff", ctypes.c_double),
                ("k", ctypes.c_uint32), ("max_batch", ctypes.c_uint32), ("slots", ctypes.c_uint32),
                ("candidate_cap", ctypes.c_uint32), ("prefix_weight", ctypes.c_double), ("weights", Weights)]


class ServerStats(ctypes.Structure):
    _fields_ = [("requests", ctypes.c_uint64), ("batches", ctypes.c_uint64), ("max_batch_seen", ctypes.c_uint64),
                ("dense_fallbacks", ctypes.c_uint64)]


_vp = ctypes.c_void_p
_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run `make -C rapidfuzz-cpp_amd` (or __graft_entry__.build())")
    L = ctypes.CDLL(path)
    L.rfa_abi_version.restype = ctypes.c_int
    L.rfa_last_error.restype = ctypes.c_char_p
    L.rfa_context_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.rfa_context_destroy.argtypes = [_vp]
    L.rfa_context_synchronize.argtypes = [_vp]
    L.rfa_context_stream.argtypes = [_vp]
    L.rfa_context_stream.restype = _vp
    L.rfa_context_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_int64]
    L.rfa_context_counter.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
    L.rfa_strings_create.argtypes = [_vp, _vp, ctypes.c_int, _vp, ctypes.c_uint64, ctypes.POINTER(_vp)]
    L.rfa_strings_destroy.argtypes = [_vp]
    L.rfa_strings_count.argtypes = [_vp]
    L.rfa_strings_count.restype = ctypes.c_uint64
    L.rfa_scorer_create.argtypes = [_vp, ctypes.c_int, _vp, ctypes.POINTER(Weights), ctypes.c_double,
                                    ctypes.POINTER(_vp)]
    L.rfa_scorer_destroy.argtypes = [_vp]
    L.rfa_scorer_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_int64]
    L.rfa_cdist.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, _vp,
                            ctypes.c_int]
    L.rfa_cdist_cutoffs.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, _vp, ctypes.c_int]
    L.rfa_pairs.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, ctypes.c_int,
                            _vp, ctypes.c_int]
    L.rfa_pairs_noalloc.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_double,
                                    _vp]
    L.rfa_pairs_autotune.argtypes = [_vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_double, _vp]
    L.rfa_cdist_autotune.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_uint64, _vp, _vp]
    L.rfa_topk.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_double, ctypes.c_uint32, _vp, _vp]
    L.rfa_editops.argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp, _vp, ctypes.c_uint64]
    L.rfa_partial_ratio_alignment.argtypes = [_vp, _vp, _vp, ctypes.c_double, _vp]
    L.rfa_server_create.argtypes = [_vp, _vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(_vp)]
    L.rfa_server_destroy.argtypes = [_vp]
    L.rfa_server_submit.argtypes = [_vp, _vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
    L.rfa_server_wait.argtypes = [_vp, ctypes.c_uint64, _vp, _vp]
    L.rfa_server_replay.argtypes = [_vp, _vp, _vp, ctypes.c_uint64, _vp, _vp, _vp, _vp]
    L.rfa_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
    for name in ("rfa_context_create", "rfa_context_destroy", "rfa_context_synchronize", "rfa_strings_create",
                 "rfa_strings_destroy", "rfa_scorer_create", "rfa_scorer_destroy", "rfa_cdist", "rfa_cdist_cutoffs",
                 "rfa_pairs", "rfa_pairs_autotune", "rfa_context_set_option", "rfa_scorer_set_option",
                 "rfa_topk", "rfa_server_create", "rfa_server_destroy", "rfa_server_submit", "rfa_server_wait",
                 "rfa_server_replay", "rfa_server_get_stats", "rfa_editops", "rfa_partial_ratio_alignment"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _check(status):
    if status != 0:
        msg = _lib.rfa_last_error().decode()
        if status == 1:
            raise InvalidArgument(status, msg)
        raise RfaError(status, msg)


# metrics whose top-k ranks the plain similarity (0-100 or 0-1 doubles)
_SIM_METRICS = (JARO, JARO_WINKLER, RATIO, PARTIAL_RATIO, TOKEN_SORT_RATIO, PARTIAL_TOKEN_SORT_RATIO, QRATIO,
                TOKEN_SET_RATIO, TOKEN_RATIO, PARTIAL_TOKEN_SET_RATIO, PARTIAL_TOKEN_RATIO, WRATIO)


def is_int_result(metric, kind):
    return metric in (LEVENSHTEIN, INDEL, LCSSEQ, OSA, HAMMING, PREFIX, POSTFIX, DAMERAU_LEVENSHTEIN) and \
        kind in (DISTANCE, SIMILARITY)


def default_cutoffs(metric, kind):
    """The reference's default cutoffs for a method (details/distance.hpp:211-301)."""
    if is_int_result(metric, kind):
        return (INT64_MAX if kind == DISTANCE else 0), 0.0
    if kind == NORMALIZED_DISTANCE or (kind == DISTANCE and metric in (JARO, JARO_WINKLER)):
        return INT64_MAX, 1.0
    return INT64_MAX, 0.0


def _as_units(s, width):
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[width]
    if isinstance(s, np.ndarray):
        return s.astype(dt, copy=False)
    if isinstance(s, (bytes, bytearray)):
        return np.frombuffer(bytes(s), dtype=np.uint8).astype(dt)
    if isinstance(s, str):
        return np.array([ord(c) for c in s], dtype=dt)
    return np.asarray(list(s), dtype=dt)


def pack(strings, width=None):
    """(chars, offsets, width) for a list of strings; width inferred from the code points."""
    arrs = []
    mx = 0
    for s in strings:
        if isinstance(s, str):
            a = np.array([ord(c) for c in s], dtype=np.uint32)
        elif isinstance(s, (bytes, bytearray)):
            a = np.frombuffer(bytes(s), dtype=np.uint8)
        else:
            a = np.asarray(s)
        arrs.append(a)
        if a.size:
            mx = max(mx, int(a.max()))
    if width is None:
        width = 1 if mx < 256 else (2 if mx < 65536 else 4)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32}[width]
    off = np.zeros(len(arrs) + 1, dtype=np.uint64)
    if arrs:
        np.cumsum([a.size for a in arrs], out=off[1:])
    chars = np.concatenate([a.astype(dt) for a in arrs]) if arrs and off[-1] else np.zeros(1, dt)
    return np.ascontiguousarray(chars), off, width


class Context:
    def __init__(self, device: int = 0):
        L = load_library()
        h = _vp()
        _check(L.rfa_context_create(device, ctypes.byref(h)))
        self.h = h

    @property
    def stream(self):
        return _lib.rfa_context_stream(self.h)

    def synchronize(self):
        _check(_lib.rfa_context_synchronize(self.h))

    def set_option(self, option: int, value: int):
        """rfa_context_set_option, e.g. OPTION_PAIRS_GRAPH: replay repeated pairs calls from a HIP graph."""
        _check(_lib.rfa_context_set_option(self.h, option, value))

    def counter(self, which: int) -> int:
        """rfa_context_counter (COUNTER_BAND_WORD_STEPS / COUNTER_BAND_PAIRS; counting is on
        while OPTION_COUNT_WORD_STEPS is 1)."""
        v = ctypes.c_uint64(0)
        _check(_lib.rfa_context_counter(self.h, which, ctypes.byref(v)))
        return int(v.value)

    def close(self):
        """Destroys the context; strings and scorers made from it become unusable
        (their device memory is released with the context's process)."""
        if self.h:
            _check(_lib.rfa_context_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def strings(self, strings=None, *, chars=None, offsets=None, width=None):
        return Strings(self, strings, chars=chars, offsets=offsets, width=width)


class Strings:
    """A device-resident packed string set."""

    def __init__(self, ctx: Context, strings=None, *, chars=None, offsets=None, width=None):
        self.ctx = ctx
        if strings is not None:
            chars, offsets, width = pack(strings, width)
        assert chars is not None and offsets is not None
        if width is None:
            width = chars.dtype.itemsize
        chars = np.ascontiguousarray(chars)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.width = width
        self.count = len(offsets) - 1
        self.lengths = np.diff(offsets)
        h = _vp()
        _check(_lib.rfa_strings_create(ctx.h, chars.ctypes.data, width, offsets.ctypes.data, self.count,
                                       ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h and self.ctx.h:
            _check(_lib.rfa_strings_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scorer:
    """Cached<metric>(q) for every q of `queries` (or Multi<metric> over them)."""

    def __init__(self, ctx: Context, metric: int, queries: Strings, weights=(1, 1, 1), prefix_weight=0.1):
        self.ctx = ctx
        self.metric = metric
        self.nq = queries.count
        w = Weights(*weights)
        h = _vp()
        _check(_lib.rfa_scorer_create(ctx.h, metric, queries.h, ctypes.byref(w), prefix_weight, ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h and self.ctx.h:
            _check(_lib.rfa_scorer_destroy(self.h))
        self.h = None

    def set_option(self, option, value):
        """rfa_scorer_set_option (SCORER_SIGNED_UNITS: 8-bit units order as signed `char`)."""
        _check(_lib.rfa_scorer_set_option(self.h, option, ctypes.c_int64(value)))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cdist(self, choices: Strings, kind=DISTANCE, cutoff_u=None, cutoff_d=None, semantics=CACHED,
              out_device_ptr=None, cutoffs_u=None):
        """All queries x all choices.  cutoffs_u: one integer cutoff per query (integer kinds)."""
        if cutoffs_u is not None:
            cut = np.ascontiguousarray(cutoffs_u, dtype=np.uint64)
            if out_device_ptr is not None:
                _check(_lib.rfa_cdist_cutoffs(self.h, choices.h, kind, semantics, cut.ctypes.data, out_device_ptr, 1))
                return None
            out = np.zeros((self.nq, choices.count), dtype=np.uint64)
            _check(_lib.rfa_cdist_cutoffs(self.h, choices.h, kind, semantics, cut.ctypes.data, out.ctypes.data, 0))
            return out
        du, dd = default_cutoffs(self.metric, kind)
        cu = du if cutoff_u is None else cutoff_u
        cd = dd if cutoff_d is None else cutoff_d
        if out_device_ptr is not None:
            _check(_lib.rfa_cdist(self.h, choices.h, kind, semantics, cu, cd, out_device_ptr, 1))
            return None
        dt = np.uint64 if is_int_result(self.metric, kind) else np.float64
        out = np.zeros((self.nq, choices.count), dtype=dt)
        _check(_lib.rfa_cdist(self.h, choices.h, kind, semantics, cu, cd, out.ctypes.data, 0))
        return out

    def cdist_autotune(self, choices: Strings, kind=DISTANCE, cutoff_u=None, cutoffs_u=None):
        """rfa_cdist_autotune: the band kernel's words per lane and workgroup size for each
        window-size bucket, measured on this cdist; returns a (6, 2) array (0: bucket unused)."""
        chosen = np.zeros(12, dtype=np.uint32)
        cu = default_cutoffs(self.metric, kind)[0] if cutoff_u is None else cutoff_u
        cut = None if cutoffs_u is None else np.ascontiguousarray(cutoffs_u, dtype=np.uint64)
        _check(_lib.rfa_cdist_autotune(self.h, choices.h, kind, cu, None if cut is None else cut.ctypes.data,
                                       chosen.ctypes.data))
        return chosen.reshape(6, 2)

    def topk(self, choices: Strings, k: int, kind=None, cutoff_d=0.0):
        """(index, score), each [nq, k]: the extract() results >= cutoff_d ordered by
        (score desc, index asc); missing entries are index -1, score 0."""
        if kind is None:
            kind = SIMILARITY if self.metric in _SIM_METRICS else NORMALIZED_SIMILARITY
        idx = np.zeros((self.nq, k), dtype=np.int64)
        sc = np.zeros((self.nq, k), dtype=np.float64)
        _check(_lib.rfa_topk(self.h, choices.h, kind, cutoff_d, k, idx.ctypes.data, sc.ctypes.data))
        return idx, sc


class Server:
    """Continuous-batching top-k request server over a resident choice set."""

    def __init__(self, ctx: Context, choices: Strings, metric=RATIO, kind=None, score_cutoff=0.0, k=5,
                 max_batch=64, slots=4, candidate_cap=65536, prefix_weight=0.1, weights=(1, 1, 1)):
        if kind is None:
            kind = SIMILARITY if metric in _SIM_METRICS else NORMALIZED_SIMILARITY
        self.ctx = ctx
        self.k = k
        self.width = choices.width
        cfg = ServerConfig(metric, kind, score_cutoff, k, max_batch, slots, candidate_cap, prefix_weight,
                           Weights(*weights))
        h = _vp()
        _check(_lib.rfa_server_create(ctx.h, choices.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def close(self):
        if self.h and self.ctx.h:
            _check(_lib.rfa_server_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, query) -> int:
        q = np.ascontiguousarray(_as_units(query, self.width))
        t = ctypes.c_uint64()
        _check(_lib.rfa_server_submit(self.h, q.ctypes.data, q.size, ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        idx = np.zeros(self.k, dtype=np.int64)
        sc = np.zeros(self.k, dtype=np.float64)
        _check(_lib.rfa_server_wait(self.h, ticket, idx.ctypes.data, sc.ctypes.data))
        return idx, sc

    def replay(self, queries, arrival_s):
        """Submits queries[i] at arrival_s[i] seconds; returns (latency_s, index, score)."""
        chars, off, _ = pack(queries, self.width)
        arr = np.ascontiguousarray(arrival_s, dtype=np.float64)
        n = len(off) - 1
        lat = np.zeros(n, dtype=np.float64)
        idx = np.zeros((n, self.k), dtype=np.int64)
        sc = np.zeros((n, self.k), dtype=np.float64)
        _check(_lib.rfa_server_replay(self.h, chars.ctypes.data, off.ctypes.data, n, arr.ctypes.data,
                                      lat.ctypes.data, idx.ctypes.data, sc.ctypes.data))
        return lat, idx, sc

    def stats(self):
        st = ServerStats()
        _check(_lib.rfa_server_get_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}


EDIT_NONE, EDIT_REPLACE, EDIT_INSERT, EDIT_DELETE = range(4)
EDITOP_DTYPE = np.dtype([("type", np.int32), ("reserved", np.int32), ("src_pos", np.uint64), ("dest_pos", np.uint64)])


def editops(ctx: Context, a: Strings, b: Strings, metric=LEVENSHTEIN):
    """Edit scripts of the pairs (a[i], b[i]): (offsets [count + 1], ops structured array);
    pair i's operations are ops[offsets[i]:offsets[i + 1]]."""
    off = np.zeros(a.count + 1, dtype=np.uint64)
    # a script never holds more operations than the two strings have units
    cap = max(1, int(a.lengths.sum()) + int(b.lengths.sum()))
    ops = np.empty(cap, dtype=EDITOP_DTYPE)
    _check(_lib.rfa_editops(ctx.h, a.h, b.h, metric, off.ctypes.data, ops.ctypes.data, cap))
    return off, ops[:int(off[-1])]


ALIGNMENT_DTYPE = np.dtype([("score", np.float64), ("src_start", np.uint64), ("src_end", np.uint64),
                            ("dest_start", np.uint64), ("dest_end", np.uint64)])


def partial_ratio_alignment(ctx: Context, a: Strings, b: Strings, score_cutoff=0.0):
    """fuzz::partial_ratio_alignment(a[i], b[i], score_cutoff) of every pair, as a structured
    array (score, src_start, src_end, dest_start, dest_end)."""
    out = np.zeros(a.count, dtype=ALIGNMENT_DTYPE)
    _check(_lib.rfa_partial_ratio_alignment(ctx.h, a.h, b.h, float(score_cutoff), out.ctypes.data))
    return out


def pairs(ctx: Context, a: Strings, b: Strings, metrics, kind=DISTANCE, cutoff_u=INT64_MAX, cutoff_d=1.0,
          schedule=SCHEDULE_BUCKETED, out_device_ptr=None):
    m = np.ascontiguousarray(np.asarray(metrics, dtype=np.int32))
    if out_device_ptr is not None:
        _check(_lib.rfa_pairs(ctx.h, a.h, b.h, m.ctypes.data, kind, cutoff_u, cutoff_d, schedule, out_device_ptr, 1))
        return None
    dt = np.uint64 if kind in (DISTANCE, SIMILARITY) else np.float64
    out = np.zeros(a.count, dtype=dt)
    _check(_lib.rfa_pairs(ctx.h, a.h, b.h, m.ctypes.data, kind, cutoff_u, cutoff_d, schedule, out.ctypes.data, 0))
    return out


def pairs_noalloc(ctx: Context, a: Strings, b: Strings, metrics_ptr, has_osa, out_ptr, kind=DISTANCE,
                  cutoff_u=INT64_MAX, cutoff_d=1.0):
    """rfa_pairs_noalloc: metrics_ptr (one uint8 code per pair) and out_ptr (8 bytes per pair) are
    device addresses; the call only enqueues work on the context's stream."""
    _check(_lib.rfa_pairs_noalloc(ctx.h, a.h, b.h, metrics_ptr, int(has_osa), kind, cutoff_u, cutoff_d, out_ptr))


def pairs_autotune(ctx: Context, a: Strings, b: Strings, metrics, kind=DISTANCE, cutoff_u=INT64_MAX, cutoff_d=1.0):
    """rfa_pairs_autotune: times the bucketed schedule for each candidate long-pair threshold on this
    batch, keeps the fastest in the context and returns it."""
    m = np.ascontiguousarray(np.asarray(metrics, dtype=np.int32))
    chosen = ctypes.c_uint32(0)
    _check(_lib.rfa_pairs_autotune(ctx.h, a.h, b.h, m.ctypes.data, kind, cutoff_u, cutoff_d, ctypes.byref(chosen)))
    return int(chosen.value)
