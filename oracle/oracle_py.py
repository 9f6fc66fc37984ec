"""ctypes bindings for the checker libraries under oracle/ (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  It loads

* ``librf_oracle.so`` -- the C restatement of the reference semantics, and
* ``_ref/librfref.so`` -- the reference's own headers compiled in place
  (optional: present where it was built from the reference sources).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "librf_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "librfref.so")

(LEVENSHTEIN, INDEL, LCSSEQ, OSA, JARO, JARO_WINKLER, RATIO, HAMMING, PREFIX, POSTFIX, PARTIAL_RATIO, TOKEN_SORT_RATIO,
 PARTIAL_TOKEN_SORT_RATIO, QRATIO, TOKEN_SET_RATIO, TOKEN_RATIO, PARTIAL_TOKEN_SET_RATIO,
 PARTIAL_TOKEN_RATIO, DAMERAU_LEVENSHTEIN) = range(19)
DISTANCE, SIMILARITY, NORM_DISTANCE, NORM_SIMILARITY = range(4)
INT64_MAX = (1 << 63) - 1

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)


def to_codes(s) -> np.ndarray:
    """A string-like value as uint32 code units."""
    if isinstance(s, np.ndarray):
        return s.astype(np.uint32, copy=False)
    if isinstance(s, (bytes, bytearray)):
        return np.frombuffer(bytes(s), dtype=np.uint8).astype(np.uint32)
    if isinstance(s, str):
        return np.array([ord(c) for c in s], dtype=np.uint32)
    return np.asarray(list(s), dtype=np.uint32)


def pack(strings, dtype=np.uint32):
    """Pack strings into (chars, offsets[n+1])."""
    arrs = [to_codes(s) for s in strings]
    lens = np.array([len(a) for a in arrs], dtype=np.uint64)
    off = np.zeros(len(arrs) + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    chars = np.concatenate(arrs).astype(dtype) if arrs and off[-1] else np.zeros(1, dtype=dtype)
    return np.ascontiguousarray(chars), off


def _p(a, t):
    return a.ctypes.data_as(t)


def is_int_result(metric: int, kind: int) -> bool:
    return metric in (LEVENSHTEIN, INDEL, LCSSEQ, OSA, HAMMING, PREFIX, POSTFIX, DAMERAU_LEVENSHTEIN) and \
        kind in (DISTANCE, SIMILARITY)


def _alignment(fn, s1, s2, cutoff):
    a = np.ascontiguousarray(np.asarray(s1, dtype=np.uint32))
    b = np.ascontiguousarray(np.asarray(s2, dtype=np.uint32))
    sc = np.zeros(1, np.float64)
    al = np.zeros(4, np.uint64)
    r = fn(a.ctypes.data_as(_u32p), len(a), b.ctypes.data_as(_u32p), len(b), float(cutoff),
           sc.ctypes.data_as(_f64p), al.ctypes.data_as(_u64p))
    if r is not None and r != 0:
        raise RuntimeError("partial_ratio_alignment failed")
    return (float(sc[0]),) + tuple(int(x) for x in al)


class Oracle:
    """The C restatement (rf_oracle.c)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run `make -C oracle`)")
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.orc_cdist.restype = ctypes.c_int
        L.orc_cdist.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, _u64p, ctypes.c_size_t,
                                _u32p, _u64p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64,
                                ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                                _u64p, _f64p]
        L.orc_pairs.restype = ctypes.c_int
        L.orc_pairs.argtypes = [_i32p, ctypes.c_int, _u32p, _u64p, _u32p, _u64p, ctypes.c_size_t,
                                ctypes.c_uint64, ctypes.c_double, _u64p, _f64p]
        L.orc_set_units.restype = None
        L.orc_set_units.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_editops.restype = ctypes.c_size_t
        L.orc_editops.argtypes = [ctypes.c_int, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, _i32p, _u64p, _u64p]
        L.orc_set_lev_block.argtypes = [ctypes.c_int]
        L.orc_sorted_join.restype = ctypes.c_size_t
        L.orc_sorted_join.argtypes = [_u32p, ctypes.c_size_t, _u32p]
        L.orc_jaro_raw.restype = None
        L.orc_jaro_raw.argtypes = [_u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, _u64p, _u64p]
        for name in ("orc_levenshtein_raw",):
            getattr(L, name).restype = ctypes.c_uint64
            getattr(L, name).argtypes = [_u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, ctypes.c_uint64,
                                         ctypes.c_uint64, ctypes.c_uint64]
        for name in ("orc_lcs_raw", "orc_lcs_bits", "orc_osa_raw"):
            getattr(L, name).restype = ctypes.c_uint64
            getattr(L, name).argtypes = [_u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t]
        L.orc_partial_ratio_alignment.restype = None
        L.orc_partial_ratio_alignment.argtypes = [_u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, ctypes.c_double,
                                                  _f64p, _u64p]

    def partial_ratio_alignment(self, s1, s2, cutoff=0.0):
        """(score, src_start, src_end, dest_start, dest_end) of fuzz::partial_ratio_alignment."""
        return _alignment(self.lib.orc_partial_ratio_alignment, s1, s2, cutoff)

    def lcs_raw(self, s1, s2):
        """Length of the longest common subsequence (plain DP)."""
        a = np.ascontiguousarray(np.asarray(s1, dtype=np.uint32))
        b = np.ascontiguousarray(np.asarray(s2, dtype=np.uint32))
        return int(self.lib.orc_lcs_raw(a.ctypes.data_as(_u32p), len(a), b.ctypes.data_as(_u32p), len(b)))

    def lcs_bits(self, s1, s2):
        """Length of the longest common subsequence (bit-parallel recurrence)."""
        a = np.ascontiguousarray(np.asarray(s1, dtype=np.uint32))
        b = np.ascontiguousarray(np.asarray(s2, dtype=np.uint32))
        return int(self.lib.orc_lcs_bits(a.ctypes.data_as(_u32p), len(a), b.ctypes.data_as(_u32p), len(b)))

    def cdist(self, metric, kind, queries, choices, weights=(1, 1, 1), prefix_weight=0.1,
              cutoff_u=None, cutoff_d=None, multi=False, char_width=4, signed8=False):
        """char_width / signed8: the code units the token scorers split and order (the
        reference's whitespace set of 8-bit units, and `char`'s signed order)."""
        self.lib.orc_set_units(char_width, int(signed8))
        if cutoff_u is None:
            cutoff_u = default_int_cutoff(kind)
        qc, qo = pack(queries)
        cc, co = pack(choices)
        nq, nc = len(qo) - 1, len(co) - 1
        if cutoff_d is None:
            cutoff_d = default_fp_cutoff(metric, kind)
        ou = np.zeros(nq * nc, dtype=np.uint64)
        od = np.zeros(nq * nc, dtype=np.float64)
        r = self.lib.orc_cdist(metric, kind, int(multi), _p(qc, _u32p), _p(qo, _u64p), nq, _p(cc, _u32p),
                               _p(co, _u64p), nc, weights[0], weights[1], weights[2], prefix_weight,
                               cutoff_u, cutoff_d, _p(ou, _u64p), _p(od, _f64p))
        if r < 0:
            raise ValueError(f"unsupported metric/kind {metric}/{kind}")
        return (ou if r == 0 else od).reshape(nq, nc)

    def cdist_mt(self, metric, kind, queries, choices, threads=None, **kw):
        """cdist over row blocks of the queries on a thread pool (ctypes releases the GIL).
        Per-query cutoffs: cutoffs_u=[...] scores row i with cutoff_u=cutoffs_u[i]."""
        from concurrent.futures import ThreadPoolExecutor

        cut = kw.pop("cutoffs_u", None)
        threads = threads or min(16, os.cpu_count() or 1)
        if cut is not None:
            rows = [(i, i + 1) for i in range(len(queries))]
        else:
            step = max(1, (len(queries) + threads - 1) // threads)
            rows = [(i, min(len(queries), i + step)) for i in range(0, len(queries), step)]

        def part(r):
            extra = {} if cut is None else {"cutoff_u": int(cut[r[0]])}
            return self.cdist(metric, kind, queries[r[0]:r[1]], choices, **kw, **extra)

        with ThreadPoolExecutor(threads) as ex:
            parts = list(ex.map(part, rows))
        return np.concatenate(parts, axis=0) if parts else np.zeros((0, len(choices)))

    def score(self, metric, kind, s1, s2, **kw):
        return self.cdist(metric, kind, [s1], [s2], **kw)[0, 0]

    def editops(self, metric, s1, s2):
        """(type, src_pos, dest_pos) arrays of levenshtein_editops (metric 0) or
        lcs_seq_editops / indel_editops (metric 1 or 2)."""
        a, b = to_codes(s1), to_codes(s2)
        n = len(a) + len(b)
        t = np.zeros(max(1, n), np.int32)
        sp = np.zeros(max(1, n), np.uint64)
        dp = np.zeros(max(1, n), np.uint64)
        k = self.lib.orc_editops(metric, _p(a, _u32p), len(a), _p(b, _u32p), len(b), _p(t, _i32p), _p(sp, _u64p),
                                 _p(dp, _u64p))
        if k == ctypes.c_size_t(-1).value:
            raise ValueError("the block band vanished (distance above the bound)")
        return t[:k], sp[:k], dp[:k]

    def set_lev_block(self, on):
        """Levenshtein editops past 64 stripped units: the reference's block band (True, the
        default) or the full-matrix walk (False)."""
        self.lib.orc_set_lev_block(int(bool(on)))

    def sorted_join(self, s, char_width=4, signed8=False):
        """detail::sorted_split(s).join() as uint32 code units."""
        self.lib.orc_set_units(char_width, int(signed8))
        a = np.ascontiguousarray(to_codes(s), dtype=np.uint32)
        out = np.zeros(max(1, len(a)), dtype=np.uint32)
        n = self.lib.orc_sorted_join(_p(a, _u32p), len(a), _p(out, _u32p))
        return out[:n]

    def pairs(self, metrics, a, b, kind=DISTANCE, cutoff_u=INT64_MAX, cutoff_d=1.0):
        ac, ao = pack(a)
        bc, bo = pack(b)
        n = len(ao) - 1
        m = np.ascontiguousarray(np.asarray(metrics, dtype=np.int32))
        ou = np.zeros(n, dtype=np.uint64)
        od = np.zeros(n, dtype=np.float64)
        r = self.lib.orc_pairs(_p(m, _i32p), kind, _p(ac, _u32p), _p(ao, _u64p), _p(bc, _u32p),
                               _p(bo, _u64p), n, cutoff_u, cutoff_d, _p(ou, _u64p), _p(od, _f64p))
        if r < 0:
            raise ValueError("unsupported")
        return ou if r == 0 else od

    def jaro_raw(self, p, t):
        pa, ta = to_codes(p), to_codes(t)
        pa = np.ascontiguousarray(pa if len(pa) else np.zeros(1, np.uint32))
        ta = np.ascontiguousarray(ta if len(ta) else np.zeros(1, np.uint32))
        cc = np.zeros(1, np.uint64)
        tr = np.zeros(1, np.uint64)
        self.lib.orc_jaro_raw(_p(pa, _u32p), len(to_codes(p)), _p(ta, _u32p), len(to_codes(t)),
                              _p(cc, _u64p), _p(tr, _u64p))
        return int(cc[0]), int(tr[0])


def default_int_cutoff(kind):
    """WorstDistance (INT64_MAX) for distance(), WorstSimilarity (0) for similarity()."""
    return INT64_MAX if kind == DISTANCE else 0


def default_fp_cutoff(metric, kind):
    """The reference's default double cutoff for a method (details/distance.hpp:211-233)."""
    if kind in (NORM_DISTANCE,):
        return 1.0
    if kind == DISTANCE and metric in (JARO, JARO_WINKLER):
        return 1.0
    return 0.0


class Reference:
    """The reference's own headers compiled in place (oracle/_ref/librfref.so)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run `make -C oracle ref`)")
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_cdist_cached.restype = ctypes.c_int
        L.ref_cdist_cached.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, _u64p,
                                       ctypes.c_size_t, _u32p, _u64p, ctypes.c_size_t, ctypes.c_uint64,
                                       ctypes.c_uint64, ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64,
                                       ctypes.c_double, _u64p, _f64p]
        L.ref_cdist_multi_w.restype = ctypes.c_int
        L.ref_cdist_multi_w.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, _u64p, ctypes.c_size_t,
                                        _u32p, _u64p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64,
                                        ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                                        _u64p, _f64p]
        L.ref_cdist_multi.restype = ctypes.c_int
        L.ref_cdist_multi.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _u32p, _u64p, ctypes.c_size_t,
                                      _u32p, _u64p, ctypes.c_size_t, ctypes.c_double, ctypes.c_uint64,
                                      ctypes.c_double, _u64p, _f64p]
        for name in ("ref_cdist_cached_u8_mt", "ref_cdist_multi_u8_mt"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, ctypes.c_int, _u8p, _u64p, ctypes.c_size_t, _u8p, _u64p,
                          ctypes.c_size_t, ctypes.c_double, ctypes.c_uint64, ctypes.c_double, _u64p, _f64p,
                          ctypes.c_int]
        L.ref_cdist_cached_u8_mt_w.restype = ctypes.c_int
        L.ref_cdist_cached_u8_mt_w.argtypes = [ctypes.c_int, ctypes.c_int, _u8p, _u64p, ctypes.c_size_t, _u8p,
                                               _u64p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64,
                                               ctypes.c_uint64, ctypes.c_double, ctypes.c_uint64, ctypes.c_double,
                                               _u64p, _f64p, ctypes.c_int]
        L.ref_editops.restype = ctypes.c_int64
        L.ref_editops.argtypes = [ctypes.c_int, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, _i32p, _u64p,
                                  _u64p]
        L.ref_opcodes.restype = ctypes.c_int64
        L.ref_opcodes.argtypes = [ctypes.c_int, _u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, _u64p]
        L.ref_partial_ratio_alignment.restype = ctypes.c_int
        L.ref_partial_ratio_alignment.argtypes = [_u32p, ctypes.c_size_t, _u32p, ctypes.c_size_t, ctypes.c_double,
                                                  _f64p, _u64p]
        L.ref_editops_algebra.restype = ctypes.c_int64
        L.ref_editops_algebra.argtypes = [ctypes.c_int, _i64p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64p, ctypes.c_size_t, _i64p,
                                          ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        L.ref_editops_u8_mt.restype = ctypes.c_int
        L.ref_editops_u8_mt.argtypes = [ctypes.c_int, _u8p, _u64p, _u8p, _u64p, ctypes.c_size_t, _u64p, ctypes.c_int]
        L.ref_partial_ratio_alignment_u8_mt.restype = ctypes.c_int
        L.ref_partial_ratio_alignment_u8_mt.argtypes = [_u8p, _u64p, _u8p, _u64p, ctypes.c_size_t, ctypes.c_double,
                                                        _f64p, _u64p, ctypes.c_int]
        L.ref_pairs_u32_mt.restype = ctypes.c_int
        L.ref_pairs_u32_mt.argtypes = [_i32p, _u32p, _u64p, _u32p, _u64p, ctypes.c_size_t, _u64p,
                                       ctypes.c_int]
        L.ref_lev_cutoff_u8_mt.restype = ctypes.c_int
        L.ref_lev_cutoff_u8_mt.argtypes = [_u8p, _u64p, ctypes.c_size_t, _u8p, _u64p, ctypes.c_size_t, _u64p,
                                           _u64p, ctypes.c_int]
        L.ref_extract_ratio_u8_mt.restype = ctypes.c_int
        L.ref_extract_ratio_u8_mt.argtypes = [_u8p, _u64p, ctypes.c_size_t, _u8p, _u64p, ctypes.c_size_t,
                                              ctypes.c_double, ctypes.c_size_t, _i64p, _f64p, _f64p,
                                              ctypes.c_int]

    def _err(self):
        return self.lib.ref_last_error().decode()

    def cdist(self, metric, kind, queries, choices, weights=(1, 1, 1), prefix_weight=0.1,
              cutoff_u=None, cutoff_d=None, char_width=4, multi_len=None):
        if cutoff_u is None:
            cutoff_u = default_int_cutoff(kind)
        qc, qo = pack(queries)
        cc, co = pack(choices)
        nq, nc = len(qo) - 1, len(co) - 1
        if cutoff_d is None:
            cutoff_d = default_fp_cutoff(metric, kind)
        ou = np.zeros(nq * nc, dtype=np.uint64)
        od = np.zeros(nq * nc, dtype=np.float64)
        if multi_len:
            r = self.lib.ref_cdist_multi_w(metric, kind, multi_len, _p(qc, _u32p), _p(qo, _u64p), nq,
                                           _p(cc, _u32p), _p(co, _u64p), nc, weights[0], weights[1], weights[2],
                                           prefix_weight, cutoff_u, cutoff_d, _p(ou, _u64p), _p(od, _f64p))
        else:
            r = self.lib.ref_cdist_cached(metric, kind, char_width, _p(qc, _u32p), _p(qo, _u64p), nq,
                                          _p(cc, _u32p), _p(co, _u64p), nc, weights[0], weights[1],
                                          weights[2], prefix_weight, cutoff_u, cutoff_d, _p(ou, _u64p),
                                          _p(od, _f64p))
        if r < 0:
            raise ValueError(f"reference error {r}: {self._err()}")
        return (ou if r == 0 else od).reshape(nq, nc)

    def cdist_u8_mt(self, metric, kind, qchars, qoff, cchars, coff, cutoff_u=None, cutoff_d=None,
                    prefix_weight=0.1, threads=None, multi=False, weights=None):
        """Threaded u8 cdist (the CPU baseline); inputs already packed as uint8.  weights:
        Levenshtein (insert, delete, replace) of the Cached* scorers."""
        if cutoff_u is None:
            cutoff_u = default_int_cutoff(kind)
        nq, nc = len(qoff) - 1, len(coff) - 1
        if cutoff_d is None:
            cutoff_d = default_fp_cutoff(metric, kind)
        threads = threads or os.cpu_count() or 1
        ou = np.zeros(nq * nc, dtype=np.uint64)
        od = np.zeros(nq * nc, dtype=np.float64)
        if weights is not None and not multi:
            r = self.lib.ref_cdist_cached_u8_mt_w(metric, kind, _p(qchars, _u8p), _p(qoff, _u64p), nq,
                                                  _p(cchars, _u8p), _p(coff, _u64p), nc, *weights, prefix_weight,
                                                  cutoff_u, cutoff_d, _p(ou, _u64p), _p(od, _f64p), threads)
        else:
            f = self.lib.ref_cdist_multi_u8_mt if multi else self.lib.ref_cdist_cached_u8_mt
            r = f(metric, kind, _p(qchars, _u8p), _p(qoff, _u64p), nq, _p(cchars, _u8p), _p(coff, _u64p), nc,
                  prefix_weight, cutoff_u, cutoff_d, _p(ou, _u64p), _p(od, _f64p), threads)
        if r < 0:
            raise ValueError(f"reference error {r}: {self._err()}")
        return (ou if is_int_result(metric, kind) else od).reshape(nq, nc)

    def editops(self, metric, s1, s2):
        a = np.ascontiguousarray(to_codes(s1), dtype=np.uint32)
        b = np.ascontiguousarray(to_codes(s2), dtype=np.uint32)
        n = len(a) + len(b)
        t = np.zeros(max(1, n), np.int32)
        sp = np.zeros(max(1, n), np.uint64)
        dp = np.zeros(max(1, n), np.uint64)
        k = self.lib.ref_editops(metric, _p(a, _u32p), len(a), _p(b, _u32p), len(b), _p(t, _i32p), _p(sp, _u64p),
                                 _p(dp, _u64p))
        if k < 0:
            raise ValueError(f"reference error {k}: {self._err()}")
        return t[:k], sp[:k], dp[:k]

    def partial_ratio_alignment(self, s1, s2, cutoff=0.0):
        """(score, src_start, src_end, dest_start, dest_end) from the reference itself."""
        return _alignment(self.lib.ref_partial_ratio_alignment, s1, s2, cutoff)

    def editops_algebra(self, op, ops, src_len, dest_len, a=0, b=0, c=1, sub=None):
        """The reference's Editops / Opcodes algebra (ref_editops_algebra): ops is an (n, 3) array
        of (type, src, dest); returns None if the reference throws, else (src_len, dest_len, rows)
        with rows (n, 3) editops or (n, 5) opcodes."""
        ops = np.ascontiguousarray(np.asarray(ops, dtype=np.int64).reshape(-1, 3))
        sub = np.ascontiguousarray(np.asarray(sub if sub is not None else np.zeros((0, 3)), dtype=np.int64).reshape(-1, 3))
        out = np.zeros(5 * (2 * len(ops) + 2) + 5, np.int64)
        sl, dl = ctypes.c_size_t(0), ctypes.c_size_t(0)
        k = self.lib.ref_editops_algebra(op, _p(ops, _i64p), len(ops), src_len, dest_len, a, b, c, _p(sub, _i64p),
                                         len(sub), _p(out, _i64p), ctypes.byref(sl), ctypes.byref(dl))
        if k < 0:
            return None
        w = 5 if op in (5, 6, 7) else 3
        return int(sl.value), int(dl.value), out[: k * w].reshape(-1, w)

    def opcodes(self, metric, s1, s2):
        """(n, 5) array: type, src_begin, src_end, dest_begin, dest_end of the reference's Opcodes."""
        a = np.ascontiguousarray(to_codes(s1), dtype=np.uint32)
        b = np.ascontiguousarray(to_codes(s2), dtype=np.uint32)
        out = np.zeros(5 * (2 * (len(a) + len(b)) + 1), np.uint64)
        k = self.lib.ref_opcodes(metric, _p(a, _u32p), len(a), _p(b, _u32p), len(b), _p(out, _u64p))
        if k < 0:
            raise ValueError(f"reference error {k}: {self._err()}")
        return out[:5 * k].reshape(k, 5)

    def editops_counts_u8_mt(self, metric, achars, aoff, bchars, boff, threads=None):
        """Threaded levenshtein_editops / lcs_seq_editops over 8-bit pairs: operations per pair."""
        n = len(aoff) - 1
        out = np.zeros(n, dtype=np.uint64)
        self.lib.ref_editops_u8_mt(metric, _p(achars, _u8p), _p(aoff, _u64p), _p(bchars, _u8p), _p(boff, _u64p), n,
                                   _p(out, _u64p), threads or os.cpu_count() or 1)
        return out

    def partial_ratio_alignment_u8_mt(self, achars, aoff, bchars, boff, cutoff=0.0, threads=None):
        """Threaded fuzz::partial_ratio_alignment over 8-bit pairs: (scores, (n, 4) windows)."""
        n = len(aoff) - 1
        score = np.zeros(n, dtype=np.float64)
        al = np.zeros((n, 4), dtype=np.uint64)
        self.lib.ref_partial_ratio_alignment_u8_mt(_p(achars, _u8p), _p(aoff, _u64p), _p(bchars, _u8p),
                                                   _p(boff, _u64p), n, float(cutoff), _p(score, _f64p),
                                                   _p(al, _u64p), threads or os.cpu_count() or 1)
        return score, al

    def pairs_mt(self, metrics, achars, aoff, bchars, boff, threads=None):
        n = len(aoff) - 1
        threads = threads or os.cpu_count() or 1
        m = np.ascontiguousarray(np.asarray(metrics, dtype=np.int32))
        out = np.zeros(n, dtype=np.uint64)
        self.lib.ref_pairs_u32_mt(_p(m, _i32p), _p(achars, _u32p), _p(aoff, _u64p), _p(bchars, _u32p),
                                  _p(boff, _u64p), n, _p(out, _u64p), threads)
        return out


    def lev_cutoff_mt(self, qchars, qoff, cchars, coff, cutoffs, threads=None):
        """CachedLevenshtein<uint8_t> distance per query with that query's cutoff; [nq, nc] uint64."""
        nq, nc = len(qoff) - 1, len(coff) - 1
        threads = threads or os.cpu_count() or 1
        cut = np.ascontiguousarray(cutoffs, dtype=np.uint64)
        out = np.zeros(nq * nc, dtype=np.uint64)
        self.lib.ref_lev_cutoff_u8_mt(_p(qchars, _u8p), _p(qoff, _u64p), nq, _p(cchars, _u8p), _p(coff, _u64p), nc,
                                      _p(cut, _u64p), _p(out, _u64p), threads)
        return out.reshape(nq, nc)

    def extract_ratio_mt(self, qchars, qoff, cchars, coff, cutoff, k, threads=None):
        """fuzz::CachedRatio extract() + top-k per query; returns (index, score, seconds per query)."""
        nq, nc = len(qoff) - 1, len(coff) - 1
        threads = threads or os.cpu_count() or 1
        idx = np.zeros(nq * k, dtype=np.int64)
        sc = np.zeros(nq * k, dtype=np.float64)
        secs = np.zeros(nq, dtype=np.float64)
        self.lib.ref_extract_ratio_u8_mt(_p(qchars, _u8p), _p(qoff, _u64p), nq, _p(cchars, _u8p),
                                         _p(coff, _u64p), nc, cutoff, k, _p(idx, _i64p), _p(sc, _f64p),
                                         _p(secs, _f64p), threads)
        return idx.reshape(nq, k), sc.reshape(nq, k), secs


def reference_available() -> bool:
    return os.path.exists(REF_SO)
