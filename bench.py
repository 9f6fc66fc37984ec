#!/usr/bin/env python3
"""Benchmark: all-pairs Levenshtein scoring on one MI355X.

Workload (the headline configuration): cdist of 4,096 queries x 65,536
choices, uint8 printable ASCII (0x20-0x7E, uniform), lengths uniform 1-64 so
only the single-word path runs.  One step = one full cdist with
CachedLevenshtein::distance semantics and no cutoff; the 2.7e8 uint64 results
stay resident in HBM.  The other scorers of that configuration (Indel, OSA,
Jaro-Winkler 0.1; cutoff off and normalized_similarity >= 0.8) are timed
after the main loop and reported under "scorers".

value = GCUPS = sum(m*n) over all pairs / seconds (full DP matrix).

Printed: one JSON line (rank 0).  The roofline object covers the dominant
kernel (the query-tile kernel, both word widths) and is measured with HIP
events on the engine's stream; the CPU baseline is the reference library
(oracle/_ref, compiled from the reference sources) on a bounded sample, or
the C restatement when that library is absent.
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rapidfuzz-cpp_amd"))

NQ, NC = 4096, 65536
LMIN, LMAX = 1, 64
SEED = 20261203
# algorithmic INT32 ops per 64-cell word-step (2 x the 64-bit operations of the
# recurrence lines): Levenshtein_impl.hpp:285-302 (22), LCSseq_impl.hpp:226-229 (5),
# OSA_impl.hpp:51-70 (27), Jaro_impl.hpp:114-128 per text char (11)
OPS = {0: 44, 1: 10, 3: 54, 5: 22}
# VALU instructions per 64-bit window word-step of the band kernel's inner loop (gfx950): the
# funnel read of the PM window, the addition with its ballot carry-lookahead, the D0 / HP / HN /
# VP / VN terms as 32-bit halves (bitop3), the D0 shift across words, and the two score counters
BAND_GFX_OPS = 29
NAMES = {0: "levenshtein", 1: "indel", 3: "osa", 5: "jaro_winkler"}


def _cores():
    """Threads for the CPU baselines: the host cores this process may use, at most 16 -- the CPU
    share of one GPU on the GPU box (its nproc shows the whole machine)."""
    try:
        return min(len(os.sched_getaffinity(0)), 16)
    except Exception:
        return min(os.cpu_count() or 1, 16)


def host_cpu():
    """Host facts every CPU baseline reports: nproc, the usable cores, the threads used, the model."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except Exception:
        usable = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "usable_cores": usable, "threads": _cores(), "cpu_model": model,
            "thread_cap": "16 threads: the CPU share of one GPU on the GPU box"}


def annotate_cpu(obj, info):
    """Adds the host facts to every cpu_baseline object of a result line."""
    if isinstance(obj, dict):
        for k, v in list(obj.items()):
            if k == "cpu_baseline" and isinstance(v, dict):
                v.setdefault("nproc", info["nproc"])
                v.setdefault("cpu_model", info["cpu_model"])
            else:
                annotate_cpu(v, info)
    elif isinstance(obj, list):
        for v in obj:
            annotate_cpu(v, info)
    return obj


def gen(rng, n):
    lens = rng.integers(LMIN, LMAX + 1, size=n).astype(np.uint64)
    off = np.zeros(n + 1, dtype=np.uint64)
    np.cumsum(lens, out=off[1:])
    chars = rng.integers(0x20, 0x7F, size=int(off[-1]), dtype=np.uint8)
    return chars, off, lens


def cpu_baseline(qc, qo, cc, co, ql, cl, target_s=12.0):
    """Reference CPU path on a bounded sample of queries (all choices)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py

    cores = _cores()
    sum_n = float(cl.sum())
    if oracle_py.reference_available():
        ref = oracle_py.Reference()
        kind = "reference"
        results = {}
        for multi in (False, True):
            nq = 8
            while True:
                idx = np.linspace(0, NQ - 1, nq).astype(np.int64)
                sub = [qc[int(qo[i]):int(qo[i + 1])] for i in idx]
                sq, so = pack_u8(sub)
                t0 = time.perf_counter()
                out = ref.cdist_u8_mt(0, 0, sq, so, cc, co, threads=cores, multi=multi)
                dt = time.perf_counter() - t0
                if dt >= target_s / 4 or nq >= NQ:
                    break
                nq = min(NQ, max(nq * 2, int(nq * (target_s / 2) / max(dt, 1e-3))))
            cells = float(ql[idx].sum()) * sum_n
            results[multi] = (cells / dt / 1e9, nq, dt, idx, out)
        best = max(results, key=lambda k: results[k][0])
        g, nq, dt, idx, out = results[best]
        return {"value": round(g, 3), "unit": "GCUPS", "cores": cores, "kind": kind,
                "sample": f"{nq} of {NQ} queries x all {NC} choices, Levenshtein distance; "
                          f"{'AVX2 experimental::MultiLevenshtein (MaxLen buckets)' if best else 'CachedLevenshtein'}"
                          f" was the faster of the two reference paths ({dt:.1f} s)",
                "other_path_gcups": round(results[not best][0], 3)}, idx, out
    orc = oracle_py.Oracle()
    nq, nc = 2, 4096
    idx = np.arange(nq)
    sub = [qc[int(qo[i]):int(qo[i + 1])] for i in idx]
    chs = [cc[int(co[j]):int(co[j + 1])] for j in range(nc)]
    t0 = time.perf_counter()
    out = orc.cdist(0, 0, sub, chs)
    dt = time.perf_counter() - t0
    cells = float(ql[idx].sum()) * float(cl[:nc].sum())
    return {"value": round(cells / dt / 1e9, 4), "unit": "GCUPS", "cores": 1, "kind": "port",
            "sample": f"{nq} queries x {nc} choices with the C restatement (naive DP), "
                      f"reference library not built"}, None, None


def read_profile_facts(path):
    """`- key: value` lines of a committed rocprofv3 summary (numbers measured on the GPU)."""
    facts = {}
    try:
        for line in open(path):
            if line.startswith("- ") and ":" in line:
                k, v = line[2:].split(":", 1)
                try:
                    facts[k.strip()] = float(v) if "." in v else int(v)
                except ValueError:
                    pass
    except OSError:
        pass
    return facts


def pack_u8(strs):
    off = np.zeros(len(strs) + 1, dtype=np.uint64)
    np.cumsum([len(s) for s in strs], out=off[1:])
    chars = np.concatenate(strs).astype(np.uint8) if off[-1] else np.zeros(1, np.uint8)
    return np.ascontiguousarray(chars), off


def zipf_units(rng, n, s=1.1, alphabet=65536):
    """n code points: Zipf(s) ranks over `alphabet` symbols through a seeded permutation."""
    ranks = np.arange(1, alphabet + 1, dtype=np.float64)
    cdf = np.cumsum(ranks ** -s)
    cdf /= cdf[-1]
    perm = rng.permutation(alphabet).astype(np.uint32)
    idx = np.searchsorted(cdf, rng.random(n), side="right")
    return perm[np.minimum(idx, alphabet - 1)]


def workload_uneven(args, emit=True):
    """1,000,000 independent pairs, log-normal lengths (median 40, sigma 1) clipped to 4-4096,
    uint32 Zipf(1.1) code points over 65,536 symbols, Levenshtein / Indel 50/50, no cutoff."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    N = args.pairs
    rng = np.random.default_rng(SEED + 4)
    la = np.clip(np.rint(np.exp(rng.normal(np.log(40), 1.0, N))), 4, 4096).astype(np.uint64)
    lb = np.clip(np.rint(np.exp(rng.normal(np.log(40), 1.0, N))), 4, 4096).astype(np.uint64)
    ao = np.zeros(N + 1, np.uint64)
    bo = np.zeros(N + 1, np.uint64)
    np.cumsum(la, out=ao[1:])
    np.cumsum(lb, out=bo[1:])
    ac = zipf_units(rng, int(ao[-1]))
    bc = zipf_units(rng, int(bo[-1]))
    metrics = rng.integers(0, 2, N).astype(np.int32)  # 0 Levenshtein, 1 Indel
    cells = float(np.dot(la.astype(np.float64), lb.astype(np.float64)))
    print(f"# uneven: {N} pairs, cells {cells:.3e}, max len {int(max(la.max(), lb.max()))}", file=sys.stderr,
          flush=True)
    ctx = rfa.Context(0)
    A = rfa.Strings(ctx, chars=ac, offsets=ao, width=4)
    B = rfa.Strings(ctx, chars=bc, offsets=bo, width=4)
    out = hip.DeviceBuffer(N * 8)
    stream = ctx.stream
    # the schedule's autotuner (untimed): picks the long-pair threshold on this batch
    t0 = time.perf_counter()
    chain = rfa.pairs_autotune(ctx, A, B, metrics)
    print(f"# uneven: autotuned long-pair threshold {chain} ({time.perf_counter() - t0:.2f} s)", file=sys.stderr,
          flush=True)
    res = {}
    for name, sched in (("bucketed", rfa.SCHEDULE_BUCKETED), ("plain", rfa.SCHEDULE_PLAIN)):
        for _ in range(args.warmup):
            rfa.pairs(ctx, A, B, metrics, schedule=sched, out_device_ptr=out.ptr.value)
        ctx.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        steps = args.steps if sched == rfa.SCHEDULE_BUCKETED else max(1, min(args.steps, 3))
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            rfa.pairs(ctx, A, B, metrics, schedule=sched, out_device_ptr=out.ptr.value)
        e1.record(stream)
        e1.synchronize()
        wall = (time.perf_counter() - t0) / steps * 1e3
        ms = e0.elapsed_ms(e1) / steps
        got = np.zeros(N, np.uint64)
        out.to_host(got)
        res[name] = {"ms": round(ms, 3), "wall_ms": round(wall, 3), "gcups": round(cells / (ms / 1e3) / 1e9, 1),
                     "out": got}
        print(f"# uneven {name}: {ms:.3f} ms ({wall:.3f} wall), {res[name]['gcups']} GCUPS", file=sys.stderr,
              flush=True)
    same = bool(np.array_equal(res["bucketed"]["out"], res["plain"]["out"]))
    cpu = None
    check = "schedules agree" if same else "SCHEDULES DISAGREE"
    # algorithmic INT32 work of the bucketed schedule: the segmented kernel takes the shorter
    # string as the pattern, or the longer one when it has chain-4096 units
    s_len, l_len = np.minimum(la, lb).astype(np.float64), np.maximum(la, lb).astype(np.float64)
    flip = (l_len >= chain) & (l_len <= 4096)
    words = np.ceil(np.where(flip, l_len, s_len) / 64)
    steps = np.where(flip, s_len, l_len)
    ops_per = np.where(metrics == 0, np.where(words == 1, OPS[0], 42.0), 10.0)
    ops = float(np.sum(words * steps * ops_per))
    cus = hip.device_attribute(hip.ATTR_CU_COUNT)
    clk_khz = hip.device_attribute(hip.ATTR_CLOCK_RATE_KHZ)
    peak_tops = cus * 64 * clk_khz * 1e3 / 1e12
    achieved_tops = ops / (res["bucketed"]["ms"] / 1e3) / 1e12
    prof = read_profile_facts(os.path.join(ROOT, "profiles", "pairs_uneven_1M_u32.md"))
    roofline = {"bound": "valu", "achieved": round(achieved_tops, 3), "peak": round(peak_tops, 3), "unit": "TOP/s",
                "frac": round(achieved_tops / peak_tops, 4), "traffic": None,
                "ops_per_step": ops,
                "ops_model": "word-steps x INT32 ops per word-step: Levenshtein 44 (one word) / 42 (block "
                             "word), LCS 10; timed over the whole call (plan + segmented kernel)",
                "kernel": "k_seg<u32, 16 symbol bits, integer epilogue>",
                "valu_issue_utilization": prof.get("valu_issue_utilization"),
                "valu_source": "profiles/pairs_uneven_1M_u32.md"}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        if oracle_py.reference_available():
            ref = oracle_py.Reference()
            cores = _cores()
            k = 20000
            while True:
                t0 = time.perf_counter()
                r = ref.pairs_mt(metrics[:k], ac, ao[:k + 1], bc, bo[:k + 1], threads=cores)
                dt = time.perf_counter() - t0
                if dt > 3 or k >= N:
                    break
                k = min(N, k * 4)
            # repeat the sample until about 10 s of CPU time so the rate is stable
            reps = 1
            while dt < 10:
                t0 = time.perf_counter()
                ref.pairs_mt(metrics[:k], ac, ao[:k + 1], bc, bo[:k + 1], threads=cores)
                dt += time.perf_counter() - t0
                reps += 1
            sc = float(np.dot(la[:k].astype(np.float64), lb[:k].astype(np.float64)))
            bad = int(np.count_nonzero(r != res["bucketed"]["out"][:k]))
            check += "; " + ("bit-exact vs reference on the CPU sample" if bad == 0 else f"{bad} MISMATCHES vs reference")
            cpu = {"value": round(sc * reps / dt / 1e9, 3), "unit": "GCUPS", "cores": cores, "kind": "reference",
                   "sample": f"first {k} of {N} pairs with the free functions levenshtein_distance / "
                             f"indel_distance, {reps} runs ({dt:.1f} s)"}
    line = {
        "metric": "uneven-batch GCUPS, length-bucketed schedule vs plain one-pair-per-lane kernel",
        "value": res["bucketed"]["gcups"], "unit": "GCUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["bucketed"]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"{N} pairs, log-normal lengths median 40 sigma 1 clipped 4-4096, uint32 Zipf(1.1) "
                               f"over 65536 symbols, Levenshtein/Indel 50/50, no cutoff", "cells": cells},
        "plain": {k: v for k, v in res["plain"].items() if k != "out"},
        "bucketed": {k: v for k, v in res["bucketed"].items() if k != "out"},
        "speedup_vs_plain": round(res["plain"]["ms"] / res["bucketed"]["ms"], 2),
        "autotuned_chain_threshold": chain,
        "roofline": roofline, "cpu_baseline": cpu, "check": check,
    }
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


SERVE_CHOICES = 2_000_000
SERVE_QPS = 2000.0
SERVE_K = 5
SERVE_CUTOFF = 70.0


def gen_serving_choices(rng, n=SERVE_CHOICES, vocab_n=5000, zipf_s=1.0):
    """n choices of 1-12 words (uniform) from a seeded vocabulary of vocab_n lowercase words of 2-12
    letters, word ranks Zipf(zipf_s), joined by single spaces; a choice shorter than 4 bytes takes one
    more word, one longer than 128 bytes is cut at 128."""
    vlen = rng.integers(2, 13, size=vocab_n).astype(np.int64)
    voff = np.zeros(vocab_n + 1, np.int64)
    np.cumsum(vlen, out=voff[1:])
    vchars = rng.integers(97, 123, size=int(voff[-1])).astype(np.uint8)
    p = 1.0 / np.arange(1, vocab_n + 1) ** zipf_s
    cdf = np.cumsum(p / p.sum())
    counts = rng.integers(1, 13, size=n).astype(np.int64)
    words = np.minimum(np.searchsorted(cdf, rng.random((n, 13)), side="right"), vocab_n - 1)
    wl = vlen[words]
    used = np.arange(13)[None, :] < counts[:, None]
    total = (wl * used).sum(1) + counts - 1
    counts += (total < 4)
    used = np.arange(13)[None, :] < counts[:, None]
    ids = words[used]                      # row-major: choice by choice
    per_choice = counts
    lens_w = vlen[ids] + 1                 # word + separator
    starts = np.zeros(len(ids) + 1, np.int64)
    np.cumsum(lens_w, out=starts[1:])
    wi = np.repeat(np.arange(len(ids), dtype=np.int64), lens_w)
    o = np.arange(int(starts[-1]), dtype=np.int64) - starts[wi]
    flat = np.where(o < vlen[ids][wi], vchars[np.minimum(voff[ids][wi] + o, len(vchars) - 1)], np.uint8(32))
    # drop each choice's trailing separator and cut choices at 128 bytes
    cend_w = np.cumsum(per_choice) - 1     # last word of each choice
    choice_of_word = np.repeat(np.arange(n, dtype=np.int64), per_choice)
    cstart = starts[np.concatenate([[0], cend_w[:-1] + 1])]
    pos_in_choice = np.arange(int(starts[-1]), dtype=np.int64) - cstart[choice_of_word[wi]]
    clen = np.minimum(starts[cend_w + 1] - cstart - 1, 128)
    keep = pos_in_choice < clen[choice_of_word[wi]]
    chars = np.ascontiguousarray(flat[keep])
    off = np.zeros(n + 1, np.uint64)
    np.cumsum(clen, out=off[1:])
    return chars, off


def gen_requests(rng, cc, co, n):
    """n queries: a uniformly chosen choice with U{0..6} random edits (substitute / insert / delete,
    lowercase letters), padded with random letters to at least 8 bytes, at most 256."""
    nc = len(co) - 1
    picks = rng.integers(0, nc, size=n)
    out = []
    for c in picks:
        s = bytearray(cc[int(co[c]):int(co[c + 1])].tobytes())
        for _ in range(int(rng.integers(0, 7))):
            op = int(rng.integers(0, 3))
            if op == 0 and s:
                s[int(rng.integers(0, len(s)))] = int(rng.integers(97, 123))
            elif op == 1:
                s.insert(int(rng.integers(0, len(s) + 1)), int(rng.integers(97, 123)))
            elif s:
                del s[int(rng.integers(0, len(s)))]
        while len(s) < 8:
            s.append(int(rng.integers(97, 123)))
        out.append(np.frombuffer(bytes(s[:256]), dtype=np.uint8))
    return out


def workload_serve(args, emit=True):
    """Serving: 2,000,000 resident choices, requests Poisson at 2,000 QPS, each the top-5 by
    fuzz::ratio >= 70 (ordered by score desc, index asc); p50/p99 latency from the scheduled arrival
    to the answer, through the library's continuous-batching server."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(SEED + 5)
    t0 = time.perf_counter()
    cc, co = gen_serving_choices(rng)
    nreq = int(SERVE_QPS * args.serve_seconds)
    queries = gen_requests(rng, cc, co, nreq)
    arrivals = np.cumsum(rng.exponential(1.0 / SERVE_QPS, size=nreq))
    clen = np.diff(co)
    print(f"# serve: {len(co) - 1} choices (mean len {clen.mean():.1f}, {int(co[-1])} bytes), {nreq} requests "
          f"over {arrivals[-1]:.1f} s; generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    ctx = rfa.Context(0)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    t0 = time.perf_counter()
    srv = rfa.Server(ctx, cs, metric=rfa.RATIO, kind=rfa.SIMILARITY, score_cutoff=SERVE_CUTOFF, k=SERVE_K,
                     max_batch=args.serve_batch, slots=args.serve_slots, candidate_cap=65536)
    print(f"# serve: server up (index built) in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    # warm-up: a short burst at the same rate
    warm = min(nreq, 400)
    srv.replay(queries[:warm], arrivals[:warm] - arrivals[0])
    st0 = srv.stats()
    t0 = time.perf_counter()
    lat, idx, sc = srv.replay(queries, arrivals)
    wall = time.perf_counter() - t0
    st1 = srv.stats()
    batches = st1["batches"] - st0["batches"]
    served_qps = nreq / max(wall, 1e-9)
    lat_ms = lat * 1e3
    pct = {p: float(np.percentile(lat_ms, p)) for p in (50, 90, 99, 99.9)}
    print(f"# serve: {nreq} requests in {wall:.2f} s ({served_qps:.0f} QPS), p50 {pct[50]:.3f} ms, "
          f"p99 {pct[99]:.3f} ms, max {lat_ms.max():.3f} ms, {batches} batches", file=sys.stderr, flush=True)
    # saturation: every request at t = 0 (the largest rate the server sustains)
    nb = min(nreq, 8192)
    t0 = time.perf_counter()
    blat, bidx, bsc = srv.replay(queries[:nb], np.zeros(nb))
    bwall = time.perf_counter() - t0
    st2 = srv.stats()
    same_burst = bool(np.array_equal(bidx, idx[:nb]) and np.array_equal(bsc.view(np.uint64), sc[:nb].view(np.uint64)))
    print(f"# serve: saturation {nb} requests in {bwall:.2f} s = {nb / bwall:.0f} QPS "
          f"(mean batch {nb / max(1, st2['batches'] - st1['batches']):.1f}); burst == paced: {same_burst}",
          file=sys.stderr, flush=True)
    check = "paced and burst answers identical" if same_burst else "PACED AND BURST ANSWERS DIFFER"
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        if oracle_py.reference_available():
            ref = oracle_py.Reference()
            cores = _cores()
            ns = cores * 2
            while True:
                qc_s, qo_s = pack_u8(queries[:ns])
                t0 = time.perf_counter()
                ri, rs, secs = ref.extract_ratio_mt(qc_s, qo_s, cc, co, SERVE_CUTOFF, SERVE_K, threads=cores)
                dt = time.perf_counter() - t0
                if dt > 5 or ns >= nreq:
                    break
                ns = min(nreq, int(ns * max(2.0, 12.0 / max(dt, 1e-3))))
            bad = int(np.count_nonzero((ri != idx[:ns]) | (rs.view(np.uint64) != sc[:ns].view(np.uint64))))
            check += "; " + (f"top-{SERVE_K} bit-exact vs reference on the CPU sample ({ns} requests)" if bad == 0
                             else f"{bad} MISMATCHES vs reference on {ns} requests")
            cpu = {"value": round(ns / dt, 2), "unit": "requests/s", "cores": cores, "kind": "reference",
                   "sample": f"first {ns} requests, fuzz::CachedRatio extract() over all {len(co) - 1} choices + "
                             f"partial sort, one request per thread ({dt:.1f} s)",
                   "single_request_ms_p50": round(float(np.median(secs)) * 1e3, 1)}
    line = {
        "metric": "served p50/p99 latency at 2000 QPS, top-5 fuzz::ratio >= 70 over 2M resident choices",
        "value": round(pct[99], 3), "unit": "ms (p99)", "n_gpus": 1, "steps": nreq, "warmup": warm,
        "ms_per_step": round(wall / nreq * 1e3, 4), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": f"{len(co) - 1} choices of 1-12 Zipf(1.0) words from a 5000-word vocabulary "
                               f"(4-128 bytes); Poisson {SERVE_QPS:.0f} QPS for {args.serve_seconds} s; "
                               f"queries = choice + 0-6 edits (8-256 bytes); top-{SERVE_K}, ratio >= {SERVE_CUTOFF}",
                   "max_batch": args.serve_batch, "slots": args.serve_slots},
        "latency_ms": {"p50": round(pct[50], 3), "p90": round(pct[90], 3), "p99": round(pct[99], 3),
                       "p99.9": round(pct[99.9], 3), "max": round(float(lat_ms.max()), 3),
                       "mean": round(float(lat_ms.mean()), 3)},
        "served_qps": round(served_qps, 1),
        "choices_scored_per_s": round(served_qps * (len(co) - 1), 0),
        "batches": batches, "mean_batch": round(nreq / max(1, batches), 2),
        "saturation_qps": round(nb / bwall, 1),
        "dense_fallbacks": st2["dense_fallbacks"],
        "cpu_baseline": cpu, "check": check,
    }
    srv.close()
    cs.close()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


def workload_latency(args, emit=True):
    """1-vs-N latency: one query (lowercase a-z, length U{8..32}) against 10,000 candidates, each the
    query with Poisson(4) random edits (substitute to another letter / insert / delete, uniform);
    Levenshtein 1/1/1 distance, no cutoff.  One pair per lane on the GPU (the few-query path)."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(SEED + 1)
    q = rng.integers(97, 123, size=int(rng.integers(8, 33))).astype(np.uint8)
    cands = []
    for _ in range(10000):
        s = list(q)
        for _ in range(int(rng.poisson(4))):
            op = int(rng.integers(0, 3))
            if op == 0 and s:
                i = int(rng.integers(0, len(s)))
                s[i] = 97 + (s[i] - 97 + int(rng.integers(1, 26))) % 26
            elif op == 1:
                s.insert(int(rng.integers(0, len(s) + 1)), int(rng.integers(97, 123)))
            elif s:
                del s[int(rng.integers(0, len(s)))]
        cands.append(np.array(s, dtype=np.uint8))
    cc, co = pack_u8(cands)
    qc, qo = pack_u8([q])
    cells = float(len(q)) * float(co[-1])
    ctx = rfa.Context(0)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    sc = rfa.Scorer(ctx, rfa.LEVENSHTEIN, rfa.Strings(ctx, chars=qc, offsets=qo, width=1))
    out = hip.DeviceBuffer(len(cands) * 8)
    for _ in range(20):
        sc.cdist(cs, rfa.DISTANCE, out_device_ptr=out.ptr.value)
    ctx.synchronize()
    reps = 500
    e0, e1 = hip.Event(), hip.Event()
    e0.record(ctx.stream)
    for _ in range(reps):
        sc.cdist(cs, rfa.DISTANCE, out_device_ptr=out.ptr.value)
    e1.record(ctx.stream)
    e1.synchronize()
    dev_us = e0.elapsed_ms(e1) / reps * 1e3
    lat = np.zeros(reps)
    for i in range(reps):
        t0 = time.perf_counter()
        got = sc.cdist(cs, rfa.DISTANCE)  # results back in host memory
        lat[i] = time.perf_counter() - t0
    p50, p99 = (float(np.percentile(lat, p)) * 1e6 for p in (50, 99))
    print(f"# latency 1x10000: device {dev_us:.1f} us/call, host round trip p50 {p50:.1f} us p99 {p99:.1f} us, "
          f"{cells / (p50 / 1e6) / 1e9:.1f} GCUPS at p50", file=sys.stderr, flush=True)
    cpu, check = None, "not checked"
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        if oracle_py.reference_available():
            ref = oracle_py.Reference()
            cut = np.array([(1 << 63) - 1], np.uint64)
            cores = _cores()

            def median_us(threads, reps=20):
                t = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    ref.lev_cutoff_mt(qc, qo, cc, co, cut, threads=threads)
                    t.append(time.perf_counter() - t0)
                return float(np.median(t)) * 1e6

            r = ref.lev_cutoff_mt(qc, qo, cc, co, cut, threads=cores)
            cpu_us = median_us(cores)
            one_us = median_us(1)
            bad = int(np.count_nonzero(r[0] != got[0]))
            check = "bit-exact vs reference" if bad == 0 else f"{bad} MISMATCHES vs reference"
            cpu = {"value": round(cpu_us, 1), "unit": "us per 1x10000 call", "cores": cores, "kind": "reference",
                   "sample": f"the full workload, CachedLevenshtein<uint8_t> over {cores} std::threads (chunks of "
                             f"256 candidates; thread start-up included), median of 20",
                   "gcups": round(cells / (cpu_us / 1e6) / 1e9, 3),
                   "one_core_us": round(one_us, 1)}
    line = {"metric": "1-vs-10,000 Levenshtein distance latency (host round trip, results in host memory)",
            "value": round(p50, 2), "unit": "us (p50)", "n_gpus": 1, "steps": reps, "warmup": 20,
            "ms_per_step": round(p50 / 1e3, 4), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "u64", "data": "synthetic",
            "config": {"workload": f"1 query (len {len(q)}) x 10000 candidates (query + Poisson(4) edits), a-z, "
                                   f"Levenshtein 1/1/1, no cutoff", "cells": cells},
            "p99_us": round(p99, 2), "device_us_per_call": round(dev_us, 2),
            "gcups_p50": round(cells / (p50 / 1e6) / 1e9, 2), "cpu_baseline": cpu, "check": check}
    out.free()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


def workload_positional(args, emit=True):
    """Hamming (padded), Prefix and Postfix cdist on the headline shape: 4096 x 65536 uint8 printable
    ASCII strings of length U{1..64}, CachedHamming distance / CachedPrefix / CachedPostfix similarity,
    no cutoff.  The kernels write 2.15 GB of results per step, so the bound is HBM."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(SEED)
    qc, qo, ql = gen(rng, NQ)
    cc, co, cl = gen(rng, NC)
    ctx = rfa.Context(0)
    qs = rfa.Strings(ctx, chars=qc, offsets=qo, width=1)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    out = hip.DeviceBuffer(NQ * NC * 8)
    pairs = float(NQ) * float(NC)
    res = {}
    for name, metric, kind in (("hamming", rfa.HAMMING, rfa.DISTANCE), ("prefix", rfa.PREFIX, rfa.SIMILARITY),
                               ("postfix", rfa.POSTFIX, rfa.SIMILARITY)):
        sc = rfa.Scorer(ctx, metric, qs)
        for _ in range(max(1, args.warmup)):
            sc.cdist(cs, kind, out_device_ptr=out.ptr.value)
        ctx.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        steps = max(1, args.steps)
        e0.record(ctx.stream)
        for _ in range(steps):
            sc.cdist(cs, kind, out_device_ptr=out.ptr.value)
        e1.record(ctx.stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        gbps = NQ * NC * 8 / (ms / 1e3) / 1e9
        entry = {"ms": round(ms, 3), "gpairs_per_s": round(pairs / (ms / 1e3) / 1e9, 2),
                 "roofline": {"bound": "hbm", "achieved": round(gbps, 1), "peak": 8000.0, "unit": "GB/s",
                              "frac": round(gbps / 8000.0, 4), "traffic": None,
                              "bytes_model": "the 8-byte results (2.15 GB); strings are L2-resident"}}
        print(f"# positional {name}: {ms:.3f} ms/step, {entry['gpairs_per_s']} Gpairs/s, {gbps:.0f} GB/s of results",
              file=sys.stderr, flush=True)
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle_py
            if oracle_py.reference_available():
                ref = oracle_py.Reference()
                cores = _cores()
                nq = 16
                while True:
                    idx = np.linspace(0, NQ - 1, nq).astype(np.int64)
                    sq, so = pack_u8([qc[int(qo[i]):int(qo[i + 1])] for i in idx])
                    t0 = time.perf_counter()
                    r = ref.cdist_u8_mt(metric, kind, sq, so, cc, co, threads=cores)
                    dt = time.perf_counter() - t0
                    if dt > 2 or nq >= NQ:
                        break
                    nq = min(NQ, nq * 4)
                rows = np.zeros(NC, np.uint64)
                bad = 0
                for k_, i in enumerate(idx):
                    out.to_host(rows, offset=int(i) * NC * 8)
                    bad += int(np.count_nonzero(rows != r[k_]))
                entry["cpu_baseline"] = {"value": round(nq * NC / dt / 1e9, 4), "unit": "Gpairs/s", "cores": cores,
                                         "kind": "reference",
                                         "sample": f"{nq} of {NQ} queries x all {NC} choices, Cached{name.title()} "
                                                   f"({dt:.1f} s)"}
                entry["check"] = "bit-exact vs reference on the CPU sample" if bad == 0 else f"{bad} MISMATCHES"
        sc.close()
        res[name] = entry
    line = {"metric": "Hamming / Prefix / Postfix cdist 4096 x 65536, Gpairs/s", "value": res["hamming"]["gpairs_per_s"],
            "unit": "Gpairs/s", "n_gpus": 1, "steps": max(1, args.steps), "warmup": max(1, args.warmup),
            "ms_per_step": res["hamming"]["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u8", "data": "synthetic",
            "config": {"workload": f"cdist {NQ} x {NC} uint8 printable, len {LMIN}-{LMAX}; Hamming distance, "
                                   f"Prefix / Postfix similarity, no cutoff"},
            **res}
    out.free()
    for x in (qs, cs):
        x.close()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


PQ, PC = 1024, 65536


def gen_partial(rng, nq=PQ, nc=PC):
    """partial_ratio workload: queries of U{8..32} and choices of U{32..128} units over a-z and space
    (space 1 in 6); a quarter of the choices carry a query with 0-2 substitutions written over a
    random position (so high scores occur)."""
    def text(n):
        t = rng.integers(97, 123, size=n).astype(np.uint8)
        t[rng.random(n) < 1 / 6] = 32
        return t

    ql = rng.integers(8, 33, size=nq)
    cl = rng.integers(32, 129, size=nc)
    qs = [text(int(m)) for m in ql]
    cs = [text(int(n)) for n in cl]
    for j in rng.choice(nc, size=nc // 4, replace=False):
        q = qs[int(rng.integers(0, nq))].copy()
        for _ in range(int(rng.integers(0, 3))):
            q[int(rng.integers(0, len(q)))] = int(rng.integers(97, 123))
        at = int(rng.integers(0, len(cs[j]) - len(q) + 1))
        cs[j][at:at + len(q)] = q
    qc, qo = pack_u8(qs)
    cc, co = pack_u8(cs)
    return qc, qo, ql.astype(np.int64), cc, co, cl.astype(np.int64)


def workload_partial(args, emit=True):
    """fuzz::partial_ratio cdist, 1024 word-like queries (8-32 units) x 65,536 text choices (32-128
    units), uint8, no cutoff: the best ratio of each query against every window of each choice.
    Queries of at most 32 units run the tile kernel with 32-bit words."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(SEED + 7)
    qc, qo, ql, cc, co, cl = gen_partial(rng)
    nq, nc = len(ql), len(cl)
    ctx = rfa.Context(0)
    qs = rfa.Strings(ctx, chars=qc, offsets=qo, width=1)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    out = hip.DeviceBuffer(nq * nc * 8)
    sc = rfa.Scorer(ctx, rfa.PARTIAL_RATIO, qs)
    for _ in range(max(1, args.warmup)):
        sc.cdist(cs, rfa.SIMILARITY, out_device_ptr=out.ptr.value)
    ctx.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    steps = max(1, args.steps)
    e0.record(ctx.stream)
    for _ in range(steps):
        sc.cdist(cs, rfa.SIMILARITY, out_device_ptr=out.ptr.value)
    e1.record(ctx.stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    pairs = float(nq) * float(nc)
    # window steps each pair needs (query as the needle: sum over window starts of
    # min(m, n - p); choice as the needle: m windows of n steps)
    mh = np.bincount(ql, minlength=129).astype(np.float64)
    nh = np.bincount(cl, minlength=129).astype(np.float64)
    mv = np.arange(129, dtype=np.float64)
    M, N = np.meshgrid(mv, mv, indexing="ij")
    s1 = np.where(M <= N, M * (M + 1) / 2 + M * (N - M), 0.0)
    s2 = np.where(M >= N, M * N, 0.0)
    wsteps = float(np.sum(np.outer(mh, nh) * (s1 + s2)))
    ops = wsteps * 5.0  # one 32-bit LCS word-step = half of the 10 INT32 ops of a 64-bit one
    tops = ops / (ms / 1e3) / 1e12
    peak = 39.322
    print(f"# partial_ratio: {ms:.3f} ms/step, {pairs / (ms / 1e3) / 1e9:.2f} Gpairs/s, "
          f"{wsteps / (ms / 1e3) / 1e12:.2f} T window-steps/s", file=sys.stderr, flush=True)
    line = {"metric": "partial_ratio cdist 1024 x 65536, Gpairs/s", "value": round(pairs / (ms / 1e3) / 1e9, 3),
            "unit": "Gpairs/s", "n_gpus": 1, "steps": steps, "warmup": max(1, args.warmup),
            "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "u32", "data": "synthetic",
            "config": {"workload": f"fuzz::CachedPartialRatio cdist {nq} x {nc}, uint8 a-z and space, queries "
                                   f"8-32 units, choices 32-128 units (a quarter embed a query), no cutoff"},
            "roofline": {"bound": "valu", "achieved": round(tops, 3), "peak": peak, "unit": "TOP/s",
                         "frac": round(tops / peak, 4), "traffic": None,
                         "ops_model": "window word-steps x 5 INT32 ops (32-bit LCS word-step); "
                                      f"{wsteps:.4g} window steps per call",
                         "kernel": "k_ptile<u8, 32-bit words>"}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        if oracle_py.reference_available():
            ref = oracle_py.Reference()
            cores = _cores()
            k = 4
            while True:
                idx = np.linspace(0, nq - 1, k).astype(np.int64)
                sq, so = pack_u8([qc[int(qo[i]):int(qo[i + 1])] for i in idx])
                t0 = time.perf_counter()
                r = ref.cdist_u8_mt(rfa.PARTIAL_RATIO, rfa.SIMILARITY, sq, so, cc, co, threads=cores)
                dt = time.perf_counter() - t0
                if dt > 3 or k >= nq:
                    break
                k = min(nq, k * 4)
            rows = np.zeros(nc, np.float64)
            bad = 0
            for k_, i in enumerate(idx):
                out.to_host(rows, offset=int(i) * nc * 8)
                bad += int(np.count_nonzero(rows.view(np.uint64) != r[k_].view(np.uint64)))
            line["cpu_baseline"] = {"value": round(k * nc / dt / 1e9, 5), "unit": "Gpairs/s", "cores": cores,
                                    "kind": "reference",
                                    "sample": f"{k} of {nq} queries x all {nc} choices, "
                                              f"fuzz::CachedPartialRatio ({dt:.1f} s)"}
            line["check"] = "bit-exact vs reference on the CPU sample" if bad == 0 else f"{bad} MISMATCHES"
    sc.close()
    out.free()
    for x in (qs, cs):
        x.close()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


WLEV_WEIGHTS = (1, 2, 3)


def workload_wlev(args, emit=True):
    """CachedLevenshtein with weights {insert 1, delete 2, replace 3} (the generalized
    Wagner-Fischer path of Levenshtein.hpp:445-477) on the headline cdist shape: 4096 x 65,536
    printable uint8 strings of 1-64 units, distance, no cutoff."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(SEED)
    qc, qo, ql = gen(rng, NQ)
    cc, co, cl = gen(rng, NC)
    cells = float(ql.sum()) * float(cl.sum())
    ctx = rfa.Context(0)
    qs = rfa.Strings(ctx, chars=qc, offsets=qo, width=1)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    out = hip.DeviceBuffer(NQ * NC * 8)
    sc = rfa.Scorer(ctx, rfa.LEVENSHTEIN, qs, weights=WLEV_WEIGHTS)
    for _ in range(max(1, args.warmup)):
        sc.cdist(cs, rfa.DISTANCE, out_device_ptr=out.ptr.value)
    ctx.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    steps = max(1, args.steps)
    e0.record(ctx.stream)
    for _ in range(steps):
        sc.cdist(cs, rfa.DISTANCE, out_device_ptr=out.ptr.value)
    e1.record(ctx.stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    gcups = cells / (ms / 1e3) / 1e9
    cus = hip.device_attribute(hip.ATTR_CU_COUNT)
    clk_khz = hip.device_attribute(hip.ATTR_CLOCK_RATE_KHZ)
    peak = cus * 64 * clk_khz * 1e3 / 1e12
    ops = cells * 5.0
    tops = ops / (ms / 1e3) / 1e12
    print(f"# weighted levenshtein {WLEV_WEIGHTS}: {ms:.3f} ms/step, {gcups:.1f} GCUPS", file=sys.stderr,
          flush=True)
    line = {"metric": "weighted Levenshtein cdist 4096 x 65536, GCUPS", "value": round(gcups, 1), "unit": "GCUPS",
            "n_gpus": 1, "steps": steps, "warmup": max(1, args.warmup), "ms_per_step": round(ms, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
            "config": {"workload": f"CachedLevenshtein weights {WLEV_WEIGHTS} cdist {NQ} x {NC}, uint8 printable, "
                                   f"lengths {LMIN}-{LMAX}, distance, no cutoff", "cells": cells},
            "roofline": {"bound": "valu", "achieved": round(tops, 3), "peak": round(peak, 3), "unit": "TOP/s",
                         "frac": round(tops / peak, 4), "traffic": None,
                         "ops_model": "5 INT32 ops per DP cell (match-bit extract, multiply-add of the replace "
                                      "weight, two adds, three-way minimum)",
                         "kernel": "k_wtile<u8, 64/32-bit PM words>"}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        if oracle_py.reference_available():
            ref = oracle_py.Reference()
            cores = _cores()
            k = 16
            while True:
                idx = np.linspace(0, NQ - 1, k).astype(np.int64)
                sq, so = pack_u8([qc[int(qo[i]):int(qo[i + 1])] for i in idx])
                t0 = time.perf_counter()
                r = ref.cdist_u8_mt(rfa.LEVENSHTEIN, rfa.DISTANCE, sq, so, cc, co, threads=cores,
                                    weights=WLEV_WEIGHTS)
                dt = time.perf_counter() - t0
                if dt > 3 or k >= NQ:
                    break
                k = min(NQ, k * 4)
            rows = np.zeros(NC, np.uint64)
            bad = 0
            for k_, i in enumerate(idx):
                out.to_host(rows, offset=int(i) * NC * 8)
                bad += int(np.count_nonzero(rows != r[k_]))
            sample_cells = float(sum(ql[int(i)] for i in idx)) * float(cl.sum())
            line["cpu_baseline"] = {"value": round(sample_cells / dt / 1e9, 3), "unit": "GCUPS", "cores": cores,
                                    "kind": "reference",
                                    "sample": f"{k} of {NQ} queries x all {NC} choices, CachedLevenshtein "
                                              f"{WLEV_WEIGHTS} ({dt:.1f} s)"}
            line["check"] = "bit-exact vs reference on the CPU sample" if bad == 0 else f"{bad} MISMATCHES"
    sc.close()
    out.free()
    for x in (qs, cs):
        x.close()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


DL_OPS_PER_CELL = 16


def workload_dl(args, emit=True):
    """experimental::CachedDamerauLevenshtein (Zhao and Sahni, DamerauLevenshtein_impl.hpp:36-121)
    on the headline cdist shape: 4096 x 65,536 printable uint8 strings of 1-64 units, distance,
    no cutoff."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(SEED)
    qc, qo, ql = gen(rng, NQ)
    cc, co, cl = gen(rng, NC)
    cells = float(ql.sum()) * float(cl.sum())
    ctx = rfa.Context(0)
    qs = rfa.Strings(ctx, chars=qc, offsets=qo, width=1)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    out = hip.DeviceBuffer(NQ * NC * 8)
    sc = rfa.Scorer(ctx, rfa.DAMERAU_LEVENSHTEIN, qs)
    for _ in range(max(1, args.warmup)):
        sc.cdist(cs, rfa.DISTANCE, out_device_ptr=out.ptr.value)
    ctx.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    steps = max(1, args.steps)
    e0.record(ctx.stream)
    for _ in range(steps):
        sc.cdist(cs, rfa.DISTANCE, out_device_ptr=out.ptr.value)
    e1.record(ctx.stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / steps
    gcups = cells / (ms / 1e3) / 1e9
    cus = hip.device_attribute(hip.ATTR_CU_COUNT)
    clk_khz = hip.device_attribute(hip.ATTR_CLOCK_RATE_KHZ)
    peak = cus * 64 * clk_khz * 1e3 / 1e12
    tops = cells * DL_OPS_PER_CELL / (ms / 1e3) / 1e12
    print(f"# damerau-levenshtein: {ms:.3f} ms/step, {gcups:.1f} GCUPS", file=sys.stderr, flush=True)
    line = {"metric": "Damerau-Levenshtein cdist 4096 x 65536, GCUPS", "value": round(gcups, 1), "unit": "GCUPS",
            "n_gpus": 1, "steps": steps, "warmup": max(1, args.warmup), "ms_per_step": round(ms, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
            "config": {"workload": f"experimental::CachedDamerauLevenshtein cdist {NQ} x {NC}, uint8 printable, "
                                   f"lengths {LMIN}-{LMAX}, distance, no cutoff", "cells": cells},
            "roofline": {"bound": "valu", "achieved": round(tops, 3), "peak": round(peak, 3), "unit": "TOP/s",
                         "frac": round(tops / peak, 4), "traffic": None,
                         "ops_model": f"{DL_OPS_PER_CELL} INT32 ops per DP cell (4 for the Levenshtein terms, 7 "
                                      "for the two transposition terms, 5 for the match bookkeeping)",
                         "kernel": "k_dltile<u8, 64/32-bit PM words>"}}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        if oracle_py.reference_available():
            ref = oracle_py.Reference()
            cores = _cores()
            k = 16
            while True:
                idx = np.linspace(0, NQ - 1, k).astype(np.int64)
                sq, so = pack_u8([qc[int(qo[i]):int(qo[i + 1])] for i in idx])
                t0 = time.perf_counter()
                r = ref.cdist_u8_mt(rfa.DAMERAU_LEVENSHTEIN, rfa.DISTANCE, sq, so, cc, co, threads=cores)
                dt = time.perf_counter() - t0
                if dt > 3 or k >= NQ:
                    break
                k = min(NQ, k * 4)
            rows = np.zeros(NC, np.uint64)
            bad = 0
            for k_, i in enumerate(idx):
                out.to_host(rows, offset=int(i) * NC * 8)
                bad += int(np.count_nonzero(rows != r[k_]))
            sample_cells = float(sum(ql[int(i)] for i in idx)) * float(cl.sum())
            line["cpu_baseline"] = {"value": round(sample_cells / dt / 1e9, 3), "unit": "GCUPS", "cores": cores,
                                    "kind": "reference",
                                    "sample": f"{k} of {NQ} queries x all {NC} choices, "
                                              f"CachedDamerauLevenshtein ({dt:.1f} s)"}
            line["check"] = "bit-exact vs reference on the CPU sample" if bad == 0 else f"{bad} MISMATCHES"
    sc.close()
    out.free()
    for x in (qs, cs):
        x.close()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


ED_PAIRS = 262144


def workload_editops(args, emit=True):
    """levenshtein_editops over 262,144 pairs (Levenshtein_impl.hpp:1229-1257): printable uint8
    first strings of 1-64 units, each against a copy with 0-8 random edits.  The call hands back
    host arrays, so the timed region is the whole rfa_editops call: item upload, the bit-row
    kernel, the rows' copy back and the host walk."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(SEED + 11)
    la = rng.integers(1, 65, size=ED_PAIRS)
    a_strs, b_strs = [], []
    for m in la:
        t = rng.integers(0x20, 0x7F, size=int(m)).astype(np.uint8)
        u = list(t)
        for _ in range(int(rng.integers(0, 9))):
            op = int(rng.integers(0, 3))
            if op == 0 and u:
                u[int(rng.integers(0, len(u)))] = int(rng.integers(0x20, 0x7F))
            elif op == 1:
                u.insert(int(rng.integers(0, len(u) + 1)), int(rng.integers(0x20, 0x7F)))
            elif u:
                del u[int(rng.integers(0, len(u)))]
        a_strs.append(t)
        b_strs.append(np.array(u, dtype=np.uint8))
    ac, ao = pack_u8(a_strs)
    bc, bo = pack_u8(b_strs)
    ctx = rfa.Context(0)
    A = rfa.Strings(ctx, chars=ac, offsets=ao, width=1)
    B = rfa.Strings(ctx, chars=bc, offsets=bo, width=1)
    off, ops = rfa.editops(ctx, A, B, rfa.LEVENSHTEIN)
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        off, ops = rfa.editops(ctx, A, B, rfa.LEVENSHTEIN)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    rows = float(np.sum(bo[1:] - bo[:-1]))
    print(f"# editops: {ms:.3f} ms per call (host-inclusive), {ED_PAIRS / (ms / 1e3) / 1e6:.2f} Mpairs/s, "
          f"{int(off[-1])} operations", file=sys.stderr, flush=True)
    line = {"metric": "levenshtein_editops 262144 pairs, Mpairs/s", "value": round(ED_PAIRS / (ms / 1e3) / 1e6, 3),
            "unit": "Mpairs/s", "n_gpus": 1, "steps": steps, "warmup": 1, "ms_per_step": round(ms, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
            "config": {"workload": f"levenshtein_editops of {ED_PAIRS} pairs, uint8 printable, first strings 1-64 "
                                   "units, second = first with 0-8 random edits; whole call (host-inclusive)",
                       "operations": int(off[-1]), "text_rows": rows},
            "roofline": None}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        if oracle_py.reference_available():
            ref = oracle_py.Reference()
            cores = _cores()
            t0 = time.perf_counter()
            counts = ref.editops_counts_u8_mt(0, ac, ao, bc, bo, threads=cores)
            dt = time.perf_counter() - t0
            bad = int(np.count_nonzero(counts != (off[1:] - off[:-1])))
            # operation-for-operation on a sample
            for i in range(0, ED_PAIRS, ED_PAIRS // 512):
                exp = ref.editops(0, a_strs[i].astype(np.uint32), b_strs[i].astype(np.uint32))
                got = ops[int(off[i]):int(off[i + 1])]
                if not (np.array_equal(got["type"], exp[0]) and np.array_equal(got["src_pos"], exp[1]) and
                        np.array_equal(got["dest_pos"], exp[2])):
                    bad += 1
            line["cpu_baseline"] = {"value": round(ED_PAIRS / dt / 1e6, 3), "unit": "Mpairs/s", "cores": cores,
                                    "kind": "reference",
                                    "sample": f"all {ED_PAIRS} pairs, rapidfuzz::levenshtein_editops ({dt:.2f} s)"}
            line["check"] = ("operation counts of every pair and 512 whole scripts equal the reference"
                             if bad == 0 else f"{bad} MISMATCHES")
    A.close()
    B.close()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


ED_LONG_PAIRS = 512


def workload_editops_long(args):
    """levenshtein_editops of 512 pairs the reference splits by Hirschberg
    (Levenshtein_impl.hpp:1185-1212): lowercase first strings of 3,000 units, each against a
    copy with 300 random substitutions, insertions and deletions.  The timed region is the whole
    rfa_editops call (the count pass, the level-by-level splits and the leaves' walk)."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(SEED + 12)
    a_strs, b_strs = [], []
    for _ in range(ED_LONG_PAIRS):
        t = rng.integers(97, 123, size=3000).astype(np.uint8)
        u = list(t)
        for _ in range(300):
            op = int(rng.integers(0, 3))
            if op == 0:
                u[int(rng.integers(0, len(u)))] = int(rng.integers(97, 123))
            elif op == 1:
                u.insert(int(rng.integers(0, len(u) + 1)), int(rng.integers(97, 123)))
            else:
                del u[int(rng.integers(0, len(u)))]
        a_strs.append(t)
        b_strs.append(np.array(u, dtype=np.uint8))
    ac, ao = pack_u8(a_strs)
    bc, bo = pack_u8(b_strs)
    ctx = rfa.Context(0)
    A = rfa.Strings(ctx, chars=ac, offsets=ao, width=1)
    B = rfa.Strings(ctx, chars=bc, offsets=bo, width=1)
    for _ in range(max(1, args.warmup)):
        off, ops = rfa.editops(ctx, A, B, rfa.LEVENSHTEIN)
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        off, ops = rfa.editops(ctx, A, B, rfa.LEVENSHTEIN)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    print(f"# editops_long: {ms:.3f} ms per call (host-inclusive), {ED_LONG_PAIRS / (ms / 1e3):.0f} pairs/s, "
          f"{int(off[-1])} operations", file=sys.stderr, flush=True)
    line = {"metric": "levenshtein_editops 512 Hirschberg pairs, pairs/s", "value": round(ED_LONG_PAIRS / (ms / 1e3), 1),
            "unit": "pairs/s", "n_gpus": 1, "steps": steps, "warmup": max(1, args.warmup), "ms_per_step": round(ms, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
            "config": {"workload": f"levenshtein_editops of {ED_LONG_PAIRS} pairs, uint8 lowercase, first strings of "
                                   "3000 units, second = first with 300 random edits (every pair split by Hirschberg); "
                                   "whole call (host-inclusive)", "operations": int(off[-1])},
            "roofline": None}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        if oracle_py.reference_available():
            ref = oracle_py.Reference()
            cores = _cores()
            t0 = time.perf_counter()
            counts = ref.editops_counts_u8_mt(0, ac, ao, bc, bo, threads=cores)
            dt = time.perf_counter() - t0
            bad = int(np.count_nonzero(counts != (off[1:] - off[:-1])))
            for i in range(0, ED_LONG_PAIRS, ED_LONG_PAIRS // 16):
                exp = ref.editops(0, a_strs[i].astype(np.uint32), b_strs[i].astype(np.uint32))
                got = ops[int(off[i]):int(off[i + 1])]
                if not (np.array_equal(got["type"], exp[0]) and np.array_equal(got["src_pos"], exp[1]) and
                        np.array_equal(got["dest_pos"], exp[2])):
                    bad += 1
            line["cpu_baseline"] = {"value": round(ED_LONG_PAIRS / dt, 1), "unit": "pairs/s", "cores": cores,
                                    "kind": "reference",
                                    "sample": f"all {ED_LONG_PAIRS} pairs, rapidfuzz::levenshtein_editops ({dt:.2f} s)"}
            line["check"] = ("operation counts of every pair and 16 whole scripts equal the reference"
                             if bad == 0 else f"{bad} MISMATCHES")
    A.close()
    B.close()
    ctx.close()
    print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


PRA_PAIRS = 65536


def gen_pralign(rng, n=PRA_PAIRS):
    """Needles of 8-128 units (log-uniform) over lowercase letters and spaces; haystacks of the
    needle's length plus 0-512 units, three in four holding a copy of the needle with up to
    one edit per eight units at a random place; every eighth pair in the swapped order."""
    a_strs, b_strs = [], []
    l1 = np.exp(rng.uniform(np.log(8), np.log(129), size=n)).astype(np.int64)
    for i in range(n):
        nd = rng.integers(96, 123, size=int(l1[i])).astype(np.uint8)
        nd[nd == 96] = 32
        extra = int(rng.integers(0, 513))
        hy = rng.integers(96, 123, size=int(l1[i]) + extra).astype(np.uint8)
        hy[hy == 96] = 32
        if i % 4 != 3:
            u = list(nd)
            for _ in range(int(rng.integers(0, 1 + len(u) // 8))):
                op = int(rng.integers(0, 3))
                if op == 0 and u:
                    u[int(rng.integers(0, len(u)))] = int(rng.integers(97, 123))
                elif op == 1:
                    u.insert(int(rng.integers(0, len(u) + 1)), int(rng.integers(97, 123)))
                elif u:
                    del u[int(rng.integers(0, len(u)))]
            at = int(rng.integers(0, extra + 1))
            hy = np.concatenate([hy[:at], np.array(u, dtype=np.uint8), hy[at:]])[:int(l1[i]) + extra]
        if i % 8 == 5:
            nd, hy = hy, nd
        a_strs.append(nd)
        b_strs.append(hy)
    return a_strs, b_strs


def workload_pralign(args, emit=True):
    """fuzz::partial_ratio_alignment over 65,536 pairs (fuzz_impl.hpp:196-230): uint8 needles
    of 8-128 units located in haystacks of up to 640 units.  The call hands back host arrays,
    so the timed region is the whole rfa_partial_ratio_alignment call: the host's grouping of
    the pairs, the item upload, the search kernels and the results' copy back."""
    import rapidfuzz_amd as rfa

    rng = np.random.default_rng(SEED + 13)
    a_strs, b_strs = gen_pralign(rng)
    ac, ao = pack_u8(a_strs)
    bc, bo = pack_u8(b_strs)
    ctx = rfa.Context(0)
    A = rfa.Strings(ctx, chars=ac, offsets=ao, width=1)
    B = rfa.Strings(ctx, chars=bc, offsets=bo, width=1)
    for _ in range(max(10, args.warmup)):  # sub-millisecond calls: let the clocks settle
        got = rfa.partial_ratio_alignment(ctx, A, B, 0.0)
    steps = max(1, args.steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        got = rfa.partial_ratio_alignment(ctx, A, B, 0.0)
    ms = (time.perf_counter() - t0) * 1e3 / steps
    print(f"# partial_ratio_alignment: {ms:.3f} ms per call (host-inclusive), "
          f"{PRA_PAIRS / (ms / 1e3) / 1e6:.3f} Mpairs/s, mean score {float(np.mean(got['score'])):.2f}",
          file=sys.stderr, flush=True)
    line = {"metric": "partial_ratio_alignment 65536 pairs, Mpairs/s",
            "value": round(PRA_PAIRS / (ms / 1e3) / 1e6, 4), "unit": "Mpairs/s", "n_gpus": 1, "steps": steps,
            "warmup": max(10, args.warmup), "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64", "data": "synthetic",
            "config": {"workload": f"fuzz::partial_ratio_alignment of {PRA_PAIRS} pairs, uint8 lowercase and "
                                   "spaces, needles 8-128 units, haystacks needle + 0-512 units (3 in 4 holding "
                                   "a mutated copy); whole call (host-inclusive)"},
            "roofline": None}
    # algorithmic LCS word-steps: every full window (l1 steps) plus the prefix and suffix passes
    # (l1 - 1 and l1 steps), times the needle's words; equal lengths below 100 search twice
    la = (ao[1:] - ao[:-1]).astype(np.float64)
    lb = (bo[1:] - bo[:-1]).astype(np.float64)
    l1, l2 = np.minimum(la, lb), np.maximum(la, lb)
    w = np.ceil(l1 / 64)
    steps_pair = w * ((l2 - l1) * l1 + np.maximum(2 * l1 - 1, 0))
    steps_pair += np.where((l1 == l2) & (l1 > 0) & (got["score"] != 100.0), w * (2 * l1 - 1), 0)
    wsteps = float(steps_pair.sum())
    ops = wsteps * 10.0  # 10 INT32 ops per 64-bit LCS word-step
    peak = 39.322
    tops = ops / (ms / 1e3) / 1e12
    prof = read_profile_facts(os.path.join(ROOT, "profiles", "partial_alignment_65536_pairs.md"))
    kus = prof.get("kernel_us_per_call")
    line["roofline"] = {"bound": "valu", "achieved": round(tops, 3), "peak": peak, "unit": "TOP/s",
                        "frac": round(tops / peak, 4), "traffic": None,
                        "time": "whole call (host-inclusive: grouping, upload, kernels, results' copy back)",
                        "ops_model": f"LCS word-steps of every candidate window x 10 INT32 ops; {wsteps:.4g} "
                                     "word-steps per call",
                        "kernel": "k_pralign<1 word>, k_pralign<2 words>",
                        "kernel_only": ({"achieved": round(ops / (kus * 1e-6) / 1e12, 3),
                                         "frac": round(ops / (kus * 1e-6) / 1e12 / peak, 4),
                                         "kernel_us_per_call": kus,
                                         "valu_issue_utilization": prof.get("valu_issue_utilization"),
                                         "source": "profiles/partial_alignment_65536_pairs.md"} if kus else None)}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        if oracle_py.reference_available():
            ref = oracle_py.Reference()
            cores = _cores()
            reps = 0
            t0 = time.perf_counter()
            while True:  # repeat the whole batch for a stable rate (about 2 s)
                sc, al = ref.partial_ratio_alignment_u8_mt(ac, ao, bc, bo, 0.0, threads=cores)
                reps += 1
                dt = time.perf_counter() - t0
                if dt > 2.0:
                    break
            mine = np.stack([got["src_start"], got["src_end"], got["dest_start"], got["dest_end"]], axis=1)
            bad = int(np.count_nonzero((sc.view(np.uint64) != got["score"].view(np.uint64)) |
                                       np.any(al != mine, axis=1)))
            line["cpu_baseline"] = {"value": round(reps * PRA_PAIRS / dt / 1e6, 4), "unit": "Mpairs/s",
                                    "cores": cores, "kind": "reference",
                                    "sample": f"all {PRA_PAIRS} pairs x {reps}, "
                                              f"rapidfuzz::fuzz::partial_ratio_alignment ({dt:.2f} s)"}
            line["check"] = ("scores and windows of every pair equal the reference" if bad == 0
                             else f"{bad} MISMATCHES")
    A.close()
    B.close()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


def workload_tokens(args, emit=True):
    """token_sort_ratio, partial_token_sort_ratio, QRatio, token_set_ratio, token_ratio, the partial
    token ratios and WRatio cdist on the partial_ratio data (1024
    word-like queries x 65,536 texts over a-z and space), no cutoff.  A token-sort scorer sorts
    the tokens of the choice set on the device on its first call and reuses that copy: the
    steady-state time is `ms`, the one-time sort is `choice_sort_ms`."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(SEED + 7)
    qc, qo, ql, cc, co, cl = gen_partial(rng)
    nq, nc = len(ql), len(cl)
    ctx = rfa.Context(0)
    qs = rfa.Strings(ctx, chars=qc, offsets=qo, width=1)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    out = hip.DeviceBuffer(nq * nc * 8)
    pairs = float(nq) * float(nc)
    res = {}
    for name, metric in (("token_sort_ratio", rfa.TOKEN_SORT_RATIO),
                         ("partial_token_sort_ratio", rfa.PARTIAL_TOKEN_SORT_RATIO), ("qratio", rfa.QRATIO),
                         ("token_set_ratio", rfa.TOKEN_SET_RATIO), ("token_ratio", rfa.TOKEN_RATIO),
                         ("partial_token_set_ratio", rfa.PARTIAL_TOKEN_SET_RATIO),
                         ("partial_token_ratio", rfa.PARTIAL_TOKEN_RATIO), ("wratio", rfa.WRATIO)):
        sc = rfa.Scorer(ctx, metric, qs)
        ctx.synchronize()
        t0 = time.perf_counter()
        sc.cdist(cs, rfa.SIMILARITY, out_device_ptr=out.ptr.value)
        ctx.synchronize()
        first_ms = (time.perf_counter() - t0) * 1e3
        for _ in range(max(0, args.warmup - 1)):
            sc.cdist(cs, rfa.SIMILARITY, out_device_ptr=out.ptr.value)
        ctx.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        steps = max(1, args.steps)
        t0 = time.perf_counter()
        e0.record(ctx.stream)
        for _ in range(steps):
            sc.cdist(cs, rfa.SIMILARITY, out_device_ptr=out.ptr.value)
        e1.record(ctx.stream)
        e1.synchronize()
        wall = (time.perf_counter() - t0) / steps * 1e3
        ms = e0.elapsed_ms(e1) / steps
        entry = {"ms": round(ms, 3), "gpairs_per_s": round(pairs / (ms / 1e3) / 1e9, 3)}
        if metric != rfa.QRATIO:
            entry["choice_sort_ms"] = round(max(0.0, first_ms - wall), 2)
        print(f"# {name}: {ms:.3f} ms/step, {entry['gpairs_per_s']} Gpairs/s"
              + (f", first call +{entry['choice_sort_ms']} ms (choice-set token sort)" if "choice_sort_ms" in entry
                 else ""), file=sys.stderr, flush=True)
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle_py
            if oracle_py.reference_available():
                ref = oracle_py.Reference()
                cores = _cores()
                k = 4
                while True:
                    idx = np.linspace(0, nq - 1, k).astype(np.int64)
                    sq, so = pack_u8([qc[int(qo[i]):int(qo[i + 1])] for i in idx])
                    t0 = time.perf_counter()
                    r = ref.cdist_u8_mt(metric, rfa.SIMILARITY, sq, so, cc, co, threads=cores)
                    dt = time.perf_counter() - t0
                    if dt > 2 or k >= nq:
                        break
                    k = min(nq, k * 4)
                rows = np.zeros(nc, np.float64)
                bad = 0
                for k_, i in enumerate(idx):
                    out.to_host(rows, offset=int(i) * nc * 8)
                    bad += int(np.count_nonzero(rows.view(np.uint64) != r[k_].view(np.uint64)))
                entry["cpu_baseline"] = {"value": round(k * nc / dt / 1e9, 5), "unit": "Gpairs/s", "cores": cores,
                                         "kind": "reference",
                                         "sample": f"{k} of {nq} queries x all {nc} choices ({dt:.1f} s)"}
                entry["check"] = "bit-exact vs reference on the CPU sample" if bad == 0 else f"{bad} MISMATCHES"
        sc.close()
        res[name] = entry
    line = {"metric": "token_sort_ratio cdist 1024 x 65536, Gpairs/s", "value": res["token_sort_ratio"]["gpairs_per_s"],
            "unit": "Gpairs/s", "n_gpus": 1, "steps": max(1, args.steps), "warmup": max(1, args.warmup),
            "ms_per_step": res["token_sort_ratio"]["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u32", "data": "synthetic",
            "config": {"workload": f"fuzz::CachedTokenSortRatio / CachedPartialTokenSortRatio / CachedQRatio / "
                                   f"CachedTokenSetRatio / CachedTokenRatio / CachedPartialTokenSetRatio / CachedPartialTokenRatio / "
                                   f"CachedWRatio cdist "
                                   f"{nq} x {nc}, uint8 a-z and space, queries 8-32 units, choices 32-128 units, "
                                   f"no cutoff"},
            **res}
    out.free()
    for x in (qs, cs):
        x.close()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


def _mutate_dna(rng, s):
    """Per position: substitution 1% (to another base), insertion 0.5% (a random base before it),
    deletion 0.5%."""
    u = rng.random(len(s))
    t = s.copy()
    sub = u < 0.01
    t[sub] = (t[sub] + rng.integers(1, 4, size=int(sub.sum()))) % 4
    ins = (u >= 0.01) & (u < 0.015)
    dele = (u >= 0.015) & (u < 0.02)
    keep = ~dele
    pos = np.flatnonzero(ins)
    t = np.insert(t, pos, rng.integers(0, 4, size=len(pos)).astype(t.dtype))
    keep = np.insert(keep, pos, True)
    return t[keep]


def workload_dna(args, emit=True):
    """Long sequences: 256 queries over {A,C,G,T} of length U{1000..10000} against 16,384 candidates
    (half derived from a random query with 1% substitutions, 0.5% insertions, 0.5% deletions; half
    unrelated, length U{1000..10000}); CachedLevenshtein distance with the cutoff k = floor(f * m) of
    each query, f = 5% and 20%."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(SEED + 3)
    NQ3, NC3 = 256, 16384
    acgt = np.frombuffer(b"ACGT", np.uint8)
    qs = [rng.integers(0, 4, size=int(rng.integers(1000, 10001))).astype(np.uint8) for _ in range(NQ3)]
    cands = []
    for i in range(NC3):
        if i % 2 == 0:
            cands.append(_mutate_dna(rng, qs[int(rng.integers(0, NQ3))]))
        else:
            cands.append(rng.integers(0, 4, size=int(rng.integers(1000, 10001))).astype(np.uint8))
    qs = [acgt[s] for s in qs]
    cands = [acgt[s] for s in cands]
    order = rng.permutation(NC3)  # interleave derived and unrelated candidates
    cands = [cands[i] for i in order]
    qc, qo = pack_u8(qs)
    cc, co = pack_u8(cands)
    ql = np.diff(qo).astype(np.float64)
    cl = np.diff(co).astype(np.float64)
    cells = float(ql.sum()) * float(cl.sum())
    print(f"# dna: {NQ3} x {NC3}, full-matrix cells {cells:.3e}", file=sys.stderr, flush=True)
    ctx = rfa.Context(0)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    # one scorer for all queries; query q is scored with its own cutoff (rfa_cdist_cutoffs)
    scorer = rfa.Scorer(ctx, rfa.LEVENSHTEIN, rfa.Strings(ctx, chars=qc, offsets=qo, width=1))
    out = hip.DeviceBuffer(NQ3 * NC3 * 8)
    res = {}
    for frac in (0.05, 0.20):
        cut = np.floor(frac * ql).astype(np.uint64)

        def step():
            scorer.cdist(cs, rfa.DISTANCE, cutoffs_u=cut, out_device_ptr=out.ptr.value)

        # the band kernel's launch settings per window-size bucket, measured untimed on this call
        t0 = time.perf_counter()
        chosen = scorer.cdist_autotune(cs, rfa.DISTANCE, cutoffs_u=cut)
        tune_s = time.perf_counter() - t0
        tuned = {f"bucket_{b}": {"words_per_lane": int(chosen[b, 0]), "threads": int(chosen[b, 1])}
                 for b in range(6) if chosen[b, 0]}
        print(f"# dna cutoff {frac:.0%}: autotuned {tuned} ({tune_s:.2f} s)", file=sys.stderr, flush=True)
        step()
        ctx.synchronize()
        e0, e1 = hip.Event(), hip.Event()
        steps = max(1, min(args.steps, 3))
        e0.record(ctx.stream)
        for _ in range(steps):
            step()
        e1.record(ctx.stream)
        e1.synchronize()
        ms = e0.elapsed_ms(e1) / steps
        got = np.zeros(NQ3 * NC3, np.uint64)
        out.to_host(got)
        got = got.reshape(NQ3, NC3)
        within = int(np.count_nonzero(got <= cut[:, None]))
        entry = {"ms": round(ms, 2), "gcups": round(cells / (ms / 1e3) / 1e9, 1), "pairs_within_cutoff": within,
                 "autotuned": tuned}
        print(f"# dna cutoff {frac:.0%}: {ms:.2f} ms, {entry['gcups']} GCUPS (full-matrix cells), "
              f"{within} pairs within the cutoff", file=sys.stderr, flush=True)
        # the band kernel's own count of the word-steps it evaluated (one more, untimed call)
        ctx.set_option(rfa.OPTION_COUNT_WORD_STEPS, 1)
        step()
        wsteps = ctx.counter(rfa.COUNTER_BAND_WORD_STEPS)
        npairs = ctx.counter(rfa.COUNTER_BAND_PAIRS)
        ctx.set_option(rfa.OPTION_COUNT_WORD_STEPS, 0)
        cus = hip.device_attribute(hip.ATTR_CU_COUNT)
        peak_tops = cus * 64 * hip.device_attribute(hip.ATTR_CLOCK_RATE_KHZ) * 1e3 / 1e12
        gfx_tops = BAND_GFX_OPS * wsteps / (ms / 1e3) / 1e12
        x86_tops = 42 * wsteps / (ms / 1e3) / 1e12
        prof = read_profile_facts(os.path.join(ROOT, "profiles", "band_dna_256x16384.md"))
        entry["roofline"] = {
            "bound": "valu", "achieved": round(gfx_tops, 3), "peak": round(peak_tops, 3),
            "unit": "TOP/s (INT32 VALU lane-operations)", "frac": round(gfx_tops / peak_tops, 4),
            "traffic": prof.get(f"traffic_bytes_{int(frac * 100)}pct"),
            "traffic_unit": "HBM bytes per call, from FETCH_SIZE / WRITE_SIZE scaled by a 1 GiB copy",
            "word_steps": wsteps, "pairs_run": npairs,
            "ops_model": f"{BAND_GFX_OPS} VALU instructions per 64-bit window word-step (the gfx950 band step of "
                         "the inner loop) x the word-steps the kernel counted on the device",
            "x86_priced": {"achieved": round(x86_tops, 3), "frac_of_peak": round(x86_tops / peak_tops, 4),
                           "ops_model": "42 INT32 ops per word-step (2 x the 21 64-bit ops of the reference's block "
                                        "word, Levenshtein_impl.hpp:700-736)"},
            "kernel": "k_band<multi-word> (diagonal window, floor(64 / words) pairs per wave)",
            "valu_instructions_per_cu_cycle": prof.get(f"valu_instructions_per_cu_cycle_{int(frac * 100)}pct"),
            "lane_instructions_per_word_step": prof.get(f"lane_instructions_per_word_step_{int(frac * 100)}pct"),
            "source": "profiles/band_dna_256x16384.md"}
        print(f"# dna cutoff {frac:.0%}: {wsteps:.4g} word-steps over {npairs} pairs run, "
              f"{gfx_tops:.2f} TOP/s ({gfx_tops / peak_tops:.3f} of peak)", file=sys.stderr, flush=True)
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle_py
            if oracle_py.reference_available():
                ref = oracle_py.Reference()
                ns = 8
                while True:
                    sq, so = pack_u8(qs[:ns])
                    t0 = time.perf_counter()
                    r = ref.lev_cutoff_mt(sq, so, cc, co, cut[:ns], threads=_cores())
                    dt = time.perf_counter() - t0
                    if dt > 4 or ns >= NQ3:
                        break
                    ns = min(NQ3, int(ns * max(2.0, 10.0 / max(dt, 1e-3))))
                bad = int(np.count_nonzero(r != got[:ns]))
                scells = float(ql[:ns].sum()) * float(cl.sum())
                entry["cpu_baseline"] = {"value": round(scells / dt / 1e9, 3), "unit": "GCUPS", "cores": _cores(),
                                         "kind": "reference",
                                         "sample": f"first {ns} of {NQ3} queries x all {NC3} candidates, "
                                                   f"CachedLevenshtein<uint8_t> with the query's cutoff ({dt:.1f} s)"}
                entry["check"] = ("bit-exact vs reference on the CPU sample" if bad == 0
                                  else f"{bad} MISMATCHES vs reference")
                print(f"# dna cutoff {frac:.0%}: CPU {entry['cpu_baseline']['value']} GCUPS; {entry['check']}",
                      file=sys.stderr, flush=True)
        res[f"cutoff_{int(frac * 100)}pct"] = entry
    line = {"metric": "DNA 256 x 16,384 Levenshtein with cutoffs 5% / 20% of the query length, GCUPS (full matrix)",
            "value": res["cutoff_5pct"]["gcups"], "unit": "GCUPS", "n_gpus": 1, "steps": max(1, min(args.steps, 3)),
            "warmup": 1, "ms_per_step": res["cutoff_5pct"]["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64", "data": "synthetic",
            "config": {"workload": f"{NQ3} ACGT queries len 1000-10000 x {NC3} candidates (half derived: 1% sub, "
                                   f"0.5% ins, 0.5% del; half unrelated), cutoff floor(f*m), f in 5%, 20%",
                       "cells": cells},
            "roofline": res["cutoff_5pct"].get("roofline"), "cpu_baseline": res["cutoff_5pct"].get("cpu_baseline"),
            "check": res["cutoff_5pct"].get("check"),
            **res}
    scorer.close()
    out.free()
    ctx.close()
    if emit:
        print(json.dumps(annotate_cpu(line, host_cpu())), flush=True)
    return line


def workload_copy(args):
    """Calibration for the HBM counters: a plain device copy of 1 GiB (torch's elementwise copy
    kernel), so FETCH_SIZE / WRITE_SIZE per dispatch can be compared with 1 GiB read + 1 GiB written."""
    import torch

    n = 1 << 28  # float32 elements = 1 GiB
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"workload": "copy 1 GiB", "ms": round(dt * 1e3, 3), "GBps": round(2 * n * 4 / dt / 1e9, 1)}),
          flush=True)


def workload_scorer(args):
    """One scorer of the per-scorer table alone on the headline inputs (4096 x 65536 printable
    ASCII, lengths 1-64): --metric, --kind and --cutoff pick it.  --warmup calls, then --steps
    timed calls; the line gives the per-call time (for profiling one kernel family)."""
    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip
    rng = np.random.default_rng(SEED)
    qc, qo, ql = gen(rng, NQ)
    cc, co, cl = gen(rng, NC)
    cells = float(ql.sum()) * float(cl.sum())
    ctx = rfa.Context(0)
    qs = rfa.Strings(ctx, chars=qc, offsets=qo, width=1)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    out = hip.DeviceBuffer(NQ * NC * 8)
    sc = rfa.Scorer(ctx, args.metric, qs, prefix_weight=0.1)

    def call():
        sc.cdist(cs, args.kind, cutoff_d=args.cutoff, out_device_ptr=out.ptr.value)

    for _ in range(args.warmup):
        call()
    ctx.synchronize()
    e0, e1 = hip.Event(), hip.Event()
    e0.record(ctx.stream)
    for _ in range(args.steps):
        call()
    e1.record(ctx.stream)
    e1.synchronize()
    ms = e0.elapsed_ms(e1) / max(1, args.steps)
    print(json.dumps({"metric": f"{NAMES.get(args.metric, args.metric)} cdist", "kind": args.kind,
                      "cutoff": args.cutoff, "value": round(cells / (ms / 1e3) / 1e9, 1), "unit": "GCUPS",
                      "ms_per_step": round(ms, 4), "steps": args.steps, "warmup": args.warmup,
                      "config": {"workload": f"cdist {NQ}x{NC} u8 printable, lengths {LMIN}-{LMAX}"}}),
          flush=True)
    for o in (sc, qs, cs):
        o.close()
    out.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="cdist",
                    choices=["cdist", "uneven", "serve", "latency", "dna", "positional", "partial", "tokens", "wlev",
                             "dl", "editops", "editops_long", "pralign", "copy", "scorer"],
                    help="cdist: the headline 4096x65536 configuration (default; also runs the serving leg); "
                         "uneven: 1M independent pairs; serve: the serving leg alone; latency: 1 x 10,000 "
                         "Levenshtein; dna: 256 x 16,384 long sequences with band cutoffs; copy: HBM counter "
                         "calibration (a plain 1 GiB device copy)")
    ap.add_argument("--pairs", type=int, default=1_000_000)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-scorers", action="store_true", help="skip the per-scorer table")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg of the default run")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the latency, DNA and uneven-batch summaries of the default run")
    ap.add_argument("--metric", type=int, default=0, help="scorer workload: metric code (0 Levenshtein, "
                                                             "1 Indel, 3 OSA, 5 Jaro-Winkler)")
    ap.add_argument("--kind", type=int, default=0, help="scorer workload: 0 distance, 1 similarity, "
                                                           "2 normalized distance, 3 normalized similarity")
    ap.add_argument("--cutoff", type=float, default=None, help="scorer workload: cutoff of a normalized kind")
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--serve-batch", type=int, default=64)
    ap.add_argument("--serve-slots", type=int, default=4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is described in DESIGN.md "
                                   "but not built (one-GPU boxes only)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "uneven":
        return workload_uneven(args)
    if args.workload == "serve":
        return workload_serve(args)
    if args.workload == "copy":
        return workload_copy(args)
    if args.workload == "scorer":
        return workload_scorer(args)
    if args.workload == "latency":
        return workload_latency(args)
    if args.workload == "dna":
        return workload_dna(args)
    if args.workload == "positional":
        return workload_positional(args)
    if args.workload == "partial":
        return workload_partial(args)
    if args.workload == "tokens":
        return workload_tokens(args)
    if args.workload == "wlev":
        return workload_wlev(args)
    if args.workload == "dl":
        return workload_dl(args)
    if args.workload == "editops":
        return workload_editops(args)
    if args.workload == "editops_long":
        return workload_editops_long(args)
    if args.workload == "pralign":
        return workload_pralign(args)

    import rapidfuzz_amd as rfa
    from rapidfuzz_amd import hip

    rng = np.random.default_rng(SEED)
    qc, qo, ql = gen(rng, NQ)
    cc, co, cl = gen(rng, NC)
    cells = float(ql.sum()) * float(cl.sum())
    print(f"# workload: cdist {NQ}x{NC} u8 printable, len {LMIN}-{LMAX}, cells {cells:.3e}", file=sys.stderr,
          flush=True)

    ctx = rfa.Context(0)
    qs = rfa.Strings(ctx, chars=qc, offsets=qo, width=1)
    cs = rfa.Strings(ctx, chars=cc, offsets=co, width=1)
    out = hip.DeviceBuffer(NQ * NC * 8)
    stream = ctx.stream

    lev = rfa.Scorer(ctx, rfa.LEVENSHTEIN, qs)

    def step(sc, kind=rfa.DISTANCE, cu=None, cd=None):
        sc.cdist(cs, kind, cutoff_u=cu, cutoff_d=cd, out_device_ptr=out.ptr.value)

    for _ in range(args.warmup):
        step(lev)
    ctx.synchronize()
    ev0, ev1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    ev0.record(stream)
    for i in range(args.steps):
        step(lev)
    ev1.record(stream)
    ev1.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_ms(ev1)
    ms_per_step = gpu_ms / args.steps
    gcups = cells / (ms_per_step / 1e3) / 1e9
    print(f"# levenshtein: {ms_per_step:.3f} ms/step (wall {wall / args.steps * 1e3:.3f}), {gcups:.1f} GCUPS",
          file=sys.stderr, flush=True)

    # roofline of the query-tile kernel: every step is two launches of it (64-bit and
    # 32-bit word tiles) and nothing else, so the event time is the kernel time.
    cus = hip.device_attribute(hip.ATTR_CU_COUNT)
    clk_khz = hip.device_attribute(hip.ATTR_CLOCK_RATE_KHZ)
    peak_tops = cus * 64 * clk_khz * 1e3 / 1e12
    # 64-cell word-steps: a query of <= 64 units is one 64-bit word per text position;
    # tiles whose queries fit in 32 units run 32-bit words, counted as half a word-step.
    n64 = float(np.count_nonzero(ql > 32))
    word_steps = (n64 + 0.5 * (NQ - n64)) * float(cl.sum())
    ops = OPS[0] * word_steps
    achieved_tops = ops / (ms_per_step / 1e3) / 1e12
    # algorithmic HBM bytes: the results, plus every query tile (64 queries) reading each
    # choice's units and its two offsets once, plus the queries
    hbm_bytes = float(NQ * NC * 8) + (NQ / 64) * float(cl.sum() + 16 * NC) + float(ql.sum())

    # spot check: rows of the resident result against the CPU baseline sample
    check = "not checked"
    cpu = None
    if not args.no_cpu:
        cpu, idx, ref_out = cpu_baseline(qc, qo, cc, co, ql, cl)
        if idx is not None:
            rows = np.zeros(NC, dtype=np.uint64)
            bad = 0
            for k, i in enumerate(idx):
                out.to_host(rows, offset=int(i) * NC * 8)
                bad += int(np.count_nonzero(rows != ref_out[k]))
            check = "bit-exact vs reference on the CPU sample" if bad == 0 else f"{bad} MISMATCHES vs reference"

    scorers = {}
    if not args.no_scorers:
        # every row is checked against the reference library on 8 query rows x all choices
        ref = None
        if not args.no_cpu:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle_py
            if oracle_py.reference_available():
                ref = oracle_py.Reference()
        sample = np.linspace(0, NQ - 1, 8).astype(np.int64)
        sq, so = pack_u8([qc[int(qo[i]):int(qo[i + 1])] for i in sample])
        for metric in (0, 1, 3, 5):
            sc = lev if metric == 0 else rfa.Scorer(ctx, metric, qs, prefix_weight=0.1)
            for mode in ("no_cutoff", "norm_sim_0.8"):
                if mode == "no_cutoff":
                    kind = rfa.SIMILARITY if metric == 5 else rfa.DISTANCE
                    kw = {}
                else:
                    kind = rfa.NORMALIZED_SIMILARITY
                    kw = {"cd": 0.8}
                step(sc, kind, **kw)
                ctx.synchronize()
                e0, e1 = hip.Event(), hip.Event()
                reps = 3
                e0.record(stream)
                for _ in range(reps):
                    step(sc, kind, **kw)
                e1.record(stream)
                e1.synchronize()
                ms = e0.elapsed_ms(e1) / reps
                row_check = "not checked"
                if ref is not None:
                    exp = ref.cdist_u8_mt(metric, kind, sq, so, cc, co, cutoff_d=kw.get("cd"), threads=_cores())
                    bad = 0
                    got_row = np.zeros(NC, dtype=np.uint64)
                    for k, i in enumerate(sample):
                        out.to_host(got_row, offset=int(i) * NC * 8)
                        bad += int(np.count_nonzero(got_row != exp[k].view(np.uint64)))
                    row_check = ("bit-exact vs reference on 8 query rows x all choices" if bad == 0
                                 else f"{bad} MISMATCHES vs reference")
                scorers[f"{NAMES[metric]}/{mode}"] = {
                    "ms": round(ms, 3), "gcups": round(cells / (ms / 1e3) / 1e9, 1),
                    "x86_priced_int_tops": round(OPS[metric] * word_steps / (ms / 1e3) / 1e12, 2),
                    "check": row_check}
            if sc is not lev:
                sc.close()

    prof = read_profile_facts(os.path.join(ROOT, "profiles", "tile_levenshtein_cdist_4096x65536.md"))
    # The roofline prices a word-step with the VALU instructions gfx950 needs for the Hyyro-2003
    # recurrence: 20 per 64-bit word (two 32-bit halves: bitop3 folds the NOR/OR/XOR terms,
    # v_lshl_add_u64 the 64-bit add, v_alignbit the carried shifts) and 10 per 32-bit word (tiles of
    # queries <= 32 units), read off this build's inner loop; the PM-row broadcast (v_readlane), the
    # LDS address and the per-choice epilogue are not counted, so the fraction is a lower bound on
    # the VALU issue rate.  Peak: CUs x 64 lanes x clock lane-operations per second.  The survey's
    # x86 pricing (44 INT32 ops per 64-cell word-step) counts operations gfx950 fuses, so it is kept
    # beside it (x86_priced), not as the fraction.
    GFX_OPS_64, GFX_OPS_32 = 20, 10
    lane_steps_64 = n64 * float(cl.sum())
    lane_steps_32 = (NQ - n64) * float(cl.sum())
    gfx_ops = GFX_OPS_64 * lane_steps_64 + GFX_OPS_32 * lane_steps_32
    gfx_tops = gfx_ops / (ms_per_step / 1e3) / 1e12
    res = {
        "metric": "Levenshtein 1-vs-N GCUPS + % INT roofline; served p50/p99 latency at QPS",
        "value": round(gcups, 1),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"cdist {NQ} queries x {NC} choices, uint8 printable ASCII, lengths uniform "
                               f"{LMIN}-{LMAX}, CachedLevenshtein distance, no cutoff (single-word path)",
                   "queries": NQ, "choices": NC, "cells": cells, "seed": SEED},
        "roofline": {"bound": "valu", "achieved": round(gfx_tops, 3), "peak": round(peak_tops, 3),
                     "unit": "TOP/s (INT32 VALU lane-operations)", "frac": round(gfx_tops / peak_tops, 4),
                     "traffic": prof.get("traffic_bytes_per_step"),
                     "traffic_unit": "HBM bytes per step (the 64-bit + 32-bit launch pair), from FETCH_SIZE / "
                                     "WRITE_SIZE scaled by a 1 GiB copy",
                     "ops_per_step": gfx_ops,
                     "ops_model": f"{GFX_OPS_64} VALU instructions per 64-bit word-step and {GFX_OPS_32} per "
                                  "32-bit word-step (the gfx950 Hyyro-2003 recurrence of the inner loop) x lane "
                                  "word-steps",
                     "x86_priced": {"achieved": round(achieved_tops, 3), "ops_per_step": ops,
                                    "ops_model": "44 INT32 ops per 64-cell word-step (2 x the 22 64-bit ops of "
                                                 "Levenshtein_impl.hpp:285-302); 32-bit-word tiles count half a "
                                                 "word-step", "frac_of_peak": round(achieved_tops / peak_tops, 4)},
                     "kernel": "k_tile<LEV,u8,64/32-bit words,integer epilogue>",
                     "valu_issue_utilization": {"w64": prof.get("valu_issue_utilization_64"),
                                                "w32": prof.get("valu_issue_utilization_32"),
                                                "source": "profiles/tile_levenshtein_cdist_4096x65536.md"},
                     "hbm_algorithmic_GBps": round(hbm_bytes / (ms_per_step / 1e3) / 1e9, 1),
                     "hbm_frac": round(hbm_bytes / (ms_per_step / 1e3) / 8e12, 4)},
        "cpu_baseline": cpu,
        "check": check,
        "scorers": scorers,
    }
    out.free()
    for s_ in (lev, qs, cs, ctx):
        s_.close()
    gc.collect()
    if not args.no_serve:
        srv = workload_serve(args, emit=False)
        res["serving"] = {k: srv[k] for k in ("latency_ms", "served_qps", "choices_scored_per_s", "mean_batch",
                                              "saturation_qps", "dense_fallbacks", "cpu_baseline", "check")}
        res["serving"]["config"] = srv["config"]["workload"]
    if not args.no_configs:
        # the other configurations, each summarized from its own workload line
        keep = ("metric", "value", "unit", "ms_per_step", "higher_is_better", "roofline", "cpu_baseline", "check")
        res["configs"] = {}
        lat = workload_latency(args, emit=False)
        res["configs"]["latency_1x10000"] = {**{k: lat.get(k) for k in keep},
                                             "p99_us": lat["p99_us"], "device_us_per_call": lat["device_us_per_call"],
                                             "config": lat["config"]["workload"]}
        dna = workload_dna(args, emit=False)
        res["configs"]["dna_256x16384"] = {**{k: dna.get(k) for k in keep}, "cutoff_5pct": dna["cutoff_5pct"],
                                           "cutoff_20pct": dna["cutoff_20pct"], "config": dna["config"]["workload"]}
        pos = workload_positional(args, emit=False)
        res["configs"]["positional_4096x65536"] = {k: pos[k] for k in ("metric", "value", "unit", "ms_per_step",
                                                                        "hamming", "prefix", "postfix")}
        par = workload_partial(args, emit=False)
        res["configs"]["partial_ratio_1024x65536"] = {**{k: par.get(k) for k in keep},
                                                      "config": par["config"]["workload"]}
        tok = workload_tokens(args, emit=False)
        res["configs"]["token_scorers_1024x65536"] = {k: tok[k] for k in ("metric", "value", "unit", "ms_per_step",
                                                                           "token_sort_ratio",
                                                                           "partial_token_sort_ratio", "qratio",
                                                                           "token_set_ratio", "token_ratio",
                                                                           "partial_token_set_ratio",
                                                                           "partial_token_ratio", "wratio")}
        wl = workload_wlev(args, emit=False)
        res["configs"]["weighted_levenshtein_4096x65536"] = {**{k: wl.get(k) for k in keep},
                                                             "config": wl["config"]["workload"]}
        dlv = workload_dl(args, emit=False)
        res["configs"]["damerau_levenshtein_4096x65536"] = {**{k: dlv.get(k) for k in keep},
                                                            "config": dlv["config"]["workload"]}
        edo = workload_editops(args, emit=False)
        res["configs"]["levenshtein_editops_262144"] = {k: edo.get(k) for k in ("metric", "value", "unit",
                                                                                 "ms_per_step", "cpu_baseline",
                                                                                 "check")}
        pra = workload_pralign(args, emit=False)
        res["configs"]["partial_ratio_alignment_65536"] = {k: pra.get(k) for k in ("metric", "value", "unit",
                                                                                    "ms_per_step", "roofline",
                                                                                    "cpu_baseline", "check")}
        une = workload_uneven(args, emit=False)
        res["configs"]["uneven_1M_pairs"] = {**{k: une.get(k) for k in keep}, "plain": une["plain"],
                                             "speedup_vs_plain": une["speedup_vs_plain"],
                                             "autotuned_chain_threshold": une["autotuned_chain_threshold"],
                                             "config": une["config"]["workload"]}
    info = host_cpu()
    res["host_cpu"] = info
    print(json.dumps(annotate_cpu(res, info)), flush=True)


if __name__ == "__main__":
    main()
