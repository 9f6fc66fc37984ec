"""Expected top-k from a score matrix: the README's extract() results
(score >= cutoff, README.md:153-214) ordered by (score desc, index asc)."""
import numpy as np


def expected_topk(scores, cutoff, k):
    nq = scores.shape[0]
    idx = np.full((nq, k), -1, dtype=np.int64)
    sc = np.zeros((nq, k), dtype=np.float64)
    for q in range(nq):
        row = scores[q]
        cand = np.flatnonzero(row >= cutoff)
        order = cand[np.lexsort((cand, -row[cand]))][:k]
        idx[q, :len(order)] = order
        sc[q, :len(order)] = row[order]
    return idx, sc


def word_strings(rng, n, vocab, lo_words=1, hi_words=6):
    """Strings of 1..6 words from `vocab` joined by spaces (uint8 code units)."""
    out = []
    for _ in range(n):
        w = int(rng.integers(lo_words, hi_words + 1))
        picks = rng.integers(0, len(vocab), size=w)
        out.append(np.frombuffer(b" ".join(vocab[i] for i in picks), dtype=np.uint8).astype(np.uint32))
    return out


def vocabulary(rng, n=60):
    return [bytes(rng.integers(97, 97 + 8, size=int(rng.integers(2, 7))).astype(np.uint8)) for _ in range(n)]
