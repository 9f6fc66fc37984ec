"""CPU checks of the top-k ordering rule the GPU tests compare against."""
import itertools

import numpy as np

from _topk import expected_topk


def test_expected_topk_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(50):
        nc = int(rng.integers(0, 12))
        row = rng.choice([0.0, 0.5, 0.7, 0.7, 1.0], size=nc)
        k = int(rng.integers(1, 6))
        cutoff = float(rng.choice([0.0, 0.6, 0.7]))
        idx, sc = expected_topk(row[None, :], cutoff, k)
        # brute force: all choices passing, sorted by a comparison of (score, index)
        passing = [(row[i], i) for i in range(nc) if row[i] >= cutoff]
        best = sorted(passing, key=lambda t: (-t[0], t[1]))[:k]
        assert [i for _, i in best] == [i for i in idx[0] if i >= 0]
        assert all(sc[0, j] == best[j][0] for j in range(len(best)))
        assert all(sc[0, j] == 0.0 for j in range(len(best), k))


def test_ties_break_by_index():
    row = np.array([0.9, 1.0, 0.9, 1.0, 0.9])
    idx, _ = expected_topk(row[None, :], 0.0, 5)
    assert list(idx[0]) == [1, 3, 0, 2, 4]
    for perm in itertools.permutations(range(3)):
        r = np.array([0.5, 0.5, 0.5])[list(perm)]
        assert list(expected_topk(r[None, :], 0.0, 3)[0][0]) == [0, 1, 2]
