"""Seeded synthetic string generators shared by the tests."""
import numpy as np


def random_strings(rng, n, lo, hi, alpha_lo=97, alpha_n=4):
    out = []
    for _ in range(n):
        m = int(rng.integers(lo, hi + 1))
        out.append(rng.integers(alpha_lo, alpha_lo + alpha_n, size=m).astype(np.uint32))
    return out


def mutate(rng, s, edits, alpha_lo=97, alpha_n=4):
    s = list(int(x) for x in s)
    for _ in range(edits):
        op = int(rng.integers(0, 3))
        if op == 0 and s:
            i = int(rng.integers(0, len(s)))
            s[i] = alpha_lo + int(rng.integers(0, alpha_n))
        elif op == 1:
            i = int(rng.integers(0, len(s) + 1))
            s.insert(i, alpha_lo + int(rng.integers(0, alpha_n)))
        elif s:
            i = int(rng.integers(0, len(s)))
            del s[i]
    return np.array(s, dtype=np.uint32)


def edge_strings():
    """Lengths the reference's tests single out (63/64/65/128/129), empties, equal prefixes."""
    base = [
        [], [97], [97, 97], [98], list(b"aaaa"), list(b"bbbb"), list(b"abaa"), list(b"baaa"),
        list(b"South Korea"), list(b"North Korea"), list(b"aabc"), list(b"cccd"),
        list(b"CA"), list(b"AC"), list(b"ABC"),
    ]
    for L in (31, 32, 33, 63, 64, 65, 127, 128, 129):
        base.append([97] * L)
        base.append([97] * (L - 1) + [98])
        base.append([98] + [97] * (L - 1))
    return [np.array(x, dtype=np.uint32) for x in base]
