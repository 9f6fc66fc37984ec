"""Loader for tests/golden/known_answers.json (vectors from the reference's tests)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def expand(spec):
    """A string spec of the JSON file as uint32 code units."""
    if isinstance(spec, str):
        return np.array([ord(c) for c in spec], dtype=np.uint32)
    if "repeat" in spec:
        return np.array([ord(spec["repeat"])] * spec["n"], dtype=np.uint32)
    if "concat" in spec:
        out = []
        for part, n in spec["concat"]:
            out.extend([ord(c) for c in part] * n)
        return np.array(out, dtype=np.uint32)
    raise ValueError(spec)


def expand_list(spec):
    if isinstance(spec, list):
        return [expand(s) for s in spec]
    out = []
    for part, n in spec["concat_list"]:
        out.extend([expand(part)] * n)
    return out


def load():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


def fixture(name, sub=None):
    a = np.fromfile(os.path.join(GOLDEN, name), dtype=np.uint8).astype(np.uint32)
    if sub:
        a = a[sub[0]:sub[0] + sub[1]]
    return a


def case_id(c):
    return c["src"].split("/")[-1]
