"""Regenerate the large byte-string fixtures of the reference's Levenshtein tests.

The reference keeps two OCR transcriptions and two python-Levenshtein issue-9
strings as uint8 arrays in C++ source (test/distance/examples/ocr.cpp,
pythonLevenshteinIssue9.cpp).  This script reads those arrays as data and
writes them as raw .bin files next to itself; it runs only where the
reference sources exist.  Known answers: tests-Levenshtein.cpp:416-468.
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def arrays(path):
    text = open(path).read()
    out = {}
    for m in re.finditer(r"std::vector<uint8_t>\s+(\w+)\s*=\s*\{([^}]*)\}", text):
        out[m.group(1)] = np.array([int(x) for x in m.group(2).replace("\n", " ").split(",") if x.strip()],
                                   dtype=np.uint8)
    return out


def main():
    ex = os.path.join(REF, "test", "distance", "examples")
    a = arrays(os.path.join(ex, "ocr.cpp"))
    b = arrays(os.path.join(ex, "pythonLevenshteinIssue9.cpp"))
    a["ocr_example1"].tofile(os.path.join(HERE, "ocr_example1.bin"))
    a["ocr_example2"].tofile(os.path.join(HERE, "ocr_example2.bin"))
    b["example1"].tofile(os.path.join(HERE, "issue9_example1.bin"))
    b["example2"].tofile(os.path.join(HERE, "issue9_example2.bin"))
    for k, v in list(a.items()) + list(b.items()):
        print(k, v.size)


if __name__ == "__main__":
    main()
