"""Extract the golden vectors of the reference's unit tests into arrangers.npz.

Source: period_search_pure_cpu_win/UnitTest_pure_cpu_win/arrangers.cpp in the
reference checkout -- printed FP64 arrays (%.30f) for nrows=6 (288 facets),
Lmax=Mmax=3 (16 coefficients) on the 118-point input of that directory, used
by its InnerFunctionsTests.cpp.  Every ``double NAME[] = {...};`` and
``int NAME[] = {...};`` initialiser is stored under ``<function>/<NAME>``
(the function the initialiser sits in), values only.

Run from the repository root with the reference checkout present:
    python tests/golden/extract_arrangers.py /root/reference
"""
import os
import re
import sys

import numpy as np

SRC = "period_search_pure_cpu_win/UnitTest_pure_cpu_win/arrangers.cpp"


def extract(text: str) -> dict:
    out = {}
    fn = None
    decl = re.compile(r"\b(double|int)\s+(\w+)\s*\[\s*\d*\s*\]\s*=\s*\{(.*?)\}\s*;", re.S)
    fn_re = re.compile(r"^\s*void\s+(\w+)\s*\(", re.M)
    fpos = [(m.start(), m.group(1)) for m in fn_re.finditer(text)]
    for m in decl.finditer(text):
        owner = None
        for pos, name in fpos:
            if pos < m.start():
                owner = name
        vals = [v for v in (x.strip() for x in m.group(3).split(",")) if v]
        dtype = np.float64 if m.group(1) == "double" else np.int64
        out[f"{owner}/{m.group(2)}"] = np.array([float(v) for v in vals], dtype=dtype)
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, SRC)).read()
    arrays = extract(text)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "arrangers.npz"), **arrays)
    print(f"{len(arrays)} arrays -> {os.path.join(here, 'arrangers.npz')}")


if __name__ == "__main__":
    main()
