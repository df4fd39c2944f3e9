"""Acceptance rule of the BOINC server-side validator, restated
(validator/period_search_validator4.cpp:24-109): two outputs match when they
have the same (non-zero) number of lines and, line by line,
|a-b|/(a+b)/2 <= 0.1 for period and rms and <= 0.5 for chi2.  Lines are
read as six numbers; NaN in period, rms or chi2 rejects the file (:52-54).
The reference's read loop (``while (!feof)``) appends the last line twice
when the file ends with a newline; that affects both files equally, so it is
not restated.
"""
import math


def read_rows(text: str):
    rows = []
    for line in text.splitlines():
        parts = line.split()
        if not parts:
            continue
        if len(parts) != 6:
            raise ValueError(f"bad line: {line!r}")
        vals = [float(x) for x in parts]
        if any(math.isnan(v) for v in vals[:3]):
            raise ValueError("NaN in period/rms/chi2")
        rows.append(vals)
    return rows


def compare(a_text: str, b_text: str, tol_per=0.1, tol_rms=0.1, tol_chisq=0.5) -> bool:
    a, b = read_rows(a_text), read_rows(b_text)
    if not a or not b or len(a) != len(b):
        return False
    for ra, rb in zip(a, b):
        for k, tol in ((0, tol_per), (1, tol_rms), (2, tol_chisq)):
            if abs((ra[k] - rb[k]) / (ra[k] + rb[k])) / 2 > tol:
                return False
    return True
