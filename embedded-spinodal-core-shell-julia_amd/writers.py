"""``writedlm(path, A, ',')`` output format (core_shell.jl:98-100).

Julia writes one matrix row per line (row i = first index), elements joined by
``,``, each Float64 printed by ``print``: the shortest round-trip digits, in
plain notation when the decimal-point position ``pt`` (value = 0.D x 10^pt)
satisfies -4 < pt <= 6, else ``d.ddde<exp>``; integral values keep ``.0``.
"""
from __future__ import annotations

import math

import numpy as np


def julia_float(x: float) -> str:
    x = float(x)
    if x != x:
        return "NaN"
    if math.isinf(x):
        return "Inf" if x > 0 else "-Inf"
    if x == 0.0:
        return "-0.0" if math.copysign(1.0, x) < 0 else "0.0"
    r = repr(x)
    sign = ""
    if r[0] == "-":
        sign, r = "-", r[1:]
    if "e" in r:
        mant, e = r.split("e")
        e = int(e)
    else:
        mant, e = r, 0
    if "." in mant:
        ip, fp = mant.split(".")
    else:
        ip, fp = mant, ""
    digits = (ip + fp).lstrip("0")
    # position of the decimal point relative to the first significant digit
    lead = len(ip + fp) - len((ip + fp).lstrip("0"))
    pt = len(ip) - lead + e
    digits = digits.rstrip("0") or "0"
    if -4 < pt <= 6:
        if pt <= 0:
            return sign + "0." + "0" * (-pt) + digits
        if pt >= len(digits):
            return sign + digits + "0" * (pt - len(digits)) + ".0"
        return sign + digits[:pt] + "." + digits[pt:]
    return sign + digits[0] + "." + (digits[1:] or "0") + "e" + str(pt - 1)


def format_matrix(a: np.ndarray) -> str:
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(-1, 1)
    return "".join(",".join(julia_float(v) for v in row) + "\n" for row in a)


def writedlm(path: str, a: np.ndarray) -> None:
    with open(path, "w") as f:
        f.write(format_matrix(a))
