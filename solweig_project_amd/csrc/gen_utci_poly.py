#!/usr/bin/env python3
"""Generate utci_poly.inc: the 210-term UTCI polynomial as straight-line code.

Reads the term table ``utci_terms.inc`` (coefficient + exponents of Ta, va,
dTmrt, Pa in summation order) and emits one statement per term, so the
device compiler keeps every power in a register (no dynamically indexed
arrays). Each term is ``coef * Ta^i * va^j * dT^k * Pa^l`` multiplied left to
right and added to the running sum in table order.
"""
import re
from pathlib import Path

HERE = Path(__file__).resolve().parent
src = (HERE / "utci_terms.inc").read_text()
coefs = re.findall(r"^\s*([-+0-9.Ee]+)f,\s*$", src, re.M)
exps = re.findall(r"\{(\d), (\d), (\d), (\d)\}", src)
assert len(coefs) == len(exps) == 210
names = ["ta", "va", "dt", "pa"]
lines = ["/* generated by gen_utci_poly.py from utci_terms.inc -- do not edit */"]
for c, e in zip(coefs, exps):
    expr = f"{c}f"
    for v, k in zip(names, map(int, e)):
        if k:
            expr = f"({expr} * {v}{k})"
    lines.append(f"u = u + {expr};")
(HERE / "utci_poly.inc").write_text("\n".join(lines) + "\n")
print("wrote", len(lines) - 1, "terms")
