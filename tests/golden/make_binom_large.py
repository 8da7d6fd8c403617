"""Writes binom_large.csv: the closed binomial sum at N = 1e5 and 1e6 
from the oracle's prefix form (oracle/oracle.cpp orc_binom_embar_prefix),
which equals the reference restatement bit for bit (checked in
tests/test_oracle.py) and matches the reference's own N=1e5 result in
ref_kernels.csv.  The reference itself needs ~1 h for N=1e6 at -O0.

    make -C oracle liboracle.so && python tests/golden/make_binom_large.py
"""
import csv
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
L = C.CDLL(os.path.join(HERE, "..", "..", "oracle", "liboracle.so"))
f = L.orc_binom_embar_prefix
f.restype = C.c_double
f.argtypes = [C.c_double] * 5 + [C.c_int, C.c_double, C.c_int]
rows = []
for pf, name in ((1.0, "call"), (-1.0, "put")):
    for N in (100000, 1000000):
        v = f(100.0, 100.0, 0.05, 0.2, 1.0, N, pf, os.cpu_count() or 1)
        rows.append(["binom_embar", name, 100, 100, 0.05, 0.2, 1, N, repr(v)])
        print(rows[-1], flush=True)
with open(os.path.join(HERE, "binom_large.csv"), "w", newline="") as fh:
    w = csv.writer(fh)
    w.writerow(["prog", "payoff", "S0", "E", "r", "sigma", "T", "N", "result"])
    w.writerows(rows)
