"""Writes binom_exact.csv: the closed binomial sum of src/binom_embar.cpp:5-50
evaluated exactly (mpmath, 40 significant digits) for the double-precision
tree parameters u, d, p (q = 1 - p) that the reference computes
(src/binom_embar.cpp:16-24, restated by oracle/oracle.cpp orc_beta_tree,
which the reference-driver fixtures pin bit for bit):

    exp(-rT) * sum_i C(N,i) p^i q^(N-i) max(pf (S0 u^i d^(N-i) - E), 0)

with log C(N,i) from mpmath's loggamma.  Only terms within 40 standard
deviations of N p carry weight (the rest are below 1e-300 relative).

    make -C oracle liboracle.so && python tests/golden/make_binom_exact.py
"""
import csv
import ctypes as C
import math
import os

import mpmath as mp

HERE = os.path.dirname(os.path.abspath(__file__))
L = C.CDLL(os.path.join(HERE, "..", "..", "oracle", "liboracle.so"))
L.orc_beta_tree.argtypes = [C.c_double] * 3 + [C.c_int, C.POINTER(C.c_double)]
mp.mp.dps = 40

CASES = [
    # S0, E, r, sigma, T, N
    (100.0, 100.0, 0.05, 0.2, 1.0, 1000),
    (100.0, 100.0, 0.05, 0.2, 1.0, 100000),
    (100.0, 100.0, 0.05, 0.2, 1.0, 1000000),
    (100.0, 110.0, 0.02, 0.75, 1.0, 10000),
]


def exact(S0, E, r, s, T, N, pf):
    o = (C.c_double * 4)()
    L.orc_beta_tree(r, s, T, N, o)
    u, d, p = mp.mpf(o[0]), mp.mpf(o[1]), mp.mpf(o[2])
    q = mp.mpf(1.0 - o[2])
    lp, lq, lu, ld = mp.log(p), mp.log(q), mp.log(u), mp.log(d)
    lN = mp.loggamma(N + 1)
    mean, sd = N * o[2], math.sqrt(N * o[2] * (1.0 - o[2]))
    lo, hi = max(0, int(mean - 40 * sd)), min(N, int(mean + 40 * sd) + 1)
    tot = mp.mpf(0)
    for i in range(lo, hi + 1):
        v = pf * (S0 * mp.exp(i * lu + (N - i) * ld) - E)
        if v <= 0:
            continue
        tot += mp.exp(lN - mp.loggamma(i + 1) - mp.loggamma(N - i + 1) + i * lp + (N - i) * lq) * v
    return mp.exp(-mp.mpf(r) * T) * tot


rows = []
for S0, E, r, s, T, N in CASES:
    for pf, name in ((1.0, "call"), (-1.0, "put")):
        v = exact(S0, E, r, s, T, N, pf)
        rows.append(["binom_embar", name, S0, E, r, s, T, N, mp.nstr(v, 25)])
        print(rows[-1], flush=True)
with open(os.path.join(HERE, "binom_exact.csv"), "w", newline="") as fh:
    w = csv.writer(fh)
    w.writerow(["prog", "payoff", "S0", "E", "r", "sigma", "T", "N", "exact"])
    w.writerows(rows)
