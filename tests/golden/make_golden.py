"""Writes the golden fixtures in this folder from the reference's own code.

Inputs come from oracle/_ref/, built by oracle/Makefile from the reference
sources (main renamed, time() pinned by PCF_FIXTIME).  Outputs:

  ref_kernels.csv  prog,payoff,S0,E,r,sigma,T,N,M,seed,result   (17 digits)
                   result = the reference kernel function's return value
  ref_rows.csv     the reference CLI's stdout row (10 digits), for the
                   drop-in row-format test; timing columns are data only

Run in the container that holds /root/reference:
    make -C oracle && python tests/golden/make_golden.py
"""
import csv
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")

CASES = []
for seed in (1, 42, 1234567):
    for N in (1000, 100000, 1000000):
        CASES.append(("mc_eur", "call", 100, 100, 0.05, 0.2, 1, N, 0, seed))
    CASES.append(("mc_eur", "put", 100, 100, 0.05, 0.2, 1, 100000, 0, seed))
    CASES.append(("mc_eur", "call", 100, 110, 0.02, 0.75, 1, 100000, 0, seed))
for seed in (1, 42):
    for N in (1024, 16384):
        CASES.append(("mc_asia", "call", 100, 100, 0.05, 0.25, 1, N, 252, seed))
    CASES.append(("mc_asia", "put", 100, 100, 0.05, 0.25, 1, 4096, 252, seed))
    CASES.append(("mc_asia", "call", 100, 110, 0.02, 0.75, 1, 4096, 200, seed))
for seed in (1, 42):
    for pf in ("call", "put"):
        for N in (1024, 16384):
            CASES.append(("mc_amer", pf, 100, 100, 0.06, 0.2, 1, N, 50, seed))
    CASES.append(("mc_amer", "call", 100, 110, 0.02, 0.75, 1, 4096, 200, seed))
for N in (1, 2, 3, 100, 101, 1000, 10000, 30000):
    CASES.append(("binom_embar", "call", 100, 100, 0.05, 0.2, 1, N, 0, 0))
CASES.append(("binom_embar", "put", 100, 100, 0.05, 0.2, 1, 1000, 0, 0))
CASES.append(("binom_embar", "call", 100, 110, 0.02, 0.75, 1, 1000, 0, 0))
for N in (1, 2, 3, 129, 640, 1000, 2000, 10000):
    CASES.append(("binom_vanilla_eur", "call", 100, 100, 0.05, 0.2, 1, N, 0, 0))
    CASES.append(("binom_vanilla_amer", "put", 100, 100, 0.06, 0.2, 1, N, 0, 0))
CASES.append(("binom_vanilla_amer", "call", 100, 110, 0.02, 0.75, 1, 1000, 0, 0))
CASES.append(("binom_vanilla_eur", "put", 100, 110, 0.02, 0.75, 1, 1000, 0, 0))
LONG = [("binom_embar", "call", 100, 100, 0.05, 0.2, 1, 100000, 0, 0)]


def run_drv(case):
    prog, pf, S0, E, r, sig, T, N, M, seed = case
    exe = os.path.join(REF, "drv_" + prog)
    args = [exe, pf, str(S0), str(E), str(r), str(sig), str(T), str(N)]
    if M:
        args.append(str(M))
    env = dict(os.environ, PCF_FIXTIME=str(seed))
    out = subprocess.run(args, env=env, check=True, capture_output=True, text=True).stdout.strip()
    return out


def run_cli(case):
    prog, pf, S0, E, r, sig, T, N, M, seed = case
    exe = os.path.join(REF, prog)
    args = [exe, pf, str(S0), str(E), str(r), str(sig), str(T), str(N)]
    if M:
        args.append(str(M))
    env = dict(os.environ, PCF_FIXTIME=str(seed))
    return subprocess.run(args, env=env, check=True, capture_output=True, text=True).stdout.strip()


def main():
    cases = CASES + (LONG if "--long" in sys.argv else [])
    rows = []
    for c in cases:
        rows.append(list(c) + [run_drv(c)])
        print(rows[-1], flush=True)
    path = os.path.join(HERE, "ref_kernels.csv")
    old = []
    if os.path.exists(path) and "--long" not in sys.argv:
        with open(path) as f:
            old = [r for r in csv.reader(f)][1:]
        keep = {tuple(r[:10]) for r in (list(map(str, x[:10])) for x in rows)}
        old = [r for r in old if tuple(r[:10]) not in keep]
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["prog", "payoff", "S0", "E", "r", "sigma", "T", "N", "M", "seed", "result"])
        for r in old:
            w.writerow(r)
        for r in rows:
            w.writerow(r)
    cli_cases = [c for c in CASES if c[7] <= 100000 and c[0] in ("mc_eur", "mc_asia", "mc_amer", "binom_embar")][:24]
    # deterministic programs: one row each
    cli_cases += [("binom_embar", "call", 100, 100, 0.05, 0.2, 1, 1000, 0, 0),
                  ("binom_vanilla_eur", "call", 100, 100, 0.05, 0.2, 1, 2000, 0, 0),
                  ("binom_vanilla_amer", "put", 100, 100, 0.06, 0.2, 1, 2000, 0, 0)]
    with open(os.path.join(HERE, "ref_rows.csv"), "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(["prog", "payoff", "S0", "E", "r", "sigma", "T", "N", "M", "seed", "row"])
        for c in cli_cases:
            w.writerow(list(c) + [run_cli(c)])


if __name__ == "__main__":
    main()
