"""Writes tests/golden/c1_faithful.json: the oracle's restatement of the
reference's plain path (Scheduler::run_one::<ClassicalMC<IsingModel,
MetropolisCore>>, random-permutation visits, xoshiro256++) on the 64x64
periodic Ising model at T = 2.269185, cold start, base seed 42, 10^4
thermalisation + 10^5 measurement sweeps, binsize 100.

Besides the five binned estimates the file keeps block means of 2000
sweeps, from which the tests take errors that cover the integrated
autocorrelation time near T_c.  Run from the repository root:

    python tests/golden/make_c1_golden.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

L, T, SEED, THERM, MEAS, BINSIZE, BLOCK = 64, 2.269185, 42, 10_000, 100_000, 100, 2000


def main() -> None:
    lat = oracle.Lattice.square(L, L, True)
    t0 = time.time()
    est, n, se, sm = oracle.run_one_ising(lat, 1.0, 1.0 / T, 0, SEED, THERM, MEAS, BINSIZE, want_series=True)
    dt = time.time() - t0
    n_sites = L * L
    series = {"Energy": se, "E2": se * se, "Magnetization": sm, "M2": sm * sm, "M4": (sm * sm) * (sm * sm)}
    out = {
        "config": dict(L=L, T=T, seed=SEED, therm=THERM, meas=MEAS, binsize=BINSIZE, start="cold"),
        "estimates": {k: dict(mean=v.mean, stderr=v.stderr, autocorr_time=v.autocorr_time, n_bins=int(v.n_bins))
                      for k, v in est.items()},
        "block_means": {k: v.reshape(-1, BLOCK).mean(axis=1).tolist() for k, v in series.items()},
        "block": BLOCK,
        "n_sites": n_sites,
        "cpu_seconds": dt,
        "flips": int(n),
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_faithful.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}: {dt:.1f} s, {n / dt / 1e9:.4f} flips/ns")


if __name__ == "__main__":
    main()
