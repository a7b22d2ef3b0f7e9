#!/usr/bin/env python3
"""Throughput of the 2-D Ising checkerboard Metropolis sweep on one MI355X.

Workload: 16384 x 16384 periodic square lattice, J = 1, h = 0, int8 +-1
spins from a seeded uniform random start (seed 42), T = 2.269185.  One step
is one full sweep (colour 0 then colour 1) of the lattice; spins stay in HBM
for the whole run.  The line printed reports attempted single-site updates
per nanosecond, the roofline of the dominant kernel (the half-sweep) and the
oracle's faithful-layout CPU sweep timed on this host as a reported
baseline.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 20]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spin-flips/ns (Ising 2D L=16384); batched-queue sweeps/s + p99 latency"
HBM_PEAK_GBS = 8000.0


def traffic_from_profile(path: str):
    """HBM bytes per half-sweep launch from a committed rocprofv3 --pmc summary
    (line 'traffic_bytes_per_launch: <int>'), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def cpu_baseline(seconds_hint: float = 15.0):
    """The oracle's restatement of the reference's plain sweep (f64 spins,
    usize CSR, 32-byte bonds, Fisher-Yates order, xoshiro256++), one thread,
    on a 4096 x 4096 periodic lattice (same per-site work and the same
    out-of-cache access pattern as 16384^2, whose reference layout needs
    ~38 GB)."""
    import oracle

    lat = oracle.Lattice.square(4096, 4096, True)
    beta = 1.0 / 2.269185
    oracle.plain_sweeps(lat, 1.0, beta, 42, 1)  # touch the pages
    t0 = time.perf_counter()
    sweeps = 0
    flips = 0
    while True:
        flips += oracle.plain_sweeps(lat, 1.0, beta, 42 + sweeps, 1)
        sweeps += 1
        if time.perf_counter() - t0 > seconds_hint or sweeps >= 8:
            break
    dt = time.perf_counter() - t0
    return {"value": flips / (dt * 1e9), "unit": "spin-flips/ns", "cores": 1, "kind": "port",
            "sample": f"{sweeps} sweeps of the faithful-layout plain Metropolis path on a "
                      f"4096^2 periodic lattice (oracle.c orc_plain_sweeps), 1 thread, "
                      f"{dt:.1f} s; host nproc={os.cpu_count()}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--L", type=int, default=16384)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build runs on a single GPU "
              "(multi-GPU sharding is described in DESIGN.md, not launched).", file=sys.stderr)
        return 2

    import scamper_amd as sc

    L = args.L
    T = 2.269185
    lat = sc.Lattice.square(L, L, True)
    st = sc.State(lat, sc.model(sc.ISING, 1.0), 1, 42)
    st.init(True)
    st.set_beta(1.0 / T)
    st.sweep(sc.METROPOLIS, args.warmup)
    st.synchronize()
    t0 = time.perf_counter()
    ms = st.sweep_timed(sc.METROPOLIS, args.steps)
    wall = time.perf_counter() - t0
    n = lat.n_sites
    flips = n * args.steps
    value = flips / (ms * 1e6)
    launches = 2 * args.steps
    launch_s = ms * 1e-3 / launches
    bytes_per_launch = 3 * (n // 2)  # read own + read other + write own colour
    achieved = bytes_per_launch / launch_s / 1e9
    e, m = st.observables()

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "spin-flips/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8",
        "data": "synthetic",
        "config": {"workload": "ising2d_checkerboard_metropolis", "L": L, "T": T, "J": 1.0,
                   "spins": "int8 colour-split", "seed": 42, "start": "random",
                   "sweeps_timed": args.steps, "wall_s": round(wall, 3),
                   "final_energy_per_site": e / n, "final_abs_m": m},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile(os.path.join(ROOT, "profiles",
                                                                 "ising2d_L16384_pmc.txt"))},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
