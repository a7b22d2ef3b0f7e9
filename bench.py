#!/usr/bin/env python3
"""Throughput of the classical lattice sweep on one MI355X.

Default line (the headline): the 2-D Ising checkerboard Metropolis sweep on a
16384 x 16384 periodic square lattice, J = 1, h = 0, int8 +-1 spins from a
seeded uniform random start (seed 42), T = 2.269185.  One step is one full
sweep (colour 0 then colour 1); spins stay in HBM for the whole run.  The
line reports attempted single-site updates per nanosecond, the roofline of
the dominant kernel (the half-sweep), the oracle's faithful-layout CPU sweep
timed on this host, and under "serving" the batched job queue: Poisson
streams of Potts requests served by the continuous-batching pool (lattice
sweeps/s, spin-flips/ns, p50/p99 completion latency) and the same requests
as plain per-job launches on the same GPU.

Other workloads, one JSON line each (not the headline):
    --workload pt3d   64-replica multi-spin parallel tempering, 3-D L=128
    --workload on3    32-replica O(3) heat bath + over-relaxation, 3-D L=96
    --workload xy     32-replica O(2) (XY) von Mises heat bath + over-relaxation, 3-D L=96
    --workload glass  64-replica 3-D Edwards-Anderson spin glass (Gaussian couplings, L=32),
                      Metropolis through the f64 graph kernels
    --workload tri    32-replica O(3) antiferromagnet on the triangular lattice (L=256),
                      heat bath + over-relaxation through the graph-lattice O(N) kernels
    --workload potts3d  the tiled Potts kernel alone: q=3, 3-D L=256, measured sweeps
    --workload cluster  Swendsen-Wang and Wolff updates of the 2-D Ising model at
                        beta_c, L=4096, and of the 3-D O(3) model near T_c, L=96
    --workload serve  the serving block alone
    --workload calibrate  kernels of known traffic for scaling rocprofv3 counters

    python bench.py [--gpus 1] [--steps 2000] [--warmup 200]
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


# wave64 VALU instructions issued per second by the whole chip: 256 CUs x 4
# SIMDs at 2.4 GHz, one wave64 instruction per SIMD every 4 cycles
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


def figure_from_profile(path: str, key: str):
    """An integer figure from a committed rocprofv3 summary (a line
    '<key>: <int>'), or None."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return int(line.split(":", 1)[1].strip())
    except OSError:
        pass
    return None


def traffic_from_profile(path: str):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (line
    'traffic_bytes_per_launch: <int>'), or None."""
    return figure_from_profile(path, "traffic_bytes_per_launch")


def host_threads() -> int:
    """Host cores this process may use: OMP_NUM_THREADS where it is set (the
    CPU share given to the job), else the affinity mask."""
    try:
        n = int(os.environ.get("OMP_NUM_THREADS", "0"))
    except ValueError:
        n = 0
    return n if n > 0 else len(os.sched_getaffinity(0))


def cpu_baseline(single_sweeps: int = 2):
    """The oracle's restatement of the reference's plain sweep (f64 spins,
    usize CSR, 32-byte bonds, Fisher-Yates order, xoshiro256++) on a 4096 x
    4096 periodic lattice (same per-site work and the same out-of-cache
    access pattern as 16384^2, whose reference layout needs ~38 GB).

    The reference's only parallelism is independent chains, one per worker
    thread (Scheduler::run_parallel, Carlo.rs/src/scheduler.rs:345-405), so
    the value is that: one chain per host core, one sweep each, all at
    once, over one shared lattice.  The single-thread rate of the same sweep
    is reported beside it."""
    import oracle

    lat = oracle.Lattice.square(4096, 4096, True)
    beta = 1.0 / 2.269185
    t0 = time.perf_counter()
    flips = sum(oracle.plain_sweeps(lat, 1.0, beta, 42 + k, 1) for k in range(single_sweeps))
    dt1 = time.perf_counter() - t0
    threads = host_threads()
    t0 = time.perf_counter()
    flips_all = oracle.plain_chains(lat, 1.0, beta, 1000, 1, threads, threads)
    dta = time.perf_counter() - t0
    return {"value": flips_all / (dta * 1e9), "unit": "spin-flips/ns", "cores": threads, "kind": "port",
            "sample": f"{threads} independent chains on {threads} threads, one sweep each of the faithful-layout "
                      f"plain Metropolis path on a 4096^2 periodic lattice (oracle.c orc_plain_chains), "
                      f"{dta:.1f} s; host nproc={os.cpu_count()}",
            "single_thread": {"value": flips / (dt1 * 1e9), "cores": 1,
                              "sample": f"{single_sweeps} sweeps of the same path on one thread, {dt1:.1f} s"}}


def _cpu_timed(step, units_per_call: float, unit: str, what: str, seconds: float = 6.0, max_calls: int = 400) -> dict:
    """A bounded single-thread timing of one oracle path (the port of the
    reference's CPU path): `step(i)` runs call i; one untimed call first."""
    step(0)
    t0 = time.perf_counter()
    calls = 0
    while True:
        calls += 1
        step(calls)
        if time.perf_counter() - t0 > seconds or calls >= max_calls:
            break
    dt = time.perf_counter() - t0
    return {"value": units_per_call * calls / (dt * 1e9), "unit": unit, "cores": 1, "kind": "port",
            "sample": f"{calls} x {what}, 1 thread, {dt:.1f} s; host nproc={os.cpu_count()}"}


def cpu_pt3d() -> dict:
    """MultiSpinIsing's sweep restated with one beta per bit (oracle
    orc_multispin_sweep), 3-D L=32, the 64 temperatures of the GPU run."""
    import numpy as np

    import oracle

    lat = oracle.Lattice.hypercubic([32, 32, 32], True)
    words = np.random.default_rng(1).integers(0, 2**63, lat.n_sites, dtype=np.uint64)
    betas = [1.0 / (4.30 * (4.72 / 4.30) ** (k / 63)) for k in range(64)]
    return _cpu_timed(lambda i: oracle.multispin_sweep(lat, words, betas, 1.0, 42, 0, i), 64 * lat.n_sites,
                      "spin-flips/ns", "64-bit multi-spin sweeps of a 32^3 periodic lattice (orc_multispin_sweep)")


def cpu_on3() -> dict:
    """The O(3) heat bath and over-relaxation in the reference's layout
    (f64 site-major spins over the CSR lattice, orc_sweep_on3_heat_bath and
    orc_sweep_over_relaxation), 3-D L=24 at T = 1.40."""
    import oracle

    lat = oracle.Lattice.hypercubic([24, 24, 24], True)
    m = oracle.model(oracle.ON, 1.0, 2, 3)
    key = oracle.replica_key(42, 0)
    ch = oracle.Chain(lat, m, oracle.random_start(lat, m, key))

    def step(i):
        ch.on3_heat_bath(1 / 1.40, key, i)
        ch.over_relaxation()
    return _cpu_timed(step, 2 * lat.n_sites, "site-updates/ns",
                      "heat-bath + over-relaxation sweeps of a 24^3 periodic lattice (f64, CSR)")


def on3_oracle_error(sc, st, L: int, betas, seed: int, replicas=(0, 15, 31)) -> dict:
    """Error of the FP32 device path against the FP64 oracle at the O(3)
    workload's own shape: after the timed sweeps, replicas first, middle and
    last are read back, one more hybrid sweep (heat bath, models.rs:617-662,
    then over-relaxation, microcanonical.rs:42-80) runs on the device with
    its fused FP64 measurement, and the oracle replays that sweep from the
    same spins with the same draws.  Per replica: the largest and the mean
    per-component spin error, and the errors of E (relative) and |M|
    (absolute, |M| per site)."""
    import numpy as np

    import oracle

    olat = oracle.Lattice.hypercubic([L, L, L], True)
    m = oracle.model(oracle.ON, 1.0, 2, 3)
    starts = {r: st.download_f64(r) for r in replicas}
    sweep = st.sweeps
    st.sweep(sc.HYBRID_HB_OR, 1, measure=True)
    rows = []
    for r in replicas:
        ch = oracle.Chain(olat, m, starts[r].copy())
        ch.on3_heat_bath(betas[r], oracle.replica_key(seed, r), sweep)
        ch.over_relaxation()
        err = np.abs(st.download_f64(r) - ch.spins)
        e, mag = st.series(r)
        e_ref, m_ref = ch.total_energy(), ch.magnetization()
        rows.append({"replica": r, "T": round(1.0 / betas[r], 4), "spin_max_abs": float(err.max()),
                     "spin_mean_abs": float(err.mean()), "energy_rel": float(abs(e[-1] - e_ref) / abs(e_ref)),
                     "abs_m_abs": float(abs(mag[-1] - m_ref))})
    return {"oracle": "orc_sweep_on3_heat_bath + orc_sweep_over_relaxation, f64, same draws and start",
            "sweep": sweep, "replicas": rows,
            "spin_max_abs": max(x["spin_max_abs"] for x in rows),
            "energy_rel_max": max(x["energy_rel"] for x in rows),
            "abs_m_abs_max": max(x["abs_m_abs"] for x in rows)}


def cpu_xy() -> dict:
    """The XY heat bath (Best-Fisher von Mises draws) and over-relaxation in
    the reference's layout (f64 site-major spins over the CSR lattice,
    orc_sweep_on2_heat_bath and orc_sweep_over_relaxation), 3-D L=24 at
    T = 2.15."""
    import oracle

    lat = oracle.Lattice.hypercubic([24, 24, 24], True)
    m = oracle.model(oracle.ON, 1.0, 2, 2)
    key = oracle.replica_key(42, 0)
    ch = oracle.Chain(lat, m, oracle.random_start(lat, m, key))

    def step(i):
        ch.on2_heat_bath(1 / 2.15, key, i)
        ch.over_relaxation()
    return _cpu_timed(step, 2 * lat.n_sites, "site-updates/ns",
                      "XY heat-bath + over-relaxation sweeps of a 24^3 periodic lattice (f64, CSR)")


def _glass_edges(L: int, seed: int = 7):
    """Periodic L^3 Edwards-Anderson couplings: edges +x, +y, +z of every
    site (build_hypercubic order), Gaussian weights."""
    import numpy as np

    rng = np.random.default_rng(seed)
    src, dst = [], []
    for z in range(L):
        for y in range(L):
            for x in range(L):
                s = x + L * (y + L * z)
                src += [s, s, s]
                dst += [(x + 1) % L + L * (y + L * z), x + L * ((y + 1) % L + L * z), x + L * (y + L * ((z + 1) % L))]
    return L ** 3, src, dst, list(rng.normal(size=len(src)))


def cpu_glass() -> dict:
    """The spin-glass Metropolis sweep in the reference's layout (f64 spins,
    CSR lattice with real weights, orc_sweep_metropolis), 3-D L=16."""
    import oracle

    n, src, dst, w = _glass_edges(16)
    lat = oracle.Lattice.from_edges(n, src, dst, w)
    m = oracle.model(oracle.ISING, 1.0, 2)
    key = oracle.replica_key(42, 0)
    ch = oracle.Chain(lat, m, oracle.random_start(lat, m, key))
    return _cpu_timed(lambda i: ch.metropolis(1.0, key, i), n, "spin-flips/ns",
                      "Metropolis sweeps of a 16^3 Gaussian spin glass (f64, CSR)")


def cpu_potts3d() -> dict:
    """The 3-state Potts Metropolis sweep in the reference's layout (f64
    spins, CSR lattice, orc_sweep_metropolis), 3-D L=32 at beta = 1."""
    import oracle

    lat = oracle.Lattice.hypercubic([32, 32, 32], True)
    m = oracle.model(oracle.POTTS, 1.0, 3)
    key = oracle.replica_key(42, 0)
    ch = oracle.Chain(lat, m, oracle.random_start(lat, m, key))
    return _cpu_timed(lambda i: ch.metropolis(1.0, key, i), lat.n_sites, "spin-flips/ns",
                      "Metropolis sweeps of 3-state Potts on a 32^3 periodic lattice (f64, CSR)")


def cpu_cluster() -> dict:
    """Swendsen-Wang updates restated (orc_cluster_steps: the reference's
    edge loop and union-find over the CSR lattice), 2-D Ising 1024^2 at
    beta_c."""
    import oracle

    lat = oracle.Lattice.square(1024, 1024, True)
    m = oracle.model(oracle.ISING)
    key = oracle.replica_key(42, 0)
    s = oracle.random_start(lat, m, key)
    return _cpu_timed(lambda i: oracle.cluster_steps(lat, m, s, 0.4406868, key, i, 1, False), lat.n_sites,
                      "site-updates/ns", "Swendsen-Wang updates of a 1024^2 periodic lattice (f64, CSR)")


def ising2d(sc, args) -> dict:
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
    value = n * args.steps / (ms * 1e6)
    launch_s = ms * 1e-3 / (2 * args.steps)
    bytes_per_launch = 3 * (n // 2)  # read own + read other + write own colour
    achieved = bytes_per_launch / launch_s / 1e9
    e, m = st.observables()
    return {
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


def serving(sc, args) -> dict:
    """Potts request streams on the pool, and the ratio against plain
    per-job launches on a shared subset (all arriving at once)."""
    from scamper_amd import serve as sv

    out = {"mix": "Potts q in {2,3,4}; 2-D L in [32,2048] or 3-D L in [16,256] log-uniform (even); "
                  "T in [0.95,1.05] T_c; sweeps log-uniform in [1e3,1e5], 10% thermalisation; "
                  "Metropolis, random start, binsize 100",
           "rates": []}
    for rate in args.serve_rates:
        reqs = sv.potts_mix(args.serve_requests, rate, seed=11)
        res = sv.serve(reqs)
        row = {"lambda_per_s": rate, **res.summary(reqs), "launch_shape": res.shape}
        out["rates"].append(row)
        print(f"# serving lambda={rate}: {json.dumps(row)}", file=sys.stderr, flush=True)
    # Same requests both ways, all arriving at once; budgets a tenth of the
    # serving mix's so that the one-job-at-a-time leg stays short (both legs
    # scale linearly with the budget).
    # Two plain forms: scamper_run_one per job (the single-job path, int8
    # spins), and the pool's own kernels on one job at a time with host-issued
    # launches (bit planes for q <= 4); the speedup is taken against the faster.
    sub = sv.potts_mix(args.plain_requests, 1e9, seed=13, sweeps=(1_000, 10_000))
    batched = sv.serve(sub, keep=len(sub))
    plain = sv.plain(sub, keep=len(sub))
    plain_same = sv.plain_pool(sub, keep=len(sub))

    def same_as(r):
        return all(batched.estimates[k][name].mean == r.estimates[k][name].mean and
                   batched.estimates[k][name].stderr == r.estimates[k][name].stderr
                   for k in range(len(sub)) for name in ("Energy", "Magnetization", "M2", "M4"))

    best = max(plain.flips_per_ns, plain_same.flips_per_ns)
    out["plain_vs_batched"] = {
        "requests": len(sub),
        "sweeps": "log-uniform in [1e3, 1e4]",
        "site_updates": sum(r.updates for r in sub),
        "batched_spin_flips_per_ns": round(batched.flips_per_ns, 3),
        "plain_spin_flips_per_ns": round(plain.flips_per_ns, 3),
        "plain_same_kernel_spin_flips_per_ns": round(plain_same.flips_per_ns, 3),
        "speedup": round(batched.flips_per_ns / best, 2),
        "speedup_vs_run_one": round(batched.flips_per_ns / plain.flips_per_ns, 2),
        "speedup_vs_same_kernel": round(batched.flips_per_ns / plain_same.flips_per_ns, 2),
        "batched_wall_s": round(batched.wall_s, 3),
        "plain_wall_s": round(plain.wall_s, 3),
        "plain_same_kernel_wall_s": round(plain_same.wall_s, 3),
        "estimates_bit_identical": bool(same_as(plain) and same_as(plain_same)),
    }
    return out


def pt3d(sc, args) -> dict:
    """64 temperatures 4.30..4.72 (geometric) of 3-D Ising, L=128, one bit per
    replica; each step one checkerboard Metropolis sweep then one even/odd
    exchange round."""
    L = 128
    lat = sc.Lattice.hypercubic([L, L, L], True)
    ms_ = sc.MultiSpin(lat, 1.0, 1, 42)
    temps = [4.30 * (4.72 / 4.30) ** (k / 63) for k in range(64)]
    ms_.set_betas([1.0 / t for t in temps])
    ms_.init(True)
    ms_.step(args.warmup, exchange=True)
    ms = ms_.step_timed(args.pt_steps, exchange=True)
    n = lat.n_sites
    flips = 64 * n * args.pt_steps
    acc, ties, ex = ms_.stats()
    bytes_per_step = 3 * 8 * n  # u64 words: read own + read other + write own colour, both colours
    # the state (16 MiB) stays in the caches; the bound is VALU issue, with
    # the instructions per step (both colours and the exchange) counted by
    # rocprofv3 on this workload
    valu_per_step = figure_from_profile(os.path.join(ROOT, "profiles", "pt3d_L128_multispin_kernel_trace.txt"),
                                        "valu_instructions_per_step")
    valu_rate = valu_per_step and valu_per_step * args.pt_steps / (ms * 1e-3)
    return {"metric": "spin-flips/ns (3-D Ising L=128, 64-replica multi-spin PT)",
            "value": round(flips / (ms * 1e6), 3), "unit": "spin-flips/ns", "n_gpus": 1,
            "steps": args.pt_steps, "warmup": args.warmup, "ms_per_step": ms / args.pt_steps,
            "higher_is_better": True, "dtype": "u64 bit-planes", "data": "synthetic",
            "config": {"workload": "pt3d_multispin", "L": L, "replicas": 64, "T": [4.30, 4.72],
                       "ladder": "geometric", "seed": 42, "exchanges_accepted": acc,
                       "near_tie_flags": ties, "exchange_steps": ex},
            "roofline": {"bound": "valu", "achieved": valu_rate and round(valu_rate / 1e12, 4),
                         "peak": round(VALU_PEAK_WAVE_INSTS / 1e12, 4), "unit": "T wave64 VALU instructions/s",
                         "frac": valu_rate and round(valu_rate / VALU_PEAK_WAVE_INSTS, 4),
                         "valu_instructions_per_step": valu_per_step,
                         "hbm_equivalent_GBps": round(bytes_per_step * args.pt_steps / (ms * 1e-3) / 1e9, 1)},
            **({} if args.no_cpu_baseline else {"cpu_baseline": cpu_pt3d()})}


def on3(sc, args) -> dict:
    """32 replicas of the 3-D O(3) model, L=96, T = 1.40 + 0.09 r/31; each
    step one heat-bath sweep then one over-relaxation sweep."""
    L, R = 96, 32
    lat = sc.Lattice.hypercubic([L, L, L], True)
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, 3), R, 42)
    betas = [1.0 / (1.40 + 0.09 * r / 31) for r in range(R)]
    for r in range(R):
        st.init(True, r)
        st.set_beta(betas[r], r)
    st.sweep(sc.HYBRID_HB_OR, args.warmup)
    st.synchronize()
    ms = st.sweep_timed(sc.HYBRID_HB_OR, args.on3_steps)
    n = lat.n_sites
    updates = 2 * R * n * args.on3_steps  # heat bath + over-relaxation
    # per site-update: read own 12 B + read 6 neighbours' planes (12 B each,
    # once per site with perfect reuse) + write 12 B = 36 B
    gbs = 36 * updates / (ms * 1e-3) / 1e9
    return {"metric": "site-updates/ns (3-D O(3) L=96, 32 replicas, HB+OR)",
            "value": round(updates / (ms * 1e6), 3), "unit": "site-updates/ns", "n_gpus": 1,
            "steps": args.on3_steps, "warmup": args.warmup, "ms_per_step": ms / args.on3_steps,
            "higher_is_better": True, "dtype": "f32 spins, f64 observables", "data": "synthetic",
            "config": {"workload": "on3_hybrid", "L": L, "replicas": R, "T": [1.40, 1.49], "seed": 42},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4),
                         "traffic": traffic_from_profile(os.path.join(ROOT, "profiles", "on3_L96_pmc.txt"))},
            **({} if args.no_cpu_baseline else {"cpu_baseline": cpu_on3(),
                                                "fp64_oracle_error": on3_oracle_error(sc, st, L, betas, 42)})}


def xy(sc, args) -> dict:
    """32 replicas of the 3-D XY model, L=96, T = 2.15 + 0.1 r/31 (around
    T_c = 2.2018); each step one von Mises heat-bath sweep then one
    over-relaxation sweep."""
    L, R = 96, 32
    lat = sc.Lattice.hypercubic([L, L, L], True)
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, 2), R, 42)
    for r in range(R):
        st.init(True, r)
        st.set_beta(1.0 / (2.15 + 0.1 * r / 31), r)
    st.sweep(sc.HYBRID_HB_OR, args.warmup)
    st.synchronize()
    ms = st.sweep_timed(sc.HYBRID_HB_OR, args.on3_steps)
    n = lat.n_sites
    updates = 2 * R * n * args.on3_steps
    # per site-update: read own 8 B + neighbours' planes 8 B (with reuse) + write 8 B = 24 B
    gbs = 24 * updates / (ms * 1e-3) / 1e9
    return {"metric": "site-updates/ns (3-D XY L=96, 32 replicas, HB+OR)",
            "value": round(updates / (ms * 1e6), 3), "unit": "site-updates/ns", "n_gpus": 1,
            "steps": args.on3_steps, "warmup": args.warmup, "ms_per_step": ms / args.on3_steps,
            "higher_is_better": True, "dtype": "f32 spins, f64 draws and observables", "data": "synthetic",
            "config": {"workload": "xy_hybrid", "L": L, "replicas": R, "T": [2.15, 2.25], "seed": 42},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4)},
            **({} if args.no_cpu_baseline else {"cpu_baseline": cpu_xy()})}


def glass(sc, args) -> dict:
    """64 replicas of the 3-D Edwards-Anderson spin glass, L=32, Gaussian
    couplings, T = 1.0; Metropolis sweeps through the f64 graph kernels (the
    same ΔE and acceptance arithmetic as the reference)."""
    L, R = 32, 64
    n, src, dst, w = _glass_edges(L)
    lat = sc.Lattice.from_edges(n, src, dst, w)
    st = sc.State(lat, sc.model(sc.ISING, 1.0), R, 42)
    for r in range(R):
        st.init(True, r)
        st.set_beta(1.0, r)
    st.sweep(sc.METROPOLIS, args.warmup)
    st.synchronize()
    ms = st.sweep_timed(sc.METROPOLIS, args.glass_steps)
    updates = R * n * args.glass_steps
    # per update: own spin read + write (2 B), 6 neighbour spins (6 B), 6
    # incidences of neighbour index + f64 weight (72 B), offsets (8 B) and
    # class list (4 B): 92 B, most of it the replica-shared graph
    gbs = 92 * updates / (ms * 1e-3) / 1e9
    return {"metric": "spin-flips/ns (3-D Gaussian spin glass L=32, 64 replicas)",
            "value": round(updates / (ms * 1e6), 3), "unit": "spin-flips/ns", "n_gpus": 1,
            "steps": args.glass_steps, "warmup": args.warmup, "ms_per_step": ms / args.glass_steps,
            "higher_is_better": True, "dtype": "int8 spins, f64 energies", "data": "synthetic",
            "config": {"workload": "spin_glass_metropolis", "L": L, "replicas": R, "T": 1.0, "couplings": "gaussian",
                       "seed": 42},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4)},
            **({} if args.no_cpu_baseline else {"cpu_baseline": cpu_glass()})}


def cpu_tri() -> dict:
    """O(3) heat bath + over-relaxation on the triangular lattice in the
    reference's layout (f64, CSR), L=32, J = -1, T = 0.5."""
    import oracle

    lat = oracle.Lattice.triangular(32, 32)
    m = oracle.model(oracle.ON, -1.0, 2, 3)
    key = oracle.replica_key(42, 0)
    ch = oracle.Chain(lat, m, oracle.random_start(lat, m, key))

    def step(i):
        ch.on3_heat_bath(2.0, key, i)
        ch.over_relaxation()
    return _cpu_timed(step, 2 * lat.n_sites, "site-updates/ns",
                      "heat-bath + over-relaxation sweeps of a 32x32 triangular antiferromagnet (f64, CSR)")


def tri(sc, args) -> dict:
    """32 replicas of the classical Heisenberg antiferromagnet (J = -1) on the
    256 x 256 triangular lattice at T = 0.5; each step one heat-bath sweep and
    one over-relaxation sweep over the lattice's colour classes."""
    L, R = 256, 32
    lat = sc.Lattice.triangular(L, L)
    st = sc.State(lat, sc.model(sc.ON, -1.0, 2, 3), R, 42)
    for r in range(R):
        st.init(True, r)
        st.set_beta(2.0, r)
    st.sweep(sc.HYBRID_HB_OR, args.warmup)
    st.synchronize()
    ms = st.sweep_timed(sc.HYBRID_HB_OR, args.on3_steps)
    updates = 2 * R * lat.n_sites * args.on3_steps
    # per update: own 12 B read + write, 6 neighbours' 12 B (with reuse), 6
    # incidences of neighbour index + f64 weight (72 B, replica-shared), 8 B
    # offsets, 4 B class list: 120 B
    gbs = 120 * updates / (ms * 1e-3) / 1e9
    return {"metric": "site-updates/ns (O(3) triangular antiferromagnet L=256, 32 replicas, HB+OR)",
            "value": round(updates / (ms * 1e6), 3), "unit": "site-updates/ns", "n_gpus": 1,
            "steps": args.on3_steps, "warmup": args.warmup, "ms_per_step": ms / args.on3_steps,
            "higher_is_better": True, "dtype": "f32 spins, f64 observables", "data": "synthetic",
            "config": {"workload": "on3_triangular_af", "L": L, "replicas": R, "J": -1.0, "T": 0.5, "seed": 42},
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4)},
            **({} if args.no_cpu_baseline else {"cpu_baseline": cpu_tri()})}


def potts3d(sc, args) -> dict:
    """The tiled int8 kernel on one large lattice through the state path (no
    graphs, so a profiler sees every launch): 3-state Potts, 3-D L=256, J=1,
    beta=1.0 (T_c is 1/0.550565), Metropolis with the fused per-sweep
    measurement on every sweep; then the same lattice as a job of the pool,
    which runs it on bit planes (pool_bit_planes)."""
    L, q = 256, 3
    lat = sc.Lattice.hypercubic([L, L, L], True)
    st = sc.State(lat, sc.model(sc.POTTS, 1.0, q), 1, 42)
    st.init(True)
    st.set_beta(1.0)
    st.sweep(sc.METROPOLIS, args.warmup, measure=True)
    st.synchronize()
    ms = st.sweep_timed(sc.METROPOLIS, args.potts_steps, measure=True)
    n = lat.n_sites
    flips = n * args.potts_steps
    # per site-update: read own 1 B + the other colour 1 B (neighbours through
    # the caches) + write own 1 B
    gbs = 3 * flips / (ms * 1e-3) / 1e9
    # the same lattice, seed and rule as one job of the pool (bit planes):
    # chunks of 32 sweeps, the job's planes packed at its first chunk and
    # unpacked at its last; the clock covers a second job of the same shape
    # (the pool's chunk graphs are captured by the first) with its admission
    pool_steps = max(32, args.potts_steps // 32 * 32)
    spec = sc.job_spec((L, L, L), sc.model(sc.POTTS, 1.0, q), 1.0, 0, pool_steps, binsize=8, seed=42,
                       random_start=True)
    pool = sc.Pool(spin_budget_bytes=1 << 26, max_jobs=4, chunk_sweeps=32, grid=2048)
    pool.submit(spec)
    pool.run()
    t = time.perf_counter()
    pool.submit(spec)
    pool.run()
    pool_rate = n * pool_steps / ((time.perf_counter() - t) * 1e9)
    return {"metric": "spin-flips/ns (3-state Potts 3-D L=256, tiled kernel, measured sweeps)",
            "value": round(flips / (ms * 1e6), 3), "unit": "spin-flips/ns", "n_gpus": 1,
            "pool_bit_planes": {"value": round(pool_rate, 3), "unit": "spin-flips/ns", "sweeps": pool_steps,
                                "chunk_sweeps": 32},
            "steps": args.potts_steps, "warmup": args.warmup, "ms_per_step": ms / args.potts_steps,
            "higher_is_better": True, "dtype": "int8 states", "data": "synthetic",
            "config": {"workload": "potts3d_tiles", "L": L, "q": q, "beta": 1.0, "seed": 42},
            "roofline": {"bound": "valu", "hbm_equivalent_GBps": round(gbs, 1),
                         "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)},
            **({} if args.no_cpu_baseline else {"cpu_baseline": cpu_potts3d()})}


def cluster(sc, args) -> dict:
    """Swendsen-Wang and Wolff updates of the 2-D Ising model at beta_c on a
    4096 x 4096 torus (bond activation, tile and face labelling, the cluster
    transform, per update over every site), and of the 3-D Heisenberg model
    (O(3) embedding clusters) on a 96^3 torus near T_c."""
    L, beta = 4096, 0.4406868
    lat = sc.Lattice.square(L, L, True)
    n = lat.n_sites
    out = {"metric": "site-updates/ns (2-D Ising L=4096 at beta_c, cluster updates)", "unit": "site-updates/ns",
           "n_gpus": 1, "higher_is_better": True, "dtype": "int8 spins, uint32 labels", "data": "synthetic",
           "config": {"workload": "cluster_ising2d", "L": L, "beta": beta, "seed": 42}}
    for name, alg, steps in (("swendsen_wang", sc.SWENDSEN_WANG, args.cluster_steps),
                             ("wolff", sc.WOLFF, args.cluster_steps)):
        st = sc.State(lat, sc.model(sc.ISING, 1.0), 1, 42)
        st.init(True)
        st.set_beta(beta)
        st.sweep(alg, 20)
        st.synchronize()
        ms = st.sweep_timed(alg, steps)
        e, m = st.observables()
        out[name] = {"updates": steps, "ms_per_update": round(ms / steps, 4),
                     "site_updates_per_ns": round(n * steps / (ms * 1e6), 3),
                     "energy_per_site": e / n, "abs_m": m}
        st.close()
    # O(3) embedding clusters: 3-D Heisenberg L=96 near its critical point
    # (T_c ~ 1.443), one replica
    L3, beta3 = 96, 1 / 1.443
    lat3 = sc.Lattice.hypercubic([L3, L3, L3], True)
    n3 = lat3.n_sites
    for name, alg in (("on3_swendsen_wang", sc.SWENDSEN_WANG), ("on3_wolff", sc.WOLFF)):
        st = sc.State(lat3, sc.model(sc.ON, 1.0, 2, 3), 1, 42)
        st.init(True)
        st.set_beta(beta3)
        st.sweep(alg, 10)
        st.synchronize()
        steps = max(1, args.cluster_steps // 4)
        ms = st.sweep_timed(alg, steps)
        e, m = st.observables()
        out[name] = {"L": L3, "beta": round(beta3, 6), "updates": steps, "ms_per_update": round(ms / steps, 4),
                     "site_updates_per_ns": round(n3 * steps / (ms * 1e6), 3), "energy_per_site": e / n3,
                     "abs_m": m}
        st.close()
    out["value"] = out["swendsen_wang"]["site_updates_per_ns"]
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_cluster()
    return out


def calibrate(sc, args) -> dict:
    """Kernels of known HBM traffic on a state larger than the Infinity Cache,
    for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE: a random start of a
    32768^2 lattice (k_init writes N = 1 GiB) and one exact recount of its
    observables (k_measure reads N bytes: each colour once, neighbours
    through the caches)."""
    L = 32768
    lat = sc.Lattice.square(L, L, True)
    st = sc.State(lat, sc.model(sc.ISING, 1.0), 1, 7)
    st.init(True)
    e, m = st.observables()
    return {"workload": "pmc_calibration", "L": L, "bytes_written_by_k_init": lat.n_sites,
            "bytes_read_by_k_measure": lat.n_sites, "energy": e, "abs_m": m}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--L", type=int, default=16384)
    ap.add_argument("--workload", choices=["ising2d", "pt3d", "on3", "xy", "glass", "tri", "potts3d", "cluster", "serve", "calibrate"],
                    default="ising2d")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-sublines", action="store_true", help="leave the pt3d and on3 lines out of the default line")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[10.0, 30.0, 100.0, 300.0, 1000.0])
    ap.add_argument("--plain-requests", type=int, default=48)
    ap.add_argument("--pt-steps", type=int, default=10_000)
    ap.add_argument("--on3-steps", type=int, default=200)
    ap.add_argument("--potts-steps", type=int, default=1000)
    ap.add_argument("--cluster-steps", type=int, default=200)
    ap.add_argument("--glass-steps", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: this build runs on a single GPU "
              "(multi-GPU sharding is designed but not launched).", file=sys.stderr)
        return 2

    import scamper_amd as sc

    if args.workload == "calibrate":
        line = calibrate(sc, args)
    elif args.workload == "pt3d":
        line = pt3d(sc, args)
    elif args.workload == "on3":
        line = on3(sc, args)
    elif args.workload == "xy":
        line = xy(sc, args)
    elif args.workload == "glass":
        line = glass(sc, args)
    elif args.workload == "tri":
        line = tri(sc, args)
    elif args.workload == "potts3d":
        line = potts3d(sc, args)
    elif args.workload == "cluster":
        line = cluster(sc, args)
    elif args.workload == "serve":
        line = {"metric": "batched-queue sweeps/s + p99 latency (Potts serving mix)", "n_gpus": 1,
                "serving": serving(sc, args)}
    else:
        line = ising2d(sc, args)
        if not args.no_sublines:
            line["pt3d"] = pt3d(sc, args)
            line["on3"] = on3(sc, args)
        if not args.no_serving:
            line["serving"] = serving(sc, args)
        if not args.no_cpu_baseline:
            line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
