"""Serving loop over the job pool: Potts simulation requests arrive as a
Poisson stream, are admitted by the pool as spin memory and job rows free
up, advance together on the GPU, and leave with their estimates.  Reports
completion latency (p50 / p99) next to throughput, and times the same
requests as plain per-job launches (scamper_run_one, one job after another)
on the same GPU.

Request mix: q uniform in {2, 3, 4}; a 2-D square lattice with L
log-uniform in [32, 2048] or a 3-D cubic lattice with L log-uniform in
[16, 256], each with probability 1/2 (L rounded to the nearest even value,
periodic); T uniform in [0.95, 1.05] T_c(q, d); sweep budget log-uniform in
[10^3, 10^5], of which one tenth is thermalisation; binsize 100; random
start; seed = a hash of (mix seed, request index).

T_c with J = 1 and energy -J delta(s_i, s_j):
  2-D: 1 / ln(1 + sqrt(q)) (Baxter; the reference's tests use the same form,
       CMC.rs/tests/physics/potts_exact.rs:482,559);
  3-D: literature inverse temperatures beta_c = 0.4433093 (q = 2: twice the
       simple-cubic Ising value 0.22165463), 0.550565 (q = 3, Janke and
       Villanova 1997) and 0.62863 (q = 4, Bazavov, Berg and Dubey 2008).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import METROPOLIS, POTTS, Lattice, Pool, job_spec, model, run_one

BETA_C_3D = {2: 0.4433093, 3: 0.550565, 4: 0.62863}


def critical_temperature(q: int, n_dims: int) -> float:
    if n_dims == 2:
        return 1.0 / math.log(1.0 + math.sqrt(q))
    return 1.0 / BETA_C_3D[q]


@dataclass
class Request:
    index: int
    arrival: float  # seconds after the start of the stream
    dims: tuple
    q: int
    beta: float
    therm: int
    meas: int
    seed: int

    @property
    def sites(self) -> int:
        return int(np.prod(self.dims))

    @property
    def updates(self) -> int:
        return self.sites * (self.therm + self.meas)


def _log_uniform(rng, lo, hi):
    return math.exp(rng.uniform(math.log(lo), math.log(hi)))


def _even(x: float, lo: int, hi: int) -> int:
    return int(min(hi, max(lo, 2 * round(x / 2))))


def potts_mix(n: int, rate: float, seed: int = 2024, l2=(32, 2048), l3=(16, 256),
              sweeps=(1_000, 100_000)) -> list[Request]:
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        q = int(rng.integers(2, 5))
        nd = 2 if rng.random() < 0.5 else 3
        lo, hi = l2 if nd == 2 else l3
        L = _even(_log_uniform(rng, lo, hi), lo, hi)
        t = rng.uniform(0.95, 1.05) * critical_temperature(q, nd)
        budget = int(round(_log_uniform(rng, *sweeps)))
        therm = budget // 10
        out.append(Request(i, float(arrivals[i]), (L,) * nd, q, 1.0 / t, therm, budget - therm,
                           int((seed * 1_000_003 + i) % (1 << 63))))
    return out


def _spec(r: Request, keep_series: bool = False):
    return job_spec(r.dims, model(POTTS, 1.0, r.q), r.beta, r.therm, r.meas, binsize=100, seed=r.seed,
                    random_start=True, keep_series=keep_series)


@dataclass
class ServeResult:
    n: int
    wall_s: float
    updates: int
    latencies: np.ndarray
    estimates: dict = field(default_factory=dict)
    shape: dict = field(default_factory=dict)  # launch shape the pool's autotuner settled on

    @property
    def flips_per_ns(self) -> float:
        return self.updates / (self.wall_s * 1e9)

    def summary(self, requests) -> dict:
        sweeps = sum(r.therm + r.meas for r in requests)
        return {"requests": self.n, "wall_s": round(self.wall_s, 3),
                "lattice_sweeps_per_s": round(sweeps / self.wall_s, 1),
                "spin_flips_per_ns": round(self.flips_per_ns, 3),
                "p50_latency_s": round(float(np.percentile(self.latencies, 50)), 4),
                "p99_latency_s": round(float(np.percentile(self.latencies, 99)), 4),
                "mean_latency_s": round(float(self.latencies.mean()), 4)}


def serve(requests: list[Request], spin_budget_bytes: int = 8 << 30, max_jobs: int = 512,
          chunk_sweeps: int = 64, grid: int = 0, keep: int = 0, bit_planes: bool = True) -> ServeResult:
    """Play the request stream in real time against one pool.  `keep` =
    number of leading requests whose estimates are returned.  With grid = 0
    the pool's tuner picks the chunk length among chunk_sweeps / 4, / 2 and
    chunk_sweeps (64: a chunk's graph of 130 launches amortises the chunk's
    table upload, plane packing and host work; 32 and 128 measured ~3% and
    ~1.5% slower on a 48-request Potts mix)."""
    pool = Pool(spin_budget_bytes=spin_budget_bytes, max_jobs=max_jobs, chunk_sweeps=chunk_sweeps, grid=grid,
                bit_planes=bit_planes)
    n = len(requests)
    id_to_req = {}
    done_at = np.zeros(n)
    estimates = {}
    t0 = time.perf_counter()
    i = done = 0
    while done < n:
        now = time.perf_counter() - t0
        while i < n and requests[i].arrival <= now:
            id_to_req[pool.submit(_spec(requests[i]))] = i
            i += 1
        if len(id_to_req) > done:
            pool.step()
            t = time.perf_counter() - t0
            for jid in pool.completed():
                k = id_to_req[jid]
                done_at[k] = t
                if k < keep:
                    estimates[k] = pool.result(jid)[0]
                pool.release(jid)
                done += 1
        elif i < n:
            time.sleep(max(0.0, min(requests[i].arrival - now, 0.002)))
    wall = time.perf_counter() - t0
    lat = done_at - np.array([r.arrival for r in requests])
    st = pool.stats()
    shape = {"grid": st.grid, "lane16_from_mi_sites": st.lane16_from_mi, "chunk_sweeps": st.chunk_sweeps} \
        if st.tuned else {"tuned": False}
    return ServeResult(n, wall, sum(r.updates for r in requests), lat, estimates, shape)


def plain(requests: list[Request], keep: int = 0) -> ServeResult:
    """The same requests as plain per-job launches: scamper_run_one, one job
    at a time in arrival order (a FIFO single server; latencies follow from
    the measured run times and the arrival times)."""
    run_s = np.zeros(len(requests))
    estimates = {}
    for k, r in enumerate(requests):
        lat = Lattice.hypercubic(list(r.dims), True)
        t = time.perf_counter()
        est = run_one(lat, model(POTTS, 1.0, r.q), METROPOLIS, r.beta, True, r.therm, r.meas, 100, r.seed)
        run_s[k] = time.perf_counter() - t
        if k < keep:
            estimates[k] = est
    finish = np.zeros(len(requests))
    clock = 0.0
    for k, r in enumerate(requests):
        clock = max(clock, r.arrival) + run_s[k]
        finish[k] = clock
    lat = finish - np.array([r.arrival for r in requests])
    return ServeResult(len(requests), float(run_s.sum()), sum(r.updates for r in requests), lat, estimates)


def plain_pool(requests: list[Request], keep: int = 0, chunk_sweeps: int = 32, bit_planes: bool = True) -> ServeResult:
    """The same requests as plain per-job launches of the pool's kernels: one
    job at a time, alone in a pool of one row, with every sweep's launches
    issued from the host (no graphs), in arrival order."""
    run_s = np.zeros(len(requests))
    estimates = {}
    for k, r in enumerate(requests):
        pool = Pool(spin_budget_bytes=max(1 << 20, 4 * r.sites), max_jobs=1, chunk_sweeps=chunk_sweeps, grid=1024,
                    use_graphs=False, bit_planes=bit_planes)
        t = time.perf_counter()
        jid = pool.submit(_spec(r))
        pool.run()
        run_s[k] = time.perf_counter() - t
        if k < keep:
            estimates[k] = pool.result(jid)[0]
        del pool
    finish = np.zeros(len(requests))
    clock = 0.0
    for k, r in enumerate(requests):
        clock = max(clock, r.arrival) + run_s[k]
        finish[k] = clock
    lat = finish - np.array([r.arrival for r in requests])
    return ServeResult(len(requests), float(run_s.sum()), sum(r.updates for r in requests), lat, estimates)
