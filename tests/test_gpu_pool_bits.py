"""Edge cases of the pool's bit-plane kernel (Ising and Potts q <= 4 jobs).

A job's words hold 32 consecutive ranks of one colour class, so rows
shorter than a word, rows of one site per colour (L = 2), words that cross
plane and lattice ends, and the last partial word all take the periodic-wrap
masks; antiferromagnetic couplings, beta = 0 and a large beta move the set
of indices with K = 2^32 and the ties.  Each job's measured series, estimates
and final spins must equal run_one on the int8 kernel bit for bit (the int8
kernel is checked against the oracle in test_gpu_stencil.py), in one mixed
pool and with every job alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EDGE = [
    # dims, kind (0 Ising, 1 Potts), q, J, beta, random start
    ((2,), 0, 2, 1.0, 0.7, True),
    ((6,), 1, 3, 1.0, 0.9, True),
    ((2, 2), 1, 4, 1.0, 1.1, True),
    ((4, 2), 0, 2, -1.0, 0.5, True),
    ((10, 4), 1, 2, 1.0, 0.88, False),
    ((64, 2), 1, 3, -1.0, 0.6, True),
    ((2, 2, 2), 1, 3, 1.0, 0.7, True),
    ((6, 6, 6), 0, 2, 1.0, 0.2216544, True),
    ((8, 4, 6), 1, 4, 1.0, 0.0, True),
    ((34, 10), 1, 4, 1.0, 9.0, True),
    ((96, 34), 0, 2, -1.0, 0.44, True),
    ((66, 64, 2), 1, 3, 1.0, 0.55, False),
    ((70, 6, 4), 1, 4, -1.0, 0.8, True),
]


def _spec(sc, dims, kind, q, j, beta, rs, seed, therm=4, meas=23):
    return sc.job_spec(dims, sc.model(kind, j, q), beta, therm, meas, binsize=5, seed=seed, random_start=rs,
                       keep_series=True, keep_spins=True)


def _run(sc, specs, **cfg):
    pool = sc.Pool(**cfg)
    ids = [pool.submit(s) for s in specs]
    pool.run()
    return [pool.result(i) for i in ids]


def _run_one(sc, dims, kind, q, j, beta, rs, seed, therm=4, meas=23):
    return sc.run_one(sc.Lattice.hypercubic(list(dims), True), sc.model(kind, j, q), sc.METROPOLIS, beta, rs,
                      therm, meas, 5, seed, series=True)


def test_edge_jobs_in_one_pool_equal_run_one(sc):
    specs = [_spec(sc, *case, 300 + i) for i, case in enumerate(EDGE)]
    res = _run(sc, specs, spin_budget_bytes=1 << 22, max_jobs=32, chunk_sweeps=8, grid=1024)
    for i, case in enumerate(EDGE):
        print(f"case {i}: {case}", flush=True)
        est, e, m = _run_one(sc, *case, 300 + i)
        assert np.array_equal(res[i][1], e) and np.array_equal(res[i][2], m), f"case {i}: series differ"
        for name in est:
            assert res[i][0][name].mean == est[name].mean and res[i][0][name].stderr == est[name].stderr, (i, name)


@pytest.mark.parametrize("i", [0, 3, 6, 9, 12])
def test_edge_job_alone_equals_mix_and_int8_pool(sc, i):
    specs = [_spec(sc, *case, 300 + k) for k, case in enumerate(EDGE)]
    mixed = _run(sc, specs, spin_budget_bytes=1 << 22, max_jobs=32, chunk_sweeps=5, grid=0)[i]
    alone = _run(sc, [specs[i]], spin_budget_bytes=1 << 20, max_jobs=2, chunk_sweeps=16, grid=8)[0]
    int8 = _run(sc, [specs[i]], spin_budget_bytes=1 << 20, max_jobs=2, chunk_sweeps=16, grid=8, bit_planes=False)[0]
    for other in (alone, int8):
        assert np.array_equal(mixed[1], other[1]) and np.array_equal(mixed[2], other[2])
        assert np.array_equal(mixed[3], other[3]), "final spins differ"
