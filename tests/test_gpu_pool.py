"""The job pool (continuous batching of independent runs) on the GPU.

A job's draws are keyed by its own (seed, sweep, colour, site), so its result
must not depend on what else shares the GPU, nor on whether the pool runs it
on bit planes (Ising and Potts q <= 4, the default) or on int8 spins: each
check below compares
measured series, estimates and final spins bit for bit -- against
run_one on the same spec (plain per-job launches), against the same job
alone in a pool, across graph and plain launches, grid sizes and chunk
lengths, and under admission pressure from a small spin budget.  The tiled
kernel itself is checked against the oracle in test_gpu_stencil.py; here the
final spins of a pool job are also checked against the oracle directly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spec(sc, dims, kind, q, beta, therm, meas, seed, binsize=10):
    return sc.job_spec(dims, sc.model(kind, 1.0, q), beta, therm, meas, binsize=binsize, seed=seed,
                       random_start=True, keep_series=True, keep_spins=True)


MIX = [
    # dims, kind, q, beta, therm, meas
    ((32, 32), 1, 3, 1.005052, 20, 70),
    ((16, 16, 16), 1, 4, 0.62863, 15, 41),
    ((64, 48), 1, 2, 0.881374, 7, 33),
    ((8, 8, 8), 0, 2, 0.2216544, 5, 60),
    ((130, 66), 1, 4, 1.098612, 3, 29),
    ((12,), 0, 2, 0.8, 0, 50),
    ((34, 20, 14), 1, 6, 0.9, 9, 16),
    ((256, 64), 1, 3, 1.0, 2, 12),
]


def _run_pool(sc, specs, submit_after=None, **cfg):
    """Run specs to completion; specs listed in submit_after (index -> steps)
    are submitted only after that many pool steps."""
    pool = sc.Pool(**cfg)
    ids = {}
    pending = dict(submit_after or {})
    for i, s in enumerate(specs):
        if i not in pending:
            ids[i] = pool.submit(s)
    steps = 0
    while pending:
        for i in [i for i, t in pending.items() if t <= steps]:
            ids[i] = pool.submit(specs[i])
            del pending[i]
        pool.step()
        steps += 1
    pool.run()
    out = {i: pool.result(jid) for i, jid in ids.items()}
    return out, pool


def _same(a, b):
    ea, sa = a[0], a[3]
    eb, sb = b[0], b[3]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), "series differ"
    assert np.array_equal(sa, sb), "final spins differ"
    for k in ea:
        assert ea[k].mean == eb[k].mean and ea[k].stderr == eb[k].stderr, k


@pytest.mark.parametrize("bit_planes", [True, False])
def test_pool_equals_run_one(sc, bit_planes):
    """Bit planes (Ising and Potts q <= 4 jobs; the q = 6 job stays on int8
    spins in the same tables) or int8 spins throughout: every job equals its
    run_one on the int8 kernel."""
    specs = [_spec(sc, d, k, q, b, t, m, 100 + i) for i, (d, k, q, b, t, m) in enumerate(MIX)]
    res, pool = _run_pool(sc, specs, chunk_sweeps=16, max_jobs=64, spin_budget_bytes=1 << 24, bit_planes=bit_planes)
    st = pool.stats()
    assert st.done == len(MIX) and st.queued == 0 and st.spin_bytes_in_use == 0
    for i, (d, k, q, b, t, m) in enumerate(MIX):
        print(f"job {i}: dims={d} kind={k} q={q} therm={t} meas={m}", flush=True)
        est, e, mm = sc.run_one(sc.Lattice.hypercubic(list(d), True), sc.model(k, 1.0, q), sc.METROPOLIS, b,
                                True, t, m, 10, 100 + i, series=True)
        got = res[i]
        assert np.array_equal(got[1], e) and np.array_equal(got[2], mm), f"job {i}: series differ from run_one"
        for name in est:
            assert got[0][name].mean == est[name].mean, (i, name)
            assert got[0][name].stderr == est[name].stderr, (i, name)
            assert got[0][name].n_bins == est[name].n_bins, (i, name)


def test_pool_final_spins_match_oracle(sc, orc):
    d, k, q, b, t, m = ((64, 34), 1, 4, 1.098612, 5, 20)
    seed = 77
    out, _ = _run_pool(sc, [_spec(sc, d, k, q, b, t, m, seed)], chunk_sweeps=8, max_jobs=4,
                       spin_budget_bytes=1 << 20)
    res = out[0]
    key = orc.replica_key(seed, 0)
    lat = orc.Lattice.hypercubic(list(d), True)
    mo = orc.model(k, 1.0, q, 1)
    start = orc.random_start(lat, mo, key).astype(np.int8)
    orc.stencil_sweeps(start, d, k, q, b, 1.0, False, key, 0, t + m)
    assert np.array_equal(res[3], start)
    chain = orc.Chain(lat, mo, start.astype(np.float64))
    assert res[1][-1] == chain.total_energy()
    assert res[2][-1] == chain.magnetization()


def test_large_mix_matches_small_launches_and_oracle(sc, orc):
    """With >= 2 Mi sites per colour in the table the pool switches to 16
    sites per lane; every job still equals its run_one (4 per lane) and the
    big job's final spins equal the oracle."""
    specs = [_spec(sc, (4100, 4100), 1, 3, 1.0, 1, 2, 61), _spec(sc, (34, 20, 14), 1, 6, 0.9, 2, 5, 62),
             _spec(sc, (130, 66), 1, 4, 1.098612, 1, 6, 63)]
    res, _ = _run_pool(sc, specs, chunk_sweeps=4, max_jobs=8, spin_budget_bytes=1 << 26)
    for i, (d, k, q, b, t, m) in enumerate([((34, 20, 14), 1, 6, 0.9, 2, 5), ((130, 66), 1, 4, 1.098612, 1, 6)]):
        est, e, mm = sc.run_one(sc.Lattice.hypercubic(list(d), True), sc.model(k, 1.0, q), sc.METROPOLIS, b,
                                True, t, m, 10, 62 + i, series=True)
        assert np.array_equal(res[i + 1][1], e) and np.array_equal(res[i + 1][2], mm)
    key = orc.replica_key(61, 0)
    lat = orc.Lattice.hypercubic([4100, 4100], True)
    mo = orc.model(1, 1.0, 3, 1)
    start = orc.random_start(lat, mo, key).astype(np.int8)
    orc.stencil_sweeps(start, (4100, 4100), 1, 3, 1.0, 1.0, False, key, 0, 3)
    assert np.array_equal(res[0][3], start)


def test_alone_equals_in_the_mix(sc):
    specs = [_spec(sc, d, k, q, b, t, m, 500 + i) for i, (d, k, q, b, t, m) in enumerate(MIX)]
    alone = {}
    for i in (0, 4, 6):
        r, _ = _run_pool(sc, [specs[i]], chunk_sweeps=32, max_jobs=8, spin_budget_bytes=1 << 24)
        alone[i] = r[0]
    # in the mix: late submission shifts the job's chunk boundaries
    mixed, _ = _run_pool(sc, specs, submit_after={0: 3, 4: 1, 6: 5}, chunk_sweeps=8, max_jobs=64,
                         spin_budget_bytes=1 << 24)
    for i in alone:
        print(f"job {i} alone vs mix", flush=True)
        _same(alone[i], mixed[i])


@pytest.mark.parametrize("cfg", [
    dict(use_graphs=False, grid=8, chunk_sweeps=5),
    dict(use_graphs=True, grid=4096, chunk_sweeps=64),
    dict(use_graphs=True, grid=0, chunk_sweeps=3),  # autotuned grid
])
def test_launch_shape_does_not_change_results(sc, cfg):
    specs = [_spec(sc, d, k, q, b, t, m, 900 + i) for i, (d, k, q, b, t, m) in enumerate(MIX)]
    ref, _ = _run_pool(sc, specs, chunk_sweeps=16, max_jobs=64, spin_budget_bytes=1 << 24, grid=1024)
    got, pool = _run_pool(sc, specs, max_jobs=64, spin_budget_bytes=1 << 24, **cfg)
    for i in ref:
        _same(ref[i], got[i])
    if cfg["grid"] == 0:
        print("autotuned grid:", pool.stats().grid, "tuned:", pool.stats().tuned)


def test_autotuner_settles_grid_lane_and_chunk(sc):
    """With grid = 0 the pool times candidates on real chunks of one resident
    set, one coordinate at a time: the lane-width crossover (1, 2 or 4 Mi
    colour-class sites for 16 sites per lane), then the grid, then the chunk
    length (chunk_sweeps / 4, / 2, chunk_sweeps).  It starts once the
    resident set holds 1 Mi colour-class sites (8 Mi for the grid), settles
    and records its choice, and every shape it tried gives the results of a
    fixed shape."""
    specs = [_spec(sc, (2048, 2048), 1, 3, 1.0, 10, 290 + 7 * i, 700 + i) for i in range(3)] + \
            [_spec(sc, (200, 200, 200), 0, 2, 0.22, 10, 300, 710)]
    ref, _ = _run_pool(sc, specs, chunk_sweeps=16, max_jobs=16, spin_budget_bytes=1 << 25, grid=1024)
    got, pool = _run_pool(sc, specs, chunk_sweeps=8, max_jobs=16, spin_budget_bytes=1 << 25, grid=0)
    st = pool.stats()
    print(f"settled: grid {st.grid} lane16 from {st.lane16_from_mi} Mi chunk {st.chunk_sweeps}", flush=True)
    assert st.tuned == 1 and st.grid in (1024, 2048, 4096)
    assert st.lane16_from_mi in (1, 2, 4) and st.chunk_sweeps in (2, 4, 8)
    for i in ref:
        _same(ref[i], got[i])


def test_autotuner_waits_for_enough_work(sc):
    """A resident set below 1 Mi colour-class sites runs the starting shape
    (4096 blocks capped at the tiles, 16 sites per lane from 2 Mi sites,
    chunk_sweeps) untuned."""
    specs = [_spec(sc, (64, 64), 1, 3, 1.0, 10, 90 + 7 * i, 720 + i) for i in range(3)]
    got, pool = _run_pool(sc, specs, chunk_sweeps=8, max_jobs=16, spin_budget_bytes=1 << 24, grid=0)
    st = pool.stats()
    assert st.tuned == 0
    ref, _ = _run_pool(sc, specs, chunk_sweeps=8, max_jobs=16, spin_budget_bytes=1 << 24, grid=1024)
    for i in ref:
        _same(ref[i], got[i])


def test_admission_by_budget(sc):
    # room for two 64x64 jobs at a time; never over budget.  A job holds its
    # two colour arrays of 2048 sites, each padded to a multiple of 256 bytes
    # with a gap of at least 16: 2 * 2304 bytes.
    job_bytes = 2 * 2304
    specs = [_spec(sc, (64, 64), 1, 3, 1.0, 5, 40 + 3 * i, 300 + i) for i in range(6)]
    pool = sc.Pool(spin_budget_bytes=2 * job_bytes, max_jobs=16, chunk_sweeps=8)
    ids = [pool.submit(s) for s in specs]
    peak_running = 0
    while True:
        st = pool.stats()
        assert st.spin_bytes_in_use <= 2 * job_bytes
        peak_running = max(peak_running, st.running)
        if st.done == len(specs):
            break
        pool.step()
    assert peak_running >= 2
    with pytest.raises(sc.ScamperError) as err:
        pool.submit(_spec(sc, (128, 128), 1, 3, 1.0, 1, 1, 1))
    assert err.value.code == 6  # SCAMPER_EBUDGET
    for i, s in enumerate(specs):
        r, _ = _run_pool(sc, [s], chunk_sweeps=32, max_jobs=4, spin_budget_bytes=1 << 20)
        _same(pool.result(ids[i]), r[0])


def test_row_limit_queues_jobs(sc):
    specs = [_spec(sc, (16, 16), 0, 2, 0.44, 3, 20 + i, 40 + i) for i in range(7)]
    pool = sc.Pool(spin_budget_bytes=1 << 20, max_jobs=2, chunk_sweeps=4)
    for s in specs:
        pool.submit(s)
    while pool.stats().done < len(specs):
        assert pool.stats().running <= 2
        pool.step()
    assert len(pool.completed()) == len(specs)
