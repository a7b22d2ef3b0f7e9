"""GPU parity of the multi-spin kernel (one temperature per bit) and of the
device-side replica exchange, against the oracle and against the
single-replica kernels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bits(words, r):
    return (((words >> np.uint64(r)) & np.uint64(1)).astype(np.int8) * 2 - 1)


# (32, 32), (32, 6, 4) and (64,): rows of a multiple of 16 sites per colour,
# where a group of 16 ranks lies in one row; the others cross rows, and
# (48, 6) ends its colour class inside a group
@pytest.mark.parametrize("dims", [(8, 8, 8), (32, 32), (16, 6, 4), (32, 6, 4), (64,), (48, 6)])
def test_bits_equal_single_replica_runs(sc, dims):
    seed, n = 5, 4
    lat = sc.Lattice.hypercubic(list(dims))
    betas = np.linspace(0.15, 0.9, 64)
    ms = sc.MultiSpin(lat, 1.0, 1, seed)
    ms.set_betas(betas)
    ms.init(True)
    st = sc.State(lat, sc.model(sc.ISING), 64, seed)
    for r in range(64):
        st.init(True, r)
        st.set_beta(float(betas[r]), r)
    w0 = ms.download()
    for r in range(64):
        assert np.array_equal(_bits(w0, r), st.download_i8(r)), f"start of replica {r}"
    ms.step(n)
    st.sweep(sc.METROPOLIS, n)
    w = ms.download()
    for r in range(64):
        assert np.array_equal(_bits(w, r), st.download_i8(r)), f"replica {r}"


def test_two_bundles_match_oracle(sc, orc):
    dims, seed, n = (4, 4, 4), 9, 3
    lat = sc.Lattice.hypercubic(list(dims))
    olat = orc.Lattice.hypercubic(list(dims))
    betas = np.linspace(0.1, 0.5, 128)
    ms = sc.MultiSpin(lat, 1.0, 2, seed)
    ms.set_betas(betas)
    ms.init(True)
    words = ms.download().reshape(2, -1).copy()
    ms.step(n)
    got = ms.download().reshape(2, -1)
    for b in range(2):
        w = words[b].copy()
        for t in range(n):
            orc.multispin_sweep(olat, w, betas[64 * b:64 * b + 64], 1.0, seed, 64 * b, t)
        assert np.array_equal(got[b], w)


@pytest.mark.parametrize("dims", [(4, 4, 4), (32, 4, 4), (64,)])
def test_exchange_trajectory_matches_oracle(sc, orc, dims):
    """Sweeps with per-label betas, then exchange rounds; the label
    permutation, energies and magnetisations follow the oracle exactly."""
    seed, steps = 21, 40
    lat = sc.Lattice.hypercubic(list(dims))
    olat = orc.Lattice.hypercubic(list(dims))
    ladder = 4.30 * (4.72 / 4.30) ** (np.arange(64) / 63.0)
    betas = 1.0 / ladder
    ms = sc.MultiSpin(lat, 1.0, 1, seed)
    ms.set_betas(betas)
    ms.init(True)
    words = ms.download().copy()
    labels = np.arange(64, dtype=np.int32)
    ms.step(steps, exchange=True, measure=True)
    e_dev, m_dev = ms.series()
    for t in range(steps):
        orc.multispin_sweep(olat, words, betas[labels], 1.0, seed, 0, t)
        energies = np.array([orc.multispin_energy(olat, words, r) for r in range(64)])
        assert np.array_equal(e_dev[t], energies), f"energies at step {t}"
        orc.pt_exchange(labels, energies, betas, seed, t + 1)
    assert np.array_equal(ms.download(), words)
    assert np.array_equal(ms.labels(), labels)
    accepted, near, rounds = ms.stats()
    assert rounds == steps and near == 0 and accepted > 0
    plus = np.array([[((words >> np.uint64(r)) & np.uint64(1)).sum() for r in range(64)]])
    m_last = np.abs(2.0 * plus[0] - words.size) / words.size
    assert np.array_equal(m_dev[-1], m_last)


def test_captured_steps_equal_single_steps(sc):
    """Chunks of 16 steps replay one captured graph whose kernels read the
    step counters from the device; 40 steps in one call (two chunks and 8
    single steps) equal 40 calls of one step each: spins, labels, series and
    exchange counters."""
    lat = sc.Lattice.hypercubic([16, 16, 16])
    betas = np.linspace(0.18, 0.26, 64)
    a, b = sc.MultiSpin(lat, 1.0, 1, 3), sc.MultiSpin(lat, 1.0, 1, 3)
    for ms in (a, b):
        ms.set_betas(betas)
        ms.init(True)
    a.step(40, exchange=True, measure=True)
    for _ in range(40):
        b.step(1, exchange=True, measure=True)
    assert np.array_equal(a.download(), b.download())
    assert np.array_equal(a.labels(), b.labels())
    ea, ma = a.series()
    eb, mb = b.series()
    assert ea.shape == eb.shape and np.array_equal(ea, eb) and np.array_equal(ma, mb)
    assert a.stats() == b.stats()


def test_bench_shape_equals_single_replica_runs(sc):
    """The tempering workload's own shape: 3-D L=128, 64 temperatures
    4.30 .. 4.72 (geometric) in one bundle.  A row holds 64 ranks, so the
    sweeps run the one-row kernels on their full grid (4,096 waves per
    launch, the tie queue and the counter copies at scale).  Two measured
    steps without exchange: every bit equals the single-replica kernels'
    run of that replica, and each step's energy and |m| of every replica
    equal that run's measured series."""
    L, seed, n = 128, 42, 2
    print(f"case multispin {L}^3 x 64 bits", flush=True)
    lat = sc.Lattice.hypercubic([L, L, L])
    betas = 1.0 / (4.30 * (4.72 / 4.30) ** (np.arange(64) / 63.0))
    ms = sc.MultiSpin(lat, 1.0, 1, seed)
    ms.set_betas(betas)
    ms.init(True)
    ms.step(n, exchange=False, measure=True)
    e_ms, m_ms = ms.series()
    st = sc.State(lat, sc.model(sc.ISING), 64, seed)
    for r in range(64):
        st.init(True, r)
        st.set_beta(float(betas[r]), r)
    st.sweep(sc.METROPOLIS, n, measure=True)
    w = ms.download()
    for r in range(64):
        assert np.array_equal(_bits(w, r), st.download_i8(r)), f"replica {r}"
        e, m = st.series(r)
        assert np.array_equal(e_ms[:, r], e) and np.array_equal(m_ms[:, r], m), f"series of replica {r}"
    st.close()
