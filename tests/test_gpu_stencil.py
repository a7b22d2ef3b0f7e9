"""GPU parity of the checkerboard kernels against the oracle, bit for bit.

Each case starts from the same spins, runs k sweeps on the device through
the C ABI and the same k sweeps in the oracle with the same Philox draws,
and compares every spin, the acceptance-driven energy and magnetisation, and
the fused per-sweep observables.  Sizes cover the per-site kernel (sides
that are not multiples of 32), the register-blocked Ising kernel, odd row
counts for its row blocking, 3-D, Potts and heat bath.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lattice(sc, dims):
    return sc.Lattice.hypercubic(list(dims), True)


def _oracle_state(orc, dims, kind, q, key):
    lat = orc.Lattice.hypercubic(list(dims), True)
    m = orc.model(kind, 1.0, q, 1)
    return lat, m, orc.random_start(lat, m, key)


def _stencil(orc, spins, dims, kind, q, beta, heat_bath, key, sweep0, n):
    s = spins.astype(np.int8).copy()
    orc.stencil_sweeps(s, dims, kind, q, beta, 1.0, heat_bath, key, sweep0, n)
    return s


CASES = [
    # dims, kind, q, algorithm, beta, sweeps
    ((8, 8), 0, 2, 0, 0.4406868, 6),
    ((6, 10), 0, 2, 0, 0.3, 5),
    ((32, 32), 0, 2, 0, 0.4406868, 8),
    ((64, 64), 0, 2, 0, 0.4406868, 8),
    ((64, 18), 0, 2, 0, 0.6, 8),
    ((256, 256), 0, 2, 0, 0.4406868, 5),
    ((8, 8), 0, 2, 1, 0.4406868, 6),
    ((64, 64), 0, 2, 1, 0.4406868, 4),
    ((32, 32, 32), 0, 2, 0, 0.2216544, 4),
    ((8, 6, 4), 0, 2, 0, 0.25, 4),
    ((16, 16, 16), 0, 2, 1, 0.2216544, 3),
    ((16, 16), 1, 3, 0, 1.005052, 5),
    ((32, 32), 1, 4, 0, 1.098612, 5),
    ((16, 16), 1, 3, 1, 1.005052, 5),
    ((8, 8, 8), 1, 4, 1, 0.6, 3),
    ((8, 8, 8), 1, 2, 0, 0.4, 3),
    ((16, 16), 1, 16, 0, 1.5, 3),
    ((16, 16), 1, 7, 1, 1.3, 3),
    ((12,), 0, 2, 0, 0.8, 5),
    # tiled kernel: several tiles, a partial last tile, rows that straddle
    # lanes, q with and without proposal straddles, enough draws for ties
    ((256, 256), 1, 4, 0, 1.098612, 10),
    ((130, 66), 1, 5, 0, 1.2, 6),
    ((34, 20, 14), 1, 6, 0, 0.9, 6),
    ((2, 2, 2), 1, 3, 0, 0.5, 4),
    ((258, 34), 0, 2, 0, 0.4406868, 6),
    ((64, 8, 6), 1, 3, 0, 0.55, 5),
    ((4098,), 0, 2, 0, 0.7, 5),
    ((96, 40), 1, 5, 0, 1.2, 5),
    ((40, 40, 40), 1, 4, 0, 0.62863, 4),
    # lanes that cross from row (ly-1, z) into row (0, z+1), where both rows
    # have the same parity (Lx/2 * Ly not a multiple of 4)
    ((10, 6, 4), 1, 3, 0, 0.55, 5),
    ((18, 10, 6), 0, 2, 0, 0.25, 4),
    ((26, 14, 6), 1, 4, 0, 0.62863, 4),
    # 16 sites per lane (launches over >= 2 Mi sites per colour)
    ((4100, 4100), 1, 3, 0, 1.0, 3),
    ((204, 206, 400), 1, 4, 0, 0.62863, 3),
    ((4098, 4100), 0, 2, 0, 0.4406868, 3),
]


@pytest.mark.parametrize("dims,kind,q,alg,beta,n", CASES)
def test_bit_exact_against_oracle(sc, orc, dims, kind, q, alg, beta, n):
    seed = 1234 + len(dims) * 7 + q
    print(f"case dims={dims} kind={kind} q={q} alg={alg} beta={beta}", flush=True)
    lat = _lattice(sc, dims)
    st = sc.State(lat, sc.model(kind, 1.0, q), 1, seed)
    st.init(True)
    key = orc.replica_key(seed, 0)
    olat, om, ospins = _oracle_state(orc, dims, kind, q, key)
    start = st.download_i8()
    assert np.array_equal(start, ospins.astype(np.int8)), "random start differs"
    st.set_beta(beta)
    st.sweep(alg, n, measure=True)
    got = st.download_i8()
    want = _stencil(orc, start, dims, kind, q, beta, alg == 1, key, 0, n)
    assert np.array_equal(got, want), f"{(got != want).sum()} spins differ"
    # fused observables of the last sweep equal the exact recomputation
    e_series, m_series = st.series()
    chain = orc.Chain(olat, om, want.astype(np.float64))
    assert e_series.size == n
    assert e_series[-1] == chain.total_energy()
    assert m_series[-1] == chain.magnetization()
    e, m = st.observables()
    assert e == chain.total_energy() and m == chain.magnetization()


@pytest.mark.parametrize("dims", [(8, 8), (16, 16), (4, 6, 4)])
def test_matches_csr_restatement(sc, orc, dims):
    """The stencil oracle equals the CSR restatement (faithful f64 rule); the
    device equals both."""
    seed = 99
    key = orc.replica_key(seed, 0)
    olat, om, spins = _oracle_state(orc, dims, 0, 2, key)
    chain = orc.Chain(olat, om, spins)
    for t in range(4):
        chain.metropolis(0.4406868, key, t)
    st = sc.State(_lattice(sc, dims), sc.model(0), 1, seed)
    st.init(True)
    st.set_beta(0.4406868)
    st.sweep(0, 4)
    assert np.array_equal(st.download_f64(), chain.spins)
    # cached energy (energy += ΔE per accepted move) equals the recomputation
    assert chain.energy.value == chain.total_energy()


def test_replicas_are_independent_streams(sc, orc):
    dims, seed = (32, 32), 7
    betas = [0.2, 0.4406868, 0.9]
    st = sc.State(_lattice(sc, dims), sc.model(0), len(betas), seed)
    starts = []
    for r, b in enumerate(betas):
        st.init(True, r)
        st.set_beta(b, r)
        starts.append(st.download_i8(r))
    st.sweep(0, 6, measure=True)
    for r, b in enumerate(betas):
        key = orc.replica_key(seed, r)
        want = _stencil(orc, starts[r], dims, 0, 2, b, False, key, 0, 6)
        assert np.array_equal(st.download_i8(r), want)


@pytest.mark.parametrize("dims,kind,q", [((34, 22), 1, 4), ((66, 10, 6), 1, 3), ((36, 14), 0, 2)])
def test_replicas_with_unaligned_colour_arrays(sc, orc, dims, kind, q):
    """Several replicas whose colour arrays start at addresses that are not
    16-byte aligned (N/2 not a multiple of 16), at different temperatures."""
    seed = 21
    betas = [0.3, 0.9, 1.4]
    st = sc.State(_lattice(sc, dims), sc.model(kind, 1.0, q), len(betas), seed)
    starts = []
    for r, b in enumerate(betas):
        st.init(True, r)
        st.set_beta(b, r)
        starts.append(st.download_i8(r))
    st.sweep(0, 5, measure=True)
    for r, b in enumerate(betas):
        key = orc.replica_key(seed, r)
        want = _stencil(orc, starts[r], dims, kind, q, b, False, key, 0, 5)
        got = st.download_i8(r)
        assert np.array_equal(got, want), f"replica {r}: {(got != want).sum()} spins differ"


def test_upload_roundtrip_and_validation(sc):
    lat = _lattice(sc, (16, 8))
    st = sc.State(lat, sc.model(0), 1, 3)
    rng = np.random.default_rng(0)
    s = rng.choice([-1, 1], size=lat.n_sites).astype(np.int8)
    st.upload_i8(s)
    assert np.array_equal(st.download_i8(), s)
    st.upload_f64(s.astype(np.float64) * -1)
    assert np.array_equal(st.download_f64(), -s.astype(np.float64))
    bad = s.astype(np.float64)
    bad[3] = 0.5
    with pytest.raises(sc.ScamperError):
        st.upload_f64(bad)


def test_cold_start_observables(sc):
    for dims in [(16, 16), (8, 8, 8)]:
        lat = _lattice(sc, dims)
        st = sc.State(lat, sc.model(0), 1, 0)
        e, m = st.observables()
        assert e == -float(lat.n_edges) and m == 1.0

