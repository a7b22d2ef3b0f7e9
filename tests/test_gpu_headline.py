"""The headline workload at its own size, bit for bit against the oracle.

2-D Ising, 16384 x 16384 periodic, J = 1, int8 spins from the seeded random
start (seed 42), checkerboard Metropolis at T = 2.269185: the register-blocked
kernel's production launch (2048 blocks of 256 threads, XCD-remapped) is
compared spin for spin with the oracle's threaded restatement of
MetropolisCore (CMC.rs/src/algorithms/metropolis.rs:67-106) with
accept_log_probability (Carlo.rs/src/acceptance.rs:11-22) on the same
Philox draws.  Mid-size cases run the same kernel with many replicas and a
grid whose block count is a multiple of 8 (the XCD-remapped order), in 2-D
and 3-D.

The fused energy and magnetisation of the measured sweep are checked
against a numpy recount of the oracle's spins, which is independent of both
the kernel and the oracle's update code.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85


def _philox_np(c0, c1, c2, c3, key):
    """Philox4x32-10 over arrays of counters (textbook round function)."""
    mask = np.uint64(0xFFFFFFFF)
    x = [np.asarray(v, np.uint64) & mask for v in (c0, c1, c2, c3)]
    k0, k1 = key & 0xFFFFFFFF, key >> 32
    for r in range(10):
        if r:
            k0, k1 = (k0 + W0) & 0xFFFFFFFF, (k1 + W1) & 0xFFFFFFFF
        p0 = np.uint64(M0) * x[0]
        p1 = np.uint64(M1) * x[2]
        x = [(p1 >> np.uint64(32)) ^ x[1] ^ np.uint64(k0), p1 & mask,
             (p0 >> np.uint64(32)) ^ x[3] ^ np.uint64(k1), p0 & mask]
    return x


def _init_spins_at(key, sites):
    """Random start of the Ising model at the given site indices: +1 iff bit 0
    of byte (site & 15) of the init-stream block site >> 4 (stream 6, sweep 0,
    colour 0) is set."""
    sites = np.asarray(sites, np.uint64)
    g = sites >> np.uint64(4)
    blk = _philox_np(g, 0, np.uint64(6 << 24), g >> np.uint64(32), key)
    byte = sites & np.uint64(15)
    word = np.choose((byte >> np.uint64(2)).astype(np.int64), blk)
    b = (word >> ((byte & np.uint64(3)) * np.uint64(8))) & np.uint64(0xFF)
    return np.where(b & np.uint64(1), 1, -1).astype(np.int8)


def _recount(s2d):
    """Bond sum (J = 1, periodic square lattice: E = -sum s_i s_j) and |m| of a
    2-D int8 configuration, in int64."""
    a = s2d.astype(np.int8)
    bonds = int(np.sum(a * np.roll(a, -1, axis=1), dtype=np.int64)) + \
        int(np.sum(a * np.roll(a, -1, axis=0), dtype=np.int64))
    m = abs(int(np.sum(a, dtype=np.int64))) / a.size
    return -float(bonds), m


def _recount3(s3d):
    a = s3d.astype(np.int8)
    bonds = sum(int(np.sum(a * np.roll(a, -1, axis=ax), dtype=np.int64)) for ax in range(3))
    return -float(bonds), abs(int(np.sum(a, dtype=np.int64))) / a.size


def test_ising2d_l16384_bit_exact(sc, orc):
    L, seed, beta = 16384, 42, 1.0 / 2.269185
    print(f"case ising2d L={L} seed={seed}: 2 sweeps, the second measured", flush=True)
    lat = sc.Lattice.square(L, L, True)
    st = sc.State(lat, sc.model(sc.ISING, 1.0), 1, seed)
    st.init(True)
    start = st.download_i8()
    key = orc.replica_key(seed, 0)
    rng = np.random.default_rng(3)
    sample = np.concatenate([np.arange(4096), rng.integers(0, L * L, 1 << 20), np.arange(L * L - 4096, L * L)])
    assert np.array_equal(start[sample], _init_spins_at(key, sample)), "random start differs from its definition"
    st.set_beta(beta)
    st.sweep(sc.METROPOLIS, 1)
    st.sweep(sc.METROPOLIS, 1, measure=True)
    got = st.download_i8()
    want = start.copy()
    orc.stencil_sweeps(want, [L, L], orc.ISING, 2, beta, 1.0, False, key, 0, 2)
    diff = int(np.count_nonzero(got != want))
    assert diff == 0, f"{diff} spins differ from the oracle"
    e_series, m_series = st.series()
    e_want, m_want = _recount(want.reshape(L, L))
    assert e_series.size == 1
    assert e_series[-1] == e_want and m_series[-1] == m_want
    e, m = st.observables()
    assert e == e_want and m == m_want
    st.close()


@pytest.mark.parametrize("dims,n_rep,sweeps", [((2048, 2048), 16, 3), ((128, 128, 128), 4, 3),
                                               ((2048, 32, 8), 3, 3), ((1024, 64), 2, 4)])
def test_remapped_grid_many_replicas(sc, orc, dims, n_rep, sweeps):
    """Replicas at temperatures from the disordered phase to deep in the
    ordered one.  2048^2 (32 blocks per replica) and 128^3 (16) use the
    XCD-remapped block order.  Rows of Lx = 2048 sites are 1024 colour bytes,
    64 lanes of 16 bytes, so the horizontal edge comes from the neighbouring
    lane (2048^2 and 2048 x 32 x 8); Lx = 1024 rows are half a wave wide and
    load it."""
    seed = 77
    print(f"case dims={dims} replicas={n_rep} sweeps={sweeps}", flush=True)
    lat = sc.Lattice.hypercubic(list(dims), True)
    st = sc.State(lat, sc.model(sc.ISING, 1.0), n_rep, seed)
    betas = np.linspace(0.15, 1.2, n_rep)
    starts = []
    for r in range(n_rep):
        st.init(True, r)
        st.set_beta(float(betas[r]), r)
        starts.append(st.download_i8(r))
    st.sweep(sc.METROPOLIS, sweeps, measure=True)
    for r in range(n_rep):
        key = orc.replica_key(seed, r)
        want = starts[r].copy()
        orc.stencil_sweeps(want, list(dims), orc.ISING, 2, float(betas[r]), 1.0, False, key, 0, sweeps)
        got = st.download_i8(r)
        diff = int(np.count_nonzero(got != want))
        assert diff == 0, f"replica {r} (beta {betas[r]:.3f}): {diff} spins differ"
        shaped = want.reshape(tuple(reversed(dims)))
        e_want, m_want = _recount(shaped) if len(dims) == 2 else _recount3(shaped)
        e_r, m_r = st.series(r)
        assert e_r[-1] == e_want and m_r[-1] == m_want, f"replica {r}: fused observables"
    st.close()
