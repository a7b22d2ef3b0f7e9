"""Cluster updates on the GPU (SCAMPER_SWENDSEN_WANG, SCAMPER_WOLFF).

Parity: from the same spins and the same counter-based draws, k device
updates leave exactly the spins of k oracle updates (orc_cluster_steps), on
stencil lattices (1-D, 2-D, 3-D, colour-split int8; sides that are
multiples of the tile labelling's 32 x 32 and 16 x 8 x 8 tiles, so that tile
faces and periodic wraps inside one tile are crossed) and graph lattices (open
boundaries, triangular, a weighted graph with a self-loop and a parallel
bond), for Ising and Potts, with several replicas at different
temperatures, and the measured per-update energy and magnetisation equal the
oracle's.  Statistics: Swendsen-Wang and Wolff chains of the 2-D Ising model
at beta_c on a 32 x 32 torus reproduce Kaufman's exact finite-torus energy.
O(3) and O(2) spins take Wolff's embedding clusters: from the same FP32
spins the device's reflected spins equal the oracle's bit for bit (both
compute in f64 in the reference's operation order and round to FP32), and
device chains on periodic rings reproduce the exact transfer-matrix energies.  J w < 0 is
refused with EUNSUPPORTED, as the reference asserts (models.rs:139-142,
321-324, 549-553).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import exact  # noqa: E402

pytestmark = pytest.mark.gpu

LATTICES = {
    "square16": (lambda m: m.Lattice.square(16, 16), None),
    # sides that are multiples of the 32 x 32 and 16 x 8 x 8 labelling tiles
    "square_tiles": (lambda m: m.Lattice.square(64, 96), None),
    "cube_tiles": (lambda m: m.Lattice.hypercubic([16, 16, 8]), None),
    "cube_tiles_wrap": (lambda m: m.Lattice.hypercubic([32, 8, 16]), None),
    "cube": (lambda m: m.Lattice.hypercubic([8, 6, 4]), None),
    "chain": (lambda m: m.Lattice.chain(14), None),
    "open": (lambda m: m.Lattice.square(7, 5, pbc=False), None),
    "tri": (lambda m: m.Lattice.triangular(6, 5), None),
    "graph": (lambda m: m.Lattice.from_edges(6, [0, 0, 1, 2, 3, 4, 5, 2], [1, 1, 2, 2, 4, 5, 0, 4],
                                             [1.0, 2.0, 1.0, 1.0, 3.0, 1.0, 2.0, 1.0]), None),
}


def _models(sc, orc, kind):
    if kind == "ising":
        return sc.model(sc.ISING, 1.0), orc.model(orc.ISING)
    return sc.model(sc.POTTS, 1.0, 3), orc.model(orc.POTTS, 1.0, 3)


@pytest.mark.parametrize("lat_name", list(LATTICES))
@pytest.mark.parametrize("kind", ["ising", "potts"])
@pytest.mark.parametrize("wolff", [False, True])
def test_bit_exact_against_oracle(sc, orc, lat_name, kind, wolff):
    make, _ = LATTICES[lat_name]
    lat, olat = make(sc), make(orc)
    m, om = _models(sc, orc, kind)
    seed, n_rep, steps = 31, 3, 7
    betas = [0.2, 0.44, 0.9] if kind == "ising" else [0.5, 1.0, 1.6]
    print(f"case {lat_name} {kind} wolff={wolff}", flush=True)
    st = sc.State(lat, m, n_rep, seed)
    for r in range(n_rep):
        st.init(True, r)
        st.set_beta(betas[r], r)
    start = [st.download_f64(r) for r in range(n_rep)]
    st.sweep(sc.WOLFF if wolff else sc.SWENDSEN_WANG, steps, measure=True)
    e_dev, m_dev = zip(*[st.series(r) for r in range(n_rep)])
    for r in range(n_rep):
        key = orc.replica_key(seed, r)
        s = start[r].copy()
        assert np.array_equal(s, orc.random_start(olat, om, key)), "random start differs"
        es, ms = [], []
        for t in range(steps):
            orc.cluster_steps(olat, om, s, betas[r], key, t, 1, wolff)
            es.append(orc.lib().orc_total_energy(orc.C.byref(om), olat.ptr, orc._p(s)))
            ms.append(orc.lib().orc_magnetization(orc.C.byref(om), orc._p(s), olat.n_sites))
        assert np.array_equal(st.download_f64(r), s), f"replica {r}: spins differ"
        assert np.allclose(e_dev[r], es, rtol=0, atol=1e-9), (r, e_dev[r], es)
        assert np.allclose(m_dev[r], ms, rtol=0, atol=1e-12), (r, m_dev[r], ms)
    assert st.sweeps == steps


@pytest.mark.parametrize("kind", ["ising", "potts"])
@pytest.mark.parametrize("wolff", [False, True])
def test_negative_coupling_with_negative_weights(sc, orc, kind, wolff):
    """J < 0 on bonds of negative weight is ferromagnetic (J w > 0), which
    the reference accepts; the coupling is |J| |w|."""
    src, dst, w = [0, 1, 2, 3, 4, 0, 2], [1, 2, 3, 4, 5, 3, 5], [-1.0, -2.0, -1.0, -3.0, -1.0, -2.0, 0.0]
    lat, olat = sc.Lattice.from_edges(6, src, dst, w), orc.Lattice.from_edges(6, src, dst, w)
    if kind == "ising":
        m, om = sc.model(sc.ISING, -0.7), orc.model(orc.ISING, -0.7)
    else:
        m, om = sc.model(sc.POTTS, -0.7, 3), orc.model(orc.POTTS, -0.7, 3)
    st = sc.State(lat, m, 2, 37)
    for r, beta in enumerate([0.4, 1.1]):
        st.init(True, r)
        st.set_beta(beta, r)
    start = [st.download_f64(r) for r in range(2)]
    st.sweep(sc.WOLFF if wolff else sc.SWENDSEN_WANG, 9)
    for r, beta in enumerate([0.4, 1.1]):
        s = start[r].copy()
        orc.cluster_steps(olat, om, s, beta, orc.replica_key(37, r), 0, 9, wolff)
        assert np.array_equal(st.download_f64(r), s)


def test_continues_after_metropolis(sc, orc):
    """Cluster draws have their own streams: Metropolis sweeps, then cluster
    updates numbered by the same counter, equal the oracle's sequence."""
    lat, olat = sc.Lattice.square(16, 16), orc.Lattice.square(16, 16)
    m, om = sc.model(sc.ISING, 1.0), orc.model(orc.ISING)
    st = sc.State(lat, m, 1, 5)
    st.init(True)
    st.set_beta(0.44)
    st.sweep(sc.METROPOLIS, 3)
    s = st.download_f64()
    st.sweep(sc.SWENDSEN_WANG, 4)
    orc.cluster_steps(olat, om, s, 0.44, orc.replica_key(5, 0), 3, 4, False)
    assert np.array_equal(st.download_f64(), s)


@pytest.mark.parametrize("alg", ["sw", "wolff"])
def test_energy_matches_kaufman_at_criticality(sc, alg):
    L, beta = 32, 0.4406868
    want = exact.kaufman_energy(L, L, beta)
    lat = sc.Lattice.square(L, L)
    st = sc.State(lat, sc.model(sc.ISING, 1.0), 4, 17)
    for r in range(4):
        st.init(True, r)
        st.set_beta(beta, r)
    a = sc.SWENDSEN_WANG if alg == "sw" else sc.WOLFF
    st.sweep(a, 500 if alg == "sw" else 3000)
    st.sweep(a, 6000 if alg == "sw" else 30000, measure=True)
    zs = []
    for r in range(4):
        e, _ = st.series(r)
        est = sc.binned_estimate(e, 200 if alg == "sw" else 1000)
        zs.append((est.mean - want) / est.stderr)
    print(alg, want, zs)
    assert all(abs(z) < 4.5 for z in zs), zs


# (64, 32) and (32, 16, 16) are labelled tile by tile (2-D tiles of 32 x 32,
# 3-D of 16 x 8 x 8 sites), the others by the global union-find
@pytest.mark.parametrize("dims", [(8, 6, 4), (16, 16), (14,), (64, 32), (32, 16, 16)])
@pytest.mark.parametrize("wolff", [False, True])
@pytest.mark.parametrize("n_comp", [3, 2])
def test_on3_bit_exact_against_oracle(sc, orc, dims, wolff, n_comp):
    seed, n_rep, steps = 23, 3, 6
    betas = [0.3, 0.8, 1.5]
    print(f"case O({n_comp}) {dims} wolff={wolff}", flush=True)
    lat, olat = sc.Lattice.hypercubic(list(dims)), orc.Lattice.hypercubic(list(dims), True)
    om = orc.model(orc.ON, 1.0, 2, n_comp)
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, n_comp), n_rep, seed)
    for r in range(n_rep):
        st.init(True, r)
        st.set_beta(betas[r], r)
    start = [st.download_f64(r) for r in range(n_rep)]
    st.sweep(sc.WOLFF if wolff else sc.SWENDSEN_WANG, steps, measure=True)
    for r in range(n_rep):
        key = orc.replica_key(seed, r)
        s = start[r].copy()
        es = []
        for t in range(steps):
            orc.cluster_steps(olat, om, s, betas[r], key, t, 1, wolff)
            es.append(orc.lib().orc_total_energy(orc.C.byref(om), olat.ptr, orc._p(s)))
        got = st.download_f64(r)
        assert np.array_equal(got, s), f"replica {r}: {np.count_nonzero(got != s)} components differ"
        e_dev, _ = st.series(r)
        assert np.allclose(e_dev, es, rtol=2e-6, atol=1e-6), (r, e_dev, es)
    assert st.sweeps == steps


@pytest.mark.parametrize("alg", ["sw", "wolff"])
def test_on3_ring_energy_matches_transfer_matrix(sc, alg):
    from scipy.special import spherical_in
    n, beta = 8, 1.0
    ls = np.arange(60)
    il, dil = spherical_in(ls, beta), spherical_in(ls, beta, derivative=True)
    want = -n * np.sum((2 * ls + 1) * il ** (n - 1) * dil) / np.sum((2 * ls + 1) * il**n)
    st = sc.State(sc.Lattice.hypercubic([n]), sc.model(sc.ON, 1.0, 2, 3), 4, 29)
    for r in range(4):
        st.init(True, r)
        st.set_beta(beta, r)
    a = sc.SWENDSEN_WANG if alg == "sw" else sc.WOLFF
    st.sweep(a, 500)
    st.sweep(a, 8000, measure=True)
    zs = []
    for r in range(4):
        e, _ = st.series(r)
        est = sc.binned_estimate(e, 200)
        zs.append((est.mean - want) / est.stderr)
    print(alg, want, zs)
    assert all(abs(z) < 4.5 for z in zs), zs


@pytest.mark.parametrize("alg", ["sw", "wolff"])
def test_xy_ring_energy_matches_transfer_spectrum(sc, alg):
    """XY ring of 8 sites: Z = sum_m I_m(K)^N over the transfer spectrum."""
    from scipy.special import iv
    n, beta = 8, 1.1
    m = np.arange(-60, 61)
    i, di = iv(m, beta), 0.5 * (iv(m - 1, beta) + iv(m + 1, beta))
    want = -n * np.sum(i ** (n - 1) * di) / np.sum(i**n)
    st = sc.State(sc.Lattice.hypercubic([n]), sc.model(sc.ON, 1.0, 2, 2), 4, 31)
    for r in range(4):
        st.init(True, r)
        st.set_beta(beta, r)
    a = sc.SWENDSEN_WANG if alg == "sw" else sc.WOLFF
    st.sweep(a, 500)
    st.sweep(a, 8000, measure=True)
    zs = []
    for r in range(4):
        e, _ = st.series(r)
        est = sc.binned_estimate(e, 200)
        zs.append((est.mean - want) / est.stderr)
    print(alg, want, zs)
    assert all(abs(z) < 4.5 for z in zs), zs


@pytest.mark.parametrize("alg", ["sw", "wolff"])
def test_run_one_with_cluster_updates(sc, alg):
    """Scheduler::run_one (scamper_run_one) drives cluster updates like any
    other algorithm: the Energy estimate is the Accumulator's over the
    returned series and matches Kaufman's finite-torus energy at beta_c."""
    L, beta = 32, 0.4406868
    a = sc.SWENDSEN_WANG if alg == "sw" else sc.WOLFF
    meas = 8000 if alg == "sw" else 40000
    res, e, _ = sc.run_one(sc.Lattice.square(L, L), sc.model(sc.ISING, 1.0), a, beta, True, 1000, meas, 100, 11,
                           series=True)
    assert res["Energy"].n_bins == meas // 100
    assert res["Energy"].mean == pytest.approx(e.mean(), rel=1e-12)
    est = sc.binned_estimate(e, meas // 40)  # 40 blocks, longer than the autocorrelation
    z = (est.mean - exact.kaufman_energy(L, L, beta)) / est.stderr
    print(alg, est.mean, z)
    assert abs(z) < 4.5


def test_unsupported_inputs(sc):
    on3 = sc.State(sc.Lattice.cubic(4), sc.model(sc.ON, -1.0, 2, 3), 1, 1)
    with pytest.raises(sc.ScamperError) as err:
        on3.sweep(sc.SWENDSEN_WANG, 1)
    assert err.value.code == 5  # SCAMPER_EUNSUPPORTED
    anti = sc.State(sc.Lattice.square(8, 8), sc.model(sc.ISING, -1.0), 1, 1)
    with pytest.raises(sc.ScamperError) as err:
        anti.sweep(sc.WOLFF, 1)
    assert err.value.code == 5
    neg = sc.State(sc.Lattice.from_edges(3, [0, 1], [1, 2], [1.0, -1.0]), sc.model(sc.ISING, 1.0), 1, 1)
    with pytest.raises(sc.ScamperError) as err:
        neg.sweep(sc.SWENDSEN_WANG, 1)
    assert err.value.code == 5
