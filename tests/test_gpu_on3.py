"""O(3) heat bath and over-relaxation on the GPU (FP32 spins, FP64
observables) against the FP64 oracle with the same draws.

Tolerances: the device does spin arithmetic in FP32, so per-site agreement
with the FP64 oracle after one sweep is checked to 2e-4 absolute per
component (single-precision rounding through the inverse CDF, the frame
rotation and the renormalisation), and E, |M| after a few hybrid sweeps to
1e-4 relative; long-run statistics agree to |z| < 4.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HEIS = dict(kind=2, q=2, spin_dim=3)


def _setup(sc, orc, dims, seed, beta):
    lat = sc.Lattice.hypercubic(list(dims))
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, 3), 1, seed)
    st.init(True)
    st.set_beta(beta)
    olat = orc.Lattice.hypercubic(list(dims))
    m = orc.model(orc.ON, 1.0, 2, 3)
    start = orc.random_start(olat, m, orc.replica_key(seed, 0))
    return lat, st, olat, m, start


@pytest.mark.parametrize("dims", [(8, 8, 8), (16, 16), (6, 4, 8)])
def test_random_start_matches(sc, orc, dims):
    _, st, _, _, start = _setup(sc, orc, dims, 3, 0.69)
    got = st.download_f64()
    assert np.max(np.abs(got - start)) < 2e-6
    assert np.allclose(np.linalg.norm(got.reshape(-1, 3), axis=1), 1.0, atol=1e-6)


@pytest.mark.parametrize("alg", [2, 3])
def test_one_sweep_per_site_error(sc, orc, alg):
    dims, seed, beta = (8, 8, 8), 11, 1 / 1.44
    _, st, olat, m, start = _setup(sc, orc, dims, seed, beta)
    st.upload_f64(start)
    chain = orc.Chain(olat, m, start)
    st.sweep(alg, 1, measure=True)
    if alg == 2:
        chain.on3_heat_bath(beta, orc.replica_key(seed, 0), 0)
    else:
        chain.over_relaxation()
    got = st.download_f64()
    err = np.max(np.abs(got - chain.spins))
    print(f"max per-component error after one sweep (alg {alg}): {err:.2e}")
    assert err < 2e-4
    e, mag = st.series()
    assert abs(e[-1] - chain.total_energy()) < 1e-4 * abs(chain.total_energy())
    assert abs(mag[-1] - chain.magnetization()) < 1e-4


def test_over_relaxation_conserves_energy(sc, orc):
    _, st, _, _, _ = _setup(sc, orc, (8, 8, 8), 4, 0.7)
    e0, _ = st.observables()
    st.sweep(3, 10)
    e1, _ = st.observables()
    assert abs(e1 - e0) < 1e-3 * abs(e0)


def test_hybrid_observable_error(sc, orc):
    dims, seed, beta = (8, 8, 8), 13, 1 / 1.45
    _, st, olat, m, start = _setup(sc, orc, dims, seed, beta)
    st.upload_f64(start)
    chain = orc.Chain(olat, m, start)
    key = orc.replica_key(seed, 0)
    st.sweep(4, 3, measure=True)
    for t in range(3):
        chain.on3_heat_bath(beta, key, t)
        chain.over_relaxation()
    e, mag = st.series()
    rel_e = abs(e[-1] - chain.total_energy()) / abs(chain.total_energy())
    rel_m = abs(mag[-1] - chain.magnetization()) / max(chain.magnetization(), 1e-3)
    print(f"hybrid x3: relative error E {rel_e:.2e}, |M| {rel_m:.2e}")
    assert rel_e < 1e-3 and rel_m < 1e-2


def test_statistics_match_fp64_oracle(sc, orc):
    """32 device replicas vs one FP64 oracle chain at L=6, T=1.45: mean
    energy per site agrees within |z| < 4 (binned errors on both sides)."""
    dims, beta, therm, meas = (6, 6, 6), 1 / 1.45, 500, 6000
    lat = sc.Lattice.hypercubic(list(dims))
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, 3), 32, 17)
    for r in range(32):
        st.init(True, r)
        st.set_beta(beta, r)
    st.sweep(4, therm)
    st.sweep(4, meas, measure=True)
    n = lat.n_sites
    means = []
    for r in range(32):
        e, _ = st.series(r)
        means.append(e.mean() / n)
    gpu_mean, gpu_se = np.mean(means), np.std(means, ddof=1) / math.sqrt(32)
    olat = orc.Lattice.hypercubic(list(dims))
    m = orc.model(orc.ON, 1.0, 2, 3)
    key = orc.replica_key(99, 0)
    chain = orc.Chain(olat, m, orc.random_start(olat, m, key))
    es = []
    for t in range(therm + meas):
        chain.on3_heat_bath(beta, key, t)
        chain.over_relaxation()
        if t >= therm:
            es.append(chain.energy.value / n)
    est = orc.binned_estimate(np.array(es), 200)
    z = (gpu_mean - est.mean) / math.sqrt(gpu_se**2 + est.stderr**2)
    print(f"E/N gpu {gpu_mean:.5f}+-{gpu_se:.5f} oracle {est.mean:.5f}+-{est.stderr:.5f} z={z:.2f}")
    assert abs(z) < 4


# ------------------------------------------------------------------ O(2) (XY)
# The device draws the von Mises angle in f64 from the same Philox words as
# the oracle and stores FP32 components, so one sweep from the same start
# agrees per component to single-precision rounding of cos/sin (2e-6), plus,
# on the second colour, the shift of atan2(h) that FP32 storage of the first
# colour's new spins causes: at most 2^-23 per neighbour component over |h|
# (1.35e-5 at a site with |h| = 0.002).  The over-relaxation sweep reflects
# in FP32 (1e-5).


def _xy_field_norm(spins, dims):
    """|h| at every site (J = 1, periodic, x fastest) from site-major spins."""
    s = spins.reshape(tuple(reversed(dims)) + (2,))
    h = sum(np.roll(s, sh, ax) for ax in range(len(dims)) for sh in (1, -1))
    return np.linalg.norm(h, axis=-1).reshape(-1)


def _setup_xy(sc, orc, dims, seed, beta, n_rep=1):
    lat = sc.Lattice.hypercubic(list(dims))
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, 2), n_rep, seed)
    for r in range(n_rep):
        st.init(True, r)
        st.set_beta(beta, r)
    olat = orc.Lattice.hypercubic(list(dims))
    m = orc.model(orc.ON, 1.0, 2, 2)
    return lat, st, olat, m


@pytest.mark.parametrize("dims", [(16, 16), (8, 4, 6), (32,)])
def test_xy_random_start_matches(sc, orc, dims):
    _, st, olat, m = _setup_xy(sc, orc, dims, 5, 1.0)
    got = st.download_f64()
    want = orc.random_start(olat, m, orc.replica_key(5, 0))
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) < 2e-7
    assert np.allclose(np.linalg.norm(got.reshape(-1, 2), axis=1), 1.0, atol=1e-6)


@pytest.mark.parametrize("dims,beta", [((16, 16), 1.12), ((8, 8, 8), 0.45), ((64,), 2.0), ((16, 16), 1e-9)])
def test_xy_one_heat_bath_sweep(sc, orc, dims, beta):
    seed = 21
    _, st, olat, m = _setup_xy(sc, orc, dims, seed, beta)
    start = st.download_f64()
    chain = orc.Chain(olat, m, start.copy())
    st.sweep(2, 1, measure=True)
    chain.on2_heat_bath(beta, orc.replica_key(seed, 0), 0)
    got = st.download_f64()
    err = np.max(np.abs(got - chain.spins).reshape(-1, 2), axis=1)
    tol = 2e-6 + 2.0**-23 * 2 * len(dims) / _xy_field_norm(chain.spins, dims)
    print(f"XY {dims} beta {beta}: max per-component error after one heat-bath sweep {err.max():.2e}")
    assert np.all(err < tol), (err.max(), np.argmax(err / tol))
    assert np.median(err) < 5e-7
    e, mag = st.series()
    assert abs(e[-1] - chain.total_energy()) < 1e-5 * max(abs(chain.total_energy()), 1.0)
    assert abs(mag[-1] - chain.magnetization()) < 1e-5


def test_xy_over_relaxation_sweep(sc, orc):
    _, st, olat, m = _setup_xy(sc, orc, (16, 16), 8, 1.0)
    start = st.download_f64()
    chain = orc.Chain(olat, m, start.copy())
    e0, _ = st.observables()
    st.sweep(3, 1)
    chain.over_relaxation()
    assert np.max(np.abs(st.download_f64() - chain.spins)) < 1e-5
    st.sweep(3, 20)
    e1, _ = st.observables()
    assert abs(e1 - e0) < 1e-4 * abs(e0)


def test_xy_hybrid_sweeps_follow_oracle(sc, orc):
    dims, seed, beta = (16, 16), 31, 1.0
    _, st, olat, m = _setup_xy(sc, orc, dims, seed, beta)
    chain = orc.Chain(olat, m, st.download_f64())
    key = orc.replica_key(seed, 0)
    st.sweep(4, 3, measure=True)
    for t in range(3):
        chain.on2_heat_bath(beta, key, t)
        chain.over_relaxation()
    e, _ = st.series()
    rel = abs(e[-1] - chain.total_energy()) / abs(chain.total_energy())
    print(f"XY hybrid x3: relative energy error {rel:.2e}")
    assert rel < 1e-3


def test_xy_ring_energy_exact(sc):
    """64 replicas of a periodic XY ring of 8 sites: <E> from the transfer
    spectrum I_m(K), Z = sum_m I_m(K)^N, within |z| < 4.5."""
    from scipy.special import iv

    n, beta, n_rep, therm, meas = 8, 1.3, 64, 200, 4000
    mm = np.arange(-60, 61)
    i = iv(mm, beta)
    di = 0.5 * (iv(mm - 1, beta) + iv(mm + 1, beta))
    want = -n * np.sum(i ** (n - 1) * di) / np.sum(i**n)
    lat = sc.Lattice.hypercubic([n])
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, 2), n_rep, 77)
    for r in range(n_rep):
        st.init(True, r)
        st.set_beta(beta, r)
    st.sweep(2, therm)
    st.sweep(2, meas, measure=True)
    means = np.array([st.series(r)[0].mean() for r in range(n_rep)])
    z = (means.mean() - want) / (means.std(ddof=1) / math.sqrt(n_rep))
    print(f"XY ring: <E> {means.mean():.5f} exact {want:.5f} z={z:.2f}")
    assert abs(z) < 4.5


def test_xy_rejects_other_components(sc):
    lat = sc.Lattice.hypercubic([8, 8])
    for n_comp in (1, 4):
        with pytest.raises(sc.ScamperError):
            sc.State(lat, sc.model(sc.ON, 1.0, 2, n_comp), 1, 1)


def test_xy_bench_shape_against_fp64_oracle(sc, orc):
    """The XY workload at its own shape: 96^3 periodic, 32 replicas at
    T = 2.15 + 0.1 r / 31.  The heat bath runs the four-sites-per-lane
    kernel on its full grid; three replicas (first, middle, last) are
    replayed from the device's spins by the FP64 oracle (sample_von_mises,
    models.rs:666-689, then microcanonical.rs:42-80) for one hybrid sweep.
    The sampler's accept/reject follows the f64 rule, so the spins differ
    only by FP32 rounding: of the stored components and of the
    over-relaxation, which reaches a spin as ~1e-7 / |h|, and a neighbour's
    error enters the field a site is drawn about.  Chains of small fields
    give a tail far wider than the bulk (the largest error, 1.8e-4, sits at
    a site with |h| ~ 1 next to sites of small fields), so the bars are the
    mean (1e-6), the 99.99% quantile (5e-5) and the maximum (1e-3, far below
    the O(1) error of a decision that went the other way), and the
    replica's energy and |M| (1e-6)."""
    L, R, seed = 96, 32, 42
    print(f"case xy {L}^3 x {R} replicas", flush=True)
    dims = (L, L, L)
    lat = sc.Lattice.hypercubic(list(dims))
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, 2), R, seed)
    betas = [1.0 / (2.15 + 0.1 * r / 31) for r in range(R)]
    for r in range(R):
        st.init(True, r)
        st.set_beta(betas[r], r)
    st.sweep(sc.HYBRID_HB_OR, 2)
    replicas = (0, 13, 31)
    starts = {r: st.download_f64(r) for r in replicas}
    sweep = st.sweeps
    st.sweep(sc.HYBRID_HB_OR, 1, measure=True)
    olat = orc.Lattice.hypercubic(list(dims))
    m = orc.model(orc.ON, 1.0, 2, 2)
    for r in replicas:
        chain = orc.Chain(olat, m, starts[r].copy())
        chain.on2_heat_bath(betas[r], orc.replica_key(seed, r), sweep)
        chain.over_relaxation()
        got = st.download_f64(r)
        err = np.abs(got - chain.spins).reshape(-1, 2).max(axis=1)
        e, mag = st.series(r)
        e_ref, m_ref = chain.total_energy(), chain.magnetization()
        rel_e = abs(e[-1] - e_ref) / abs(e_ref)
        print(f"replica {r}: per-site max {err.max():.2e}, mean {err.mean():.2e}, "
              f"99.99% {np.quantile(err, 0.9999):.2e}; E rel {rel_e:.2e}, |M| abs {abs(mag[-1] - m_ref):.2e}")
        assert err.max() < 1e-3 and np.quantile(err, 0.9999) < 5e-5 and err.mean() < 1e-6
        assert rel_e < 1e-6 and abs(mag[-1] - m_ref) < 1e-6
    st.close()


def _fields(spins, L):
    """Local field h = sum of the six neighbours of every site of a periodic
    L^3 lattice, from site-major f64 spins (x fastest)."""
    a = spins.reshape(L, L, L, 3)
    return sum(np.roll(a, sh, axis=ax) for ax in range(3) for sh in (1, -1)).reshape(-1, 3)


def test_c4_shape_against_fp64_oracle(sc, orc):
    """The 3-D Heisenberg workload at its own shape: 96^3 periodic, 32
    replicas at T = 1.40 + 0.09 r / 31, FP32 spins from the seeded isotropic
    start.  Lx/2 = 48 is even, so the heat-bath half runs the two-ranks-per-
    thread kernel on its full XCD-remapped grid.  Three replicas (first,
    middle, last) are replayed from the device's spins by the FP64 oracle
    (models.rs:617-662 heat bath, microcanonical.rs:42-80 over-relaxation)
    for one hybrid sweep.

    Error model: both halves build the new spin from the direction h/|h| of
    the local field (the heat bath's frame, the reflection), so FP32
    rounding of order 1e-7 reaches a spin as about 1e-7 / |h|, and a
    neighbour's error enters h.  Among 884,736 sites a few see small fields,
    so the tail of the per-component error is wider than its bulk: the
    bars are the maximum (2e-4), the 99.99% quantile (5e-5) and the mean
    (1e-6) per component, and the replica's energy and |M| (1e-6)."""
    L, R, seed = 96, 32, 42
    print(f"case on3 bench shape {L}^3 x {R} replicas", flush=True)
    lat = sc.Lattice.hypercubic([L, L, L])
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, 3), R, seed)
    betas = [1.0 / (1.40 + 0.09 * r / 31) for r in range(R)]
    for r in range(R):
        st.init(True, r)
        st.set_beta(betas[r], r)
    st.sweep(sc.HYBRID_HB_OR, 2)  # leave the exact random start
    replicas = (0, 13, 31)
    starts = {r: st.download_f64(r) for r in replicas}
    sweep = st.sweeps
    st.sweep(sc.HYBRID_HB_OR, 1, measure=True)
    olat = orc.Lattice.hypercubic([L, L, L])
    m = orc.model(orc.ON, 1.0, 2, 3)
    idx = np.arange(L ** 3)
    colour = ((idx % L) + (idx // L % L) + (idx // (L * L))) & 1
    for r in replicas:
        chain = orc.Chain(olat, m, starts[r].copy())
        chain.on3_heat_bath(betas[r], orc.replica_key(seed, r), sweep)
        after_hb = chain.spins.copy()
        chain.over_relaxation()
        # the smallest |h| each site saw: heat bath (colour 0 from the start,
        # colour 1 from the heat-bathed colour 0), over-relaxation (colour 0
        # from the heat-bathed colour 1, colour 1 from the final colour 0)
        hb = np.where(colour == 0, np.linalg.norm(_fields(starts[r], L), axis=1),
                      np.linalg.norm(_fields(after_hb, L), axis=1))
        orl = np.where(colour == 0, np.linalg.norm(_fields(after_hb, L), axis=1),
                       np.linalg.norm(_fields(chain.spins, L), axis=1))
        hmin = np.minimum(hb, orl)
        got = st.download_f64(r)
        err = np.abs(got - chain.spins).reshape(-1, 3).max(axis=1)
        e, mag = st.series(r)
        e_ref, m_ref = chain.total_energy(), chain.magnetization()
        rel_e = abs(e[-1] - e_ref) / abs(e_ref)
        worst = int(np.argmax(err))
        print(f"replica {r}: per-site max {err.max():.2e} (|h|min {hmin[worst]:.2e} there), mean {err.mean():.2e}, "
              f"99.99% {np.quantile(err, 0.9999):.2e}; "
              f"E rel {rel_e:.2e}, |M| abs {abs(mag[-1] - m_ref):.2e}")
        assert err.max() < 2e-4 and np.quantile(err, 0.9999) < 5e-5 and err.mean() < 1e-6
        assert rel_e < 1e-6 and abs(mag[-1] - m_ref) < 1e-6
    st.close()
