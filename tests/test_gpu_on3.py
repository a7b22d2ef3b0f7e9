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
