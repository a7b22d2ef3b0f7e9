"""Pins the oracle's restated rules against the reference's own known
answers and against mathematics (closed forms, enumeration, invariants).

Each check is chosen so that a plausible slip in the restatement (dropped
term, wrong sign, index or order) fails it.
"""
import math
import sys
import os

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import exact  # noqa: E402


# ---------------------------------------------------------- acceptance rule

def test_acceptance_boundaries(orc):
    # Carlo.rs/src/acceptance.rs:30-44, tests/unit/pt_exchange_rule.rs:59-90
    for u in [0.0, 1e-300, 0.25, 0.999999]:
        for d in [0.0, -0.0, np.finfo(float).eps, np.finfo(float).tiny, 1.0, math.inf]:
            assert orc.accept_log_probability(d, u)
        assert not orc.accept_log_probability(-math.inf, u)
        assert not orc.accept_log_probability(math.nan, u)


def test_acceptance_rate_equals_exp(orc):
    # ln u < ln 0.3  <=>  u < 0.3 up to rounding at the boundary
    d = math.log(0.3)
    u = (np.arange(200000) + 0.5) / 200000
    acc = np.array([orc.accept_log_probability(d, float(x)) for x in u])
    assert abs(acc.mean() - 0.3) < 1e-5


def test_log_acceptance_is_minus_beta_delta(orc):
    L = orc.lib()
    assert L.orc_log_acceptance(0.5, 4.0) == -2.0
    assert L.orc_log_acceptance(0.5, math.inf) == -math.inf
    assert L.orc_log_acceptance(0.5, -math.inf) == math.inf
    assert math.copysign(1.0, L.orc_log_acceptance(0.5, 0.0)) == 1.0


# ------------------------------------------------------------ energies

def test_model_anchor_energies(orc):
    # CMC.rs/src/lattice/models.rs:698-711
    chain = orc.Lattice.chain(4, True)
    c = orc.Chain(chain, orc.model(orc.ISING, 1.0), np.ones(4))
    assert c.total_energy() == -4.0
    w = orc.Lattice.from_edges(2, [0], [1], [2.5])
    c = orc.Chain(w, orc.model(orc.ISING, 2.0), np.ones(2))
    assert c.total_energy() == -5.0


def _multigraph(orc):
    # weighted multigraph with a self-loop and parallel bonds
    # (tests/physics/lattice_exact.rs:23-96)
    src = [0, 0, 1, 2, 3, 3, 0, 2]
    dst = [1, 1, 2, 3, 0, 3, 2, 2]
    w = [1.0, 0.5, -1.25, 2.0, 0.75, 1.5, -0.5, 3.0]
    return orc.Lattice.from_edges(4, src, dst, w), src, dst, w


def _direct(kind, spins, src, dst, w, j, d=1):
    e = 0.0
    for a, b, ww in zip(src, dst, w):
        la, lb = spins[a * d:(a + 1) * d], spins[b * d:(b + 1) * d]
        if kind == 0:
            e += -j * ww * la[0] * lb[0]
        elif kind == 1:
            e += -j * ww if int(la[0]) == int(lb[0]) else 0.0
        else:
            e += -j * ww * float(np.dot(la, lb))
    return e


def test_delta_energy_equals_direct_sum_ising(orc):
    lat, src, dst, w = _multigraph(orc)
    m = orc.model(orc.ISING, 1.3)
    for cfg in range(16):
        s = np.array([1.0 if (cfg >> k) & 1 else -1.0 for k in range(4)])
        c = orc.Chain(lat, m, s)
        assert abs(c.total_energy() - _direct(0, s, src, dst, w, 1.3)) < 1e-12
        for site in range(4):
            t = s.copy()
            t[site] = -t[site]
            want = _direct(0, t, src, dst, w, 1.3) - _direct(0, s, src, dst, w, 1.3)
            assert abs(c.delta_energy(site, t[site]) - want) < 1e-12


def test_delta_energy_equals_direct_sum_potts(orc):
    lat, src, dst, w = _multigraph(orc)
    q = 3
    m = orc.model(orc.POTTS, 0.7, q)
    rng = np.random.default_rng(1)
    for _ in range(30):
        s = rng.integers(0, q, 4).astype(float)
        c = orc.Chain(lat, m, s)
        for site in range(4):
            for new in range(q):
                t = s.copy()
                t[site] = new
                want = _direct(1, t, src, dst, w, 0.7) - _direct(1, s, src, dst, w, 0.7)
                assert abs(c.delta_energy(site, float(new)) - want) < 1e-12


def test_delta_energy_equals_direct_sum_on3(orc):
    lat, src, dst, w = _multigraph(orc)
    m = orc.model(orc.ON, 1.0, 2, 3)
    rng = np.random.default_rng(2)
    s = rng.normal(size=12)
    s = (s.reshape(4, 3) / np.linalg.norm(s.reshape(4, 3), axis=1)[:, None]).ravel()
    c = orc.Chain(lat, m, s)
    for site in range(4):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        t = s.copy()
        t[3 * site:3 * site + 3] = v
        want = _direct(2, t, src, dst, w, 1.0, 3) - _direct(2, s, src, dst, w, 1.0, 3)
        assert abs(c.delta_energy(site, v) - want) < 1e-12


# ------------------------------------------------------------ binning

def test_accumulator_known_answers(orc):
    # Carlo.rs/tests/unit/accumulator.rs
    e = orc.binned_estimate(np.arange(1, 11), 10)
    assert abs(e.mean - 5.5) < 1e-10
    e = orc.binned_estimate(np.arange(1, 16), 5)
    assert abs(e.mean - 8.0) < 1e-10 and e.stderr > 0 and e.n_bins == 3
    assert orc.binned_estimate(np.arange(1, 16), 10).n_bins == 2
    e = orc.binned_estimate(np.zeros(0), 10)
    assert e.n_bins == 0 and e.mean == 0.0
    e = orc.binned_estimate(np.array([1.0]), 10)
    assert e.n_bins == 1 and abs(e.mean - 1.0) < 1e-10
    e = orc.binned_estimate(np.arange(-3, 1), 3)
    assert abs(e.mean + 1.0) < 1e-10
    e = orc.binned_estimate(np.array([1e10, 1e10]), 2)
    assert abs(e.mean - 1e10) < 1.0


def test_estimate_closed_forms(orc):
    # bins of one sample each: mean, stderr = s/sqrt(n); tau from rebinned variance
    x = np.array([1.0, 3.0, 2.0, 6.0, 4.0, 8.0])
    e = orc.binned_estimate(x, 1)
    assert e.mean == pytest.approx(x.mean(), abs=1e-15)
    assert e.stderr == pytest.approx(x.std(ddof=1) / math.sqrt(6), rel=1e-14)
    r = (x[0::2] + x[1::2]) / 2
    tau = 0.5 * max(r.var(ddof=1) / (x.var(ddof=1) / 2) - 1, 0)
    assert e.autocorr_time == pytest.approx(max(tau, 1.0), rel=1e-14)


# ------------------------------------------------------------ samplers

def test_threshold_rule_matches_exponential(orc):
    """ln(max(u,tiny)) < -beta dE  <=>  u < exp(-beta dE): count the accepting
    32-bit words near the boundary."""
    beta = 1 / 2.269185
    for de in [4.0, 8.0, 12.0]:
        p = math.exp(-beta * de)
        k = int(p * 2**32)
        for kk in range(k - 3, k + 4):
            acc = orc.accept_log_probability(-beta * de, kk * 2.0**-32)
            assert acc == (kk * 2.0**-32 < p) or abs(kk * 2.0**-32 - p) < 1e-15


def _small_square(L):
    src, dst = [], []
    for y in range(L):
        for x in range(L):
            s = x + L * y
            src += [s, s]
            dst += [(x + 1) % L + L * y, x + L * ((y + 1) % L)]
    return np.array(src), np.array(dst)


@pytest.mark.parametrize("alg", ["metropolis", "heat_bath"])
def test_colour_ordered_sweep_samples_boltzmann(orc, alg):
    """4x4 periodic Ising: long colour-ordered chains with Philox draws
    reproduce the enumerated <E>, <m^2> (|z| < 4 over 8 seeds, mean |z| < 2)."""
    L, beta = 4, 0.35
    src, dst = _small_square(L)
    ex = exact.enumerate_ising(src, dst, L * L, beta)
    lat = orc.Lattice.square(L, L)
    zs = []
    for seed in range(8):
        key = orc.replica_key(seed, 0)
        c = orc.Chain(lat, orc.model(orc.ISING), orc.random_start(lat, orc.model(orc.ISING), key))
        es, ms = [], []
        for t in range(12000):
            if alg == "metropolis":
                c.metropolis(beta, key, t)
            else:
                c.heat_bath(beta, key, t)
            if t >= 1000:
                es.append(c.energy.value)
                ms.append(c.magnetization() ** 2)
        for series, want in [(es, ex["E"]), (ms, ex["M2"])]:
            est = orc.binned_estimate(np.array(series), 200)
            zs.append((est.mean - want) / est.stderr)
    zs = np.array(zs)
    assert np.all(np.abs(zs) < 4), zs
    assert abs(zs.mean()) < 2 * 4 / math.sqrt(len(zs)) + 0.5


@pytest.mark.parametrize("j", [1.0, -1.0])
def test_potts_heat_bath_samples_boltzmann(orc, j):
    # potts_exact.rs: q=3 on a 3x3 torus vs q^N enumeration; J = -1 pins the
    # antiferromagnetic case, whose max_log is the smallest count's
    L, q, beta = 3, 3, 0.8
    src, dst = _small_square(L)
    ex = exact.enumerate_potts(src, dst, L * L, q, beta, j)
    lat = orc.Lattice.square(L, L)
    m = orc.model(orc.POTTS, j, q)
    zs = []
    for seed in range(4):
        key = orc.replica_key(100 + seed, 0)
        c = orc.Chain(lat, m, orc.random_start(lat, m, key))
        es = []
        for t in range(15000):
            c.heat_bath(beta, key, t)
            if t >= 500:
                es.append(c.energy.value)
        est = orc.binned_estimate(np.array(es), 250)
        zs.append((est.mean - ex["E"]) / est.stderr)
    assert np.all(np.abs(zs) < 4), zs


def test_potts_metropolis_samples_boltzmann(orc):
    L, q, beta = 3, 4, 0.9
    src, dst = _small_square(L)
    ex = exact.enumerate_potts(src, dst, L * L, q, beta)
    lat = orc.Lattice.square(L, L)
    m = orc.model(orc.POTTS, 1.0, q)
    zs = []
    for seed in range(4):
        key = orc.replica_key(200 + seed, 0)
        c = orc.Chain(lat, m, orc.random_start(lat, m, key))
        es = []
        for t in range(15000):
            c.metropolis(beta, key, t)
            if t >= 500:
                es.append(c.energy.value)
        est = orc.binned_estimate(np.array(es), 250)
        zs.append((est.mean - ex["E"]) / est.stderr)
    assert np.all(np.abs(zs) < 4), zs


def test_cached_energy_never_drifts(orc):
    # tests/exact/classic_models.rs:280-315: 3-D Ising, 6^3, cache == recompute
    lat = orc.Lattice.hypercubic([6, 6, 6])
    m = orc.model(orc.ISING)
    key = orc.replica_key(3, 0)
    c = orc.Chain(lat, m, orc.random_start(lat, m, key))
    for t in range(300):
        c.metropolis(0.25, key, t)
    assert c.energy.value == c.total_energy()


def test_stencil_oracle_equals_csr_oracle(orc):
    for dims, kind, q, hb in [((8, 8), 0, 2, False), ((8, 8), 0, 2, True), ((4, 6, 4), 0, 2, False),
                              ((8, 8), 1, 3, False), ((8, 8), 1, 4, True), ((6, 4, 4), 1, 3, True)]:
        lat = orc.Lattice.hypercubic(list(dims))
        m = orc.model(kind, 1.0, q)
        key = orc.replica_key(11, 0)
        s0 = orc.random_start(lat, m, key)
        c = orc.Chain(lat, m, s0)
        beta = 0.6
        for t in range(5):
            (c.heat_bath if hb else c.metropolis)(beta, key, t)
        s = s0.astype(np.int8)
        orc.stencil_sweeps(s, dims, kind, q, beta, 1.0, hb, key, 0, 5)
        assert np.array_equal(s.astype(float), c.spins), (dims, kind, q, hb)


# ------------------------------------------------------------ O(3)

def test_on3_heat_bath_conditional_is_langevin(orc):
    """For a single site in a fixed field the sampled cos(theta) has mean
    coth(kappa) - 1/kappa (heat_bath_exact.rs:183-226)."""
    # star graph: site 0 bonded to one frozen neighbour; sweep colour 0 only
    lat = orc.Lattice.from_edges(2, [0], [1])
    m = orc.model(orc.ON, 1.0, 2, 3)
    beta = 1.7
    field = np.array([0.0, 0.6, 0.8])
    cos = []
    for t in range(20000):
        s = np.concatenate([[1.0, 0.0, 0.0], field])
        c = orc.Chain(lat, m, s)
        c.on3_heat_bath(beta, orc.replica_key(5, 0), t)
        cos.append(np.dot(c.spins[:3], field))
    # site 1 is updated too (colour 1) but site 0's sample is taken against
    # the frozen field of the original site 1
    kappa = beta
    want = 1 / math.tanh(kappa) - 1 / kappa
    cos = np.array(cos)
    assert abs(cos.mean() - want) < 5 * cos.std() / math.sqrt(cos.size)


def test_over_relaxation_conserves_energy_and_is_involution(orc):
    lat = orc.Lattice.hypercubic([4, 4, 4])
    m = orc.model(orc.ON, 1.0, 2, 3)
    s = orc.random_start(lat, m, orc.replica_key(8, 0))
    c = orc.Chain(lat, m, s)
    e0 = c.total_energy()
    c.over_relaxation()
    assert abs(c.total_energy() - e0) < 1e-9
    assert abs(c.energy.value - e0) < 1e-9
    norms = np.linalg.norm(c.spins.reshape(-1, 3), axis=1)
    assert np.allclose(norms, 1.0, atol=1e-12)


# ------------------------------------------------------------ multi-spin, PT

def test_multispin_bits_equal_plain_replicas(orc):
    """Bit r of the per-bit-beta multi-spin sweep follows exactly the plain
    Metropolis chain of replica r (same key, same draws)."""
    lat = orc.Lattice.hypercubic([4, 4, 4])
    m = orc.model(orc.ISING)
    seed = 17
    betas = 0.2 + 0.01 * np.arange(64)
    chains = [orc.Chain(lat, m, orc.random_start(lat, m, orc.replica_key(seed, r))) for r in range(64)]
    words = np.zeros(lat.n_sites, np.uint64)
    for r, c in enumerate(chains):
        words |= (c.spins > 0).astype(np.uint64) << np.uint64(r)
    for t in range(3):
        orc.multispin_sweep(lat, words, betas, 1.0, seed, 0, t)
        for r, c in enumerate(chains):
            c.metropolis(betas[r], orc.replica_key(seed, r), t)
    for r, c in enumerate(chains):
        bit = ((words >> np.uint64(r)) & np.uint64(1)).astype(float) * 2 - 1
        assert np.array_equal(bit, c.spins)
        assert orc.multispin_energy(lat, words, r) == c.total_energy()


def test_pt_two_state_acceptance(orc):
    """tests/mpi/pt_dynamics.rs: betas (0.3, 0.7), x in {0,1} drawn from
    pi_beta(x=1) = 1/(1+e^beta): mean acceptance of min(1, e^R), R =
    (b_a - b_b)(x_a - x_b), is 0.9063."""
    betas = np.array([0.3, 0.7])
    rng = np.random.default_rng(3)
    acc = 0
    n = 40000
    for step in range(1, n + 1):
        labels = np.array([0, 1], np.int32)
        x = [(rng.random() < 1 / (1 + math.exp(b))) * 1.0 for b in betas]
        acc += orc.pt_exchange(labels, x, betas, 9, 2 * step - 1)
    want = 0.0
    for xa in (0, 1):
        for xb in (0, 1):
            pa = 1 / (1 + math.exp(0.3)) if xa else 1 - 1 / (1 + math.exp(0.3))
            pb = 1 / (1 + math.exp(0.7)) if xb else 1 - 1 / (1 + math.exp(0.7))
            want += pa * pb * min(1.0, math.exp((0.3 - 0.7) * (xa - xb)))
    assert abs(want - 0.9063) < 1e-3
    assert abs(acc / n - want) < 5 * math.sqrt(want * (1 - want) / n)


def test_pt_pairing_alternates(orc):
    betas = np.linspace(0.1, 0.8, 6)
    e = np.zeros(6)  # equal energies: every proposed swap is accepted
    labels = np.arange(6, dtype=np.int32)
    assert orc.pt_exchange(labels, e, betas, 1, 1) == 3  # pairs (0,1)(2,3)(4,5)
    assert list(labels) == [1, 0, 3, 2, 5, 4]
    assert orc.pt_exchange(labels, e, betas, 1, 2) == 2  # pairs (1,2)(3,4)
    assert sorted(labels) == list(range(6))


# ------------------------------------------------------------ faithful path

def test_faithful_plain_path_matches_transfer_matrix(orc):
    L, beta = 6, 0.4
    lat = orc.Lattice.square(L, L)
    want = exact.transfer_matrix_energy(L, L, beta)
    est, n = orc.run_one_ising(lat, 1.0, beta, 0, 42, 2000, 40000, 400)
    z = (est["Energy"].mean - want) / est["Energy"].stderr
    assert abs(z) < 4
    assert n == L * L * 42000
