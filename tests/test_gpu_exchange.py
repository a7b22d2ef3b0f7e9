"""Replica exchange for batches of replicas (scamper_state_set_ladder /
scamper_state_exchange): ParallelTemperingCompatible for ClassicalMC
(CMC.rs/src/classical_mc.rs:203-229) driven by PtExchange::try_exchange
(Carlo.rs/src/parallel_tempering.rs:362-419).

Discrete models: every replica's trajectory is replayed in the oracle with
the beta of the label it holds at each step, the oracle's exchange rule
(orc_pt_exchange) moves the labels from the oracle's own energies, and the
device must agree on spins, energies and labels after every step.  O(3)
spins are FP32 on the device: there the rule is checked on the device's own
energies, and the energies against the oracle's FP64 recount of the device
spins.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ladder(n, lo, hi):
    return np.array([lo * (hi / lo) ** (k / (n - 1)) for k in range(n)])


CASES = [
    # lattice builder, args, kind, q, J, algorithm, replicas, beta range, steps
    ("hypercubic", [32, 32], 0, 2, 1.0, 0, 16, (0.30, 0.60), 12),   # register-blocked kernel
    ("hypercubic", [12, 10], 0, 2, 1.0, 0, 8, (0.30, 0.60), 10),    # tiled kernel
    ("hypercubic", [32, 32], 0, 2, 1.0, 1, 8, (0.30, 0.60), 8),     # heat bath
    ("hypercubic", [16, 16], 1, 3, 1.0, 0, 8, (0.85, 1.15), 10),    # Potts, tiled kernel
    ("hypercubic", [8, 8], 1, 3, 1.0, 1, 8, (0.85, 1.15), 8),       # Potts heat bath (weight tables)
    ("hypercubic", [8, 6, 4], 0, 2, 1.0, 0, 12, (0.15, 0.30), 8),   # 3-D
    ("triangular", [12, 12], 0, 2, -1.0, 0, 8, (0.20, 1.00), 8),    # graph lattice, antiferromagnet
]


@pytest.mark.parametrize("builder,args,kind,q,j,alg,n_rep,brange,steps", CASES)
def test_exchange_trajectory_matches_oracle(sc, orc, builder, args, kind, q, j, alg, n_rep, brange, steps):
    seed = 31 + n_rep
    print(f"case {builder}{args} kind={kind} q={q} J={j} alg={alg} replicas={n_rep}", flush=True)
    lat = getattr(sc.Lattice, builder)(*([args] if builder == "hypercubic" else args))
    olat = getattr(orc.Lattice, builder)(*([args] if builder == "hypercubic" else args))
    st = sc.State(lat, sc.model(kind, j, q), n_rep, seed)
    ladder = _ladder(n_rep, *brange)
    st.set_ladder(ladder)
    om = orc.model(kind, j, q)
    chains = []
    for r in range(n_rep):
        st.init(True, r)
        chains.append(orc.Chain(olat, om, st.download_f64(r)))
    labels = np.arange(n_rep, dtype=np.int32)
    for t in range(steps):
        st.sweep(alg, 1)
        for r, ch in enumerate(chains):
            beta = float(ladder[labels[r]])
            key = orc.replica_key(seed, r)
            (ch.metropolis if alg == 0 else ch.heat_bath)(beta, key, t)
        st.exchange(t + 1)
        energies = np.array([ch.total_energy() for ch in chains])
        orc.pt_exchange(labels, energies, ladder, seed, t + 1)
        assert np.array_equal(st.exchange_energies(), energies), f"step {t}: energies"
        assert np.array_equal(st.labels(), labels), f"step {t}: labels {st.labels()} vs {labels}"
        for r, ch in enumerate(chains):
            assert np.array_equal(st.download_f64(r), ch.spins), f"step {t} replica {r}: spins"
    accepted, ties, rounds = st.exchange_stats()
    assert rounds == steps and ties == 0
    assert 0 < accepted
    for r in range(n_rep):
        assert st.get_beta(r) == ladder[labels[r]]


def test_exchange_rule_on3(sc, orc):
    """O(3), 32 replicas at T = 1.40 .. 1.49 (the 96^3 workload's ladder) on 8^3: hybrid
    sweeps and exchange rounds.  The labels follow the oracle's rule applied
    to the device's energies; those energies equal the FP64 recount of the
    device's spins to FP32 accuracy."""
    seed, n_rep, L = 5, 32, 8
    lat = sc.Lattice.hypercubic([L, L, L])
    olat = orc.Lattice.hypercubic([L, L, L])
    om = orc.model(orc.ON, 1.0, 2, 3)
    st = sc.State(lat, sc.model(sc.ON, 1.0, 2, 3), n_rep, seed)
    ladder = np.array([1.0 / (1.40 + 0.09 * r / 31) for r in range(n_rep)])
    st.set_ladder(ladder)
    for r in range(n_rep):
        st.init(True, r)
    labels = np.arange(n_rep, dtype=np.int32)
    for t in range(20):
        st.sweep(sc.HYBRID_HB_OR, 1)
        st.exchange(t + 1)
        e_dev = st.exchange_energies()
        orc.pt_exchange(labels, e_dev, ladder, seed, t + 1)
        assert np.array_equal(st.labels(), labels), f"step {t}"
        if t % 5 == 4:
            for r in range(0, n_rep, 7):
                want = orc.Chain(olat, om, st.download_f64(r)).total_energy()
                assert abs(e_dev[r] - want) <= 2e-5 * abs(want), f"replica {r}: {e_dev[r]} vs {want}"
    accepted, ties, rounds = st.exchange_stats()
    assert rounds == 20 and accepted > 0
    for r in range(n_rep):
        assert st.get_beta(r) == ladder[labels[r]]


def test_set_beta_ends_tempering(sc):
    lat = sc.Lattice.square(32, 32)
    st = sc.State(lat, sc.model(sc.ISING), 4, 3)
    ladder = np.array([0.2, 0.3, 0.45, 0.6])
    st.set_ladder(ladder)
    for r in range(4):
        st.init(True, r)
    for t in range(6):
        st.sweep(sc.METROPOLIS, 1)
        st.exchange(t + 1)
    labels = st.labels()
    st.set_beta(1.0, 0)
    with pytest.raises(sc.ScamperError):
        st.labels()
    with pytest.raises(sc.ScamperError):
        st.exchange(7)
    assert st.get_beta(0) == 1.0
    for r in range(1, 4):
        assert st.get_beta(r) == ladder[labels[r]]


@pytest.mark.parametrize("case", ["on3", "ising_fast", "potts_tiles"])
def test_pt_step_equals_sweep_then_exchange(sc, case):
    """scamper_state_pt_step (captured graphs of 16 steps where the kernels
    read the device clock, step by step elsewhere) equals the same steps made
    one call at a time: sweep, then exchange round t + 1; spins, labels,
    exchange energies and the measured series agree exactly."""
    n_rep, steps, seed = 8, 37, 61
    if case == "on3":
        lat, m, alg, ladder = sc.Lattice.hypercubic([8, 8, 8]), sc.model(sc.ON, 1.0, 2, 3), sc.HYBRID_HB_OR, \
            np.linspace(0.62, 0.74, n_rep)
    elif case == "ising_fast":
        lat, m, alg, ladder = sc.Lattice.square(64, 32), sc.model(sc.ISING), sc.METROPOLIS, np.linspace(0.38, 0.5, n_rep)
    else:
        lat, m, alg, ladder = sc.Lattice.square(20, 12), sc.model(sc.POTTS, 1.0, 3), sc.METROPOLIS, \
            np.linspace(0.9, 1.1, n_rep)
    a, b = sc.State(lat, m, n_rep, seed), sc.State(lat, m, n_rep, seed)
    for st in (a, b):
        st.set_ladder(ladder)
        for r in range(n_rep):
            st.init(True, r)
    a.pt_step(alg, steps, measure=True)
    energies, mags = [], []
    for t in range(steps):
        b.sweep(alg, 1)
        b.exchange(t + 1)
        energies.append(b.exchange_energies())
        mags.append([b.observables(r)[1] for r in range(n_rep)])
    # O(N) recounts add FP64 partial sums with atomics, in no fixed order:
    # equal to the last bits; integer counts are exact
    same = (lambda x, y: np.allclose(x, y, rtol=1e-12, atol=0)) if case == "on3" else np.array_equal
    assert (a.pt_graph_launches() > 0) == (case != "potts_tiles")
    assert np.array_equal(a.labels(), b.labels())
    assert same(a.exchange_energies(), b.exchange_energies())
    for r in range(n_rep):
        if case == "on3":
            assert np.array_equal(a.download_f32(r), b.download_f32(r)), f"replica {r}"
        else:
            assert np.array_equal(a.download_i8(r), b.download_i8(r)), f"replica {r}"
        e, mm = a.series(r)
        assert e.size == steps
        assert same(e, [energies[t][r] for t in range(steps)]), f"replica {r}: series energies"
        assert same(mm, [mags[t][r] for t in range(steps)]), f"replica {r}: series |m|"
    assert a.sweeps == b.sweeps == steps
    assert a.exchange_stats() == b.exchange_stats()
