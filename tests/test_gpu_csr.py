"""GPU parity of the graph-lattice (CSR) path against the oracle's CSR
restatement, bit for bit: non-bipartite lattices (greedy colouring with 3+
classes), open boundaries, and a weighted multigraph with parallel bonds and
self-loops (the reference's tests/physics/lattice_exact.rs:23-96 shapes).

Real weights (non-integer, |w| > 127 or weighted degree > 32) and the Potts
heat bath on weighted graphs run on the f64 kernels, which restate the
reference's f64 rules: spins equal the oracle's bit for bit, and E agrees to
1e-9 relative (the device sums bonds per site, the oracle per edge)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(sc, orc, name, *args, weights=None):
    if name == "edges":
        n, src, dst = args
        return (sc.Lattice.from_edges(n, src, dst, weights), orc.Lattice.from_edges(n, src, dst, weights))
    return getattr(sc.Lattice, name)(*args), getattr(orc.Lattice, name)(*args)


MULTI = (6, [0, 0, 1, 2, 3, 3, 0, 2, 4, 5, 1, 4], [1, 1, 2, 3, 0, 3, 2, 2, 5, 0, 4, 4])
MULTI_W = [1.0, 2.0, -1.0, 3.0, 1.0, 2.0, -2.0, 1.0, 1.0, 2.0, 1.0, 5.0]

LATTICES = [
    ("triangular", (6, 6), None),
    ("honeycomb", (6, 4), None),
    ("kagome", (4, 4), None),
    ("square", (5, 5), None),          # odd periodic: 3+ colours
    ("hypercubic", ([6, 4], False), None),  # open boundaries
    ("hypercubic", ([3, 3, 4], True), None),
    ("chain", (7,), None),
    ("edges", MULTI, None),
    ("edges", MULTI, MULTI_W),
]


@pytest.mark.parametrize("name,args,weights", LATTICES)
@pytest.mark.parametrize("kind,q,alg", [(0, 2, 0), (0, 2, 1), (1, 3, 0), (1, 4, 1)])
def test_csr_bit_exact(sc, orc, name, args, weights, kind, q, alg):
    print(f"case {name}{args} w={weights is not None} kind={kind} q={q} alg={alg}", flush=True)
    lat, olat = _pair(sc, orc, name, *args, weights=weights)
    seed, beta, n = 77, 0.7, 6
    st = sc.State(lat, sc.model(kind, 1.0, q), 1, seed)
    st.init(True)
    key = orc.replica_key(seed, 0)
    m = orc.model(kind, 1.0, q)
    start = orc.random_start(olat, m, key)
    assert np.array_equal(st.download_f64(), start)
    chain = orc.Chain(olat, m, start)
    st.set_beta(beta)
    st.sweep(alg, n, measure=True)
    for t in range(n):
        (chain.metropolis if alg == 0 else chain.heat_bath)(beta, key, t)
    assert np.array_equal(st.download_f64(), chain.spins)
    e, mag = st.series()
    assert e[-1] == chain.total_energy() == chain.energy.value
    assert mag[-1] == chain.magnetization()


def _glass(L, seed):
    """Periodic L x L Edwards-Anderson couplings, Gaussian (edges +x, +y)."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for y in range(L):
        for x in range(L):
            s = x + L * y
            src += [s, s]
            dst += [(x + 1) % L + L * y, x + L * ((y + 1) % L)]
    return L * L, src, dst, list(rng.normal(size=len(src)))


REAL = [
    ("multigraph", MULTI[0], MULTI[1], MULTI[2], [0.5, 1.25, -0.75, 2.5, 1.0, 0.3, -1.7, 1.1, 0.9, 2.2, 1.0, 5.5]),
    ("glass", *_glass(8, 3)),
    ("wide", 4, [0, 1, 2, 3, 0], [1, 2, 3, 0, 2], [200.0, -150.0, 3.0, 1.0, 0.25]),
]


@pytest.mark.parametrize("case", REAL, ids=[c[0] for c in REAL])
@pytest.mark.parametrize("kind,q,alg", [(0, 2, 0), (0, 2, 1), (1, 3, 0), (1, 4, 1)])
def test_real_weights_bit_exact(sc, orc, case, kind, q, alg):
    name, n_sites, src, dst, w = case
    print(f"case real {name} kind={kind} q={q} alg={alg}", flush=True)
    lat, olat = sc.Lattice.from_edges(n_sites, src, dst, w), orc.Lattice.from_edges(n_sites, src, dst, w)
    seed, beta, n = 78, 0.9, 8
    st = sc.State(lat, sc.model(kind, 1.0, q), 1, seed)
    st.init(True)
    key = orc.replica_key(seed, 0)
    m = orc.model(kind, 1.0, q)
    start = orc.random_start(olat, m, key)
    assert np.array_equal(st.download_f64(), start)
    chain = orc.Chain(olat, m, start)
    st.set_beta(beta)
    st.sweep(alg, n, measure=True)
    for t in range(n):
        (chain.metropolis if alg == 0 else chain.heat_bath)(beta, key, t)
    assert np.array_equal(st.download_f64(), chain.spins)
    e, mag = st.series()
    assert abs(e[-1] - chain.total_energy()) <= 1e-9 * (1 + abs(chain.total_energy()))
    assert mag[-1] == chain.magnetization()
    e1, m1 = st.observables()
    assert e1 == e[-1] and m1 == mag[-1]


@pytest.mark.parametrize("graph", ["multigraph", "glass"])
@pytest.mark.parametrize("kind,q", [(0, 2), (1, 3)])
@pytest.mark.parametrize("wolff", [False, True])
def test_real_weight_clusters_bit_exact(sc, orc, graph, kind, q, wolff):
    """Swendsen-Wang / Wolff on random-bond ferromagnets (|w| of the real
    weights): one threshold per edge, spins equal the oracle's bit for bit."""
    name, n_sites, src, dst, w = next(c for c in REAL if c[0] == graph)
    w = [abs(x) for x in w]
    print(f"case real clusters {graph} kind={kind} wolff={wolff}", flush=True)
    lat, olat = sc.Lattice.from_edges(n_sites, src, dst, w), orc.Lattice.from_edges(n_sites, src, dst, w)
    seed, steps, betas = 81, 5, [0.3, 0.8]
    st = sc.State(lat, sc.model(kind, 1.0, q), 2, seed)
    for r in range(2):
        st.init(True, r)
        st.set_beta(betas[r], r)
    m = orc.model(kind, 1.0, q)
    start = [st.download_f64(r) for r in range(2)]
    st.sweep(6 if wolff else 5, steps)
    for r in range(2):
        s = start[r].copy()
        orc.cluster_steps(olat, m, s, betas[r], orc.replica_key(seed, r), 0, steps, wolff)
        assert np.array_equal(st.download_f64(r), s), f"replica {r}"


def test_real_weight_clusters_refuse_negative_bonds(sc):
    lat = sc.Lattice.from_edges(3, [0, 1], [1, 2], [0.5, -1.25])
    st = sc.State(lat, sc.model(0), 1, 1)
    with pytest.raises(sc.ScamperError):
        st.sweep(5, 1)


def test_run_one_real_weight_chain(sc):
    """Scheduler::run_one, Ising Metropolis on an open chain with real
    weights (a tree): <E> = -sum w tanh(beta w)."""
    import math
    w = [0.35, 1.7, 0.8, 2.25, 0.6, 1.1, 0.45, 1.9]
    n = len(w) + 1
    beta, meas = 0.7, 40000
    want = -sum(x * math.tanh(beta * x) for x in w)
    res, e, _ = sc.run_one(sc.Lattice.from_edges(n, list(range(n - 1)), list(range(1, n)), w), sc.model(sc.ISING, 1.0),
                           sc.METROPOLIS, beta, True, 500, meas, 100, 17, series=True)
    assert res["Energy"].mean == pytest.approx(e.mean(), rel=1e-12)
    est = sc.binned_estimate(e, meas // 40)
    z = (est.mean - want) / est.stderr
    print(f"run_one real chain: {est.mean:.5f} exact {want:.5f} z={z:.2f}")
    assert abs(z) < 4.5
