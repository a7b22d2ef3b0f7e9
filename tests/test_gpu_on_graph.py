"""O(N) spins (N = 3 Heisenberg, N = 2 XY) on graph lattices on the GPU:
triangular, honeycomb, kagome, an open square and a weighted graph with a
self-loop, a parallel bond and non-integer weights (any CsrLattice, as
ONModel<D> runs on, lattice/models.rs:384-689).

Parity with the oracle, which sweeps colour classes in order and numbers each
site's draws by its rank in its class (orc_sweep_on3_heat_bath,
orc_sweep_on2_heat_bath, orc_sweep_over_relaxation), from the device's own
FP32 start:
* random start to 2e-6 (O(3), FP32 Box-Muller) and 2e-7 (XY, f64);
* one heat-bath sweep per component to 2e-4 (O(3), the stencil bar) and 1e-4
  (XY; FP32 storage of earlier classes moves atan2(h) by up to 2^-23 per
  neighbour component over |h|), with the median below 1e-6;
* one over-relaxation sweep to 1e-5;
* E (self-loops add -J w) and |M| of the device recount to 1e-5 relative;
* Swendsen-Wang and Wolff updates bit for bit (f64 bond probabilities and
  reflections, rounded to FP32, as on the stencil).
Statistics: a weighted tree's energy, -sum w L(beta w) (O(3)) and
-sum w I1/I0(beta w) (XY), every bond independent.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SRC = [0, 1, 2, 3, 4, 0, 2, 5, 1]
DST = [1, 2, 3, 4, 5, 3, 2, 0, 2]      # 2-2 is a self-loop, 1-2 a parallel bond
W = [1.0, 0.5, 2.25, 1.5, 0.75, 1.25, 0.5, 3.0, 0.25]

LATTICES = {
    "triangular": (lambda sc: sc.Lattice.triangular(6, 6), lambda orc: orc.Lattice.triangular(6, 6)),
    "honeycomb": (lambda sc: sc.Lattice.honeycomb(4, 4), lambda orc: orc.Lattice.honeycomb(4, 4)),
    "kagome": (lambda sc: sc.Lattice.kagome(3, 3), lambda orc: orc.Lattice.kagome(3, 3)),
    "open_square": (lambda sc: sc.Lattice.hypercubic([6, 5], False), lambda orc: orc.Lattice.hypercubic([6, 5], False)),
    "weighted": (lambda sc: sc.Lattice.from_edges(6, SRC, DST, W), lambda orc: orc.Lattice.from_edges(6, SRC, DST, W)),
}


def _pair(sc, orc, name, n_comp, j=1.0, seed=5, n_rep=1, beta=0.8):
    mk, mko = LATTICES[name]
    lat, olat = mk(sc), mko(orc)
    st = sc.State(lat, sc.model(sc.ON, j, 2, n_comp), n_rep, seed)
    for r in range(n_rep):
        st.init(True, r)
        st.set_beta(beta, r)
    return lat, st, olat, orc.model(orc.ON, j, 2, n_comp)


@pytest.mark.parametrize("n_comp", [3, 2])
@pytest.mark.parametrize("name", list(LATTICES))
def test_sweeps_follow_oracle(sc, orc, name, n_comp):
    print(f"case O({n_comp}) on {name}", flush=True)
    j = -1.0 if name == "triangular" else 1.0  # the frustrated antiferromagnet too
    beta, seed = 0.8, 5
    lat, st, olat, m = _pair(sc, orc, name, n_comp, j, seed, beta=beta)
    start = st.download_f64()
    want = orc.random_start(olat, m, orc.replica_key(seed, 0))
    assert np.max(np.abs(start - want)) < (2e-6 if n_comp == 3 else 2e-7)
    e0, m0 = st.observables()
    chain = orc.Chain(olat, m, start.copy())
    assert abs(e0 - chain.total_energy()) < 1e-5 * max(1.0, abs(chain.total_energy()))
    assert abs(m0 - chain.magnetization()) < 1e-5
    key = orc.replica_key(seed, 0)
    st.sweep(2, 1, measure=True)
    (chain.on3_heat_bath if n_comp == 3 else chain.on2_heat_bath)(beta, key, 0)
    err = np.abs(st.download_f64() - chain.spins)
    print(f"  heat bath: max {err.max():.2e} median {np.median(err):.2e}")
    assert err.max() < (2e-4 if n_comp == 3 else 1e-4) and np.median(err) < 1e-6
    e, mag = st.series()
    assert abs(e[-1] - chain.total_energy()) < 1e-4 * max(1.0, abs(chain.total_energy()))
    assert abs(mag[-1] - chain.magnetization()) < 1e-4
    chain = orc.Chain(olat, m, st.download_f64())
    st.sweep(3, 1)
    chain.over_relaxation()
    assert np.max(np.abs(st.download_f64() - chain.spins)) < 1e-5
    e1, _ = st.observables()
    assert abs(e1 - chain.total_energy()) < 1e-5 * max(1.0, abs(chain.total_energy()))


@pytest.mark.parametrize("n_comp", [3, 2])
@pytest.mark.parametrize("name", ["triangular", "weighted", "open_square"])
@pytest.mark.parametrize("wolff", [False, True])
def test_clusters_bit_exact(sc, orc, name, n_comp, wolff):
    print(f"case O({n_comp}) clusters on {name} wolff={wolff}", flush=True)
    seed, steps, betas = 29, 5, [0.4, 1.3]
    lat, st, olat, m = _pair(sc, orc, name, n_comp, 1.0, seed, n_rep=2)
    for r in range(2):
        st.set_beta(betas[r], r)
    start = [st.download_f64(r) for r in range(2)]
    st.sweep(6 if wolff else 5, steps, measure=True)
    for r in range(2):
        s = start[r].copy()
        es = []
        for t in range(steps):
            orc.cluster_steps(olat, m, s, betas[r], orc.replica_key(seed, r), t, 1, wolff)
            es.append(orc.lib().orc_total_energy(orc.C.byref(m), olat.ptr, orc._p(s)))
        got = st.download_f64(r)
        assert np.array_equal(got, s), f"replica {r}: {np.count_nonzero(got != s)} components differ"
        e_dev, _ = st.series(r)
        assert np.allclose(e_dev, es, rtol=1e-5, atol=1e-5)


def test_clusters_refuse_negative_coupling(sc):
    st = sc.State(sc.Lattice.triangular(4, 4), sc.model(sc.ON, -1.0, 2, 3), 1, 1)
    with pytest.raises(sc.ScamperError):
        st.sweep(5, 1)


@pytest.mark.parametrize("n_comp", [3, 2])
def test_weighted_tree_energy(sc, n_comp):
    from scipy.special import iv
    src, dst, w = [0, 1, 1, 3, 3], [1, 2, 3, 4, 5], [1.0, 2.0, 0.5, 1.5, 0.75]
    beta, n_rep, meas = 0.7, 64, 3000
    if n_comp == 3:
        want = -sum(x * (1 / math.tanh(beta * x) - 1 / (beta * x)) for x in w)
    else:
        want = -sum(x * iv(1, beta * x) / iv(0, beta * x) for x in w)
    st = sc.State(sc.Lattice.from_edges(6, src, dst, w), sc.model(sc.ON, 1.0, 2, n_comp), n_rep, 41)
    for r in range(n_rep):
        st.init(True, r)
        st.set_beta(beta, r)
    st.sweep(2, 100)
    st.sweep(2, meas, measure=True)
    means = np.array([st.series(r)[0].mean() for r in range(n_rep)])
    z = (means.mean() - want) / (means.std(ddof=1) / math.sqrt(n_rep))
    print(f"O({n_comp}) tree: <E> {means.mean():.5f} exact {want:.5f} z={z:.2f}")
    assert abs(z) < 4.5


def test_run_one_xy_tree(sc):
    """Scheduler::run_one over an XY state on a weighted tree (hybrid heat
    bath + over-relaxation): the Energy estimate matches -sum w I1/I0(beta w)."""
    from scipy.special import iv
    src, dst, w = [0, 1, 1, 3], [1, 2, 3, 4], [1.0, 2.5, 0.5, 1.25]
    beta, meas = 0.8, 40000
    want = -sum(x * iv(1, beta * x) / iv(0, beta * x) for x in w)
    res, e, _ = sc.run_one(sc.Lattice.from_edges(5, src, dst, w), sc.model(sc.ON, 1.0, 2, 2), sc.HYBRID_HB_OR,
                           beta, True, 500, meas, 100, 13, series=True)
    assert res["Energy"].mean == pytest.approx(e.mean(), rel=1e-12)
    est = sc.binned_estimate(e, meas // 40)
    z = (est.mean - want) / est.stderr
    print(f"run_one XY tree: {est.mean:.5f} exact {want:.5f} z={z:.2f}")
    assert abs(z) < 4.5
