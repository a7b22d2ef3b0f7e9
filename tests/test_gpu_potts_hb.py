"""Potts heat bath with an antiferromagnetic coupling, bit for bit.

PottsModel::heat_bath_sample_site (CMC.rs/src/lattice/models.rs:340-380)
subtracts max_log, the largest log weight, before exponentiating.  For
J < 0 that is the log weight of the smallest neighbour count, not of the
largest: with the wrong reference the weights exceed 1, and at large beta
they overflow.  These cases run J = -1 on a stencil and on a graph lattice
with unit weights (the integer-count kernels), at moderate and at large
beta, against the oracle's restatement with the same draws.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims,q,beta", [((8, 8), 3, 0.8), ((16, 16), 4, 1.3), ((8, 8), 3, 150.0),
                                         ((6, 4, 4), 5, 40.0), ((16, 16), 3, 0.0)])
def test_stencil_antiferro_heat_bath(sc, orc, dims, q, beta):
    seed, n = 17, 5
    print(f"case stencil {dims} q={q} J=-1 beta={beta}", flush=True)
    lat = sc.Lattice.hypercubic(list(dims))
    st = sc.State(lat, sc.model(sc.POTTS, -1.0, q), 1, seed)
    st.init(True)
    start = st.download_i8()
    st.set_beta(beta)
    st.sweep(sc.HEAT_BATH, n, measure=True)
    want = start.copy()
    orc.stencil_sweeps(want, list(dims), orc.POTTS, q, beta, -1.0, True, orc.replica_key(seed, 0), 0, n)
    got = st.download_i8()
    assert np.array_equal(got, want), f"{int(np.count_nonzero(got != want))} sites differ"
    chain = orc.Chain(orc.Lattice.hypercubic(list(dims)), orc.model(orc.POTTS, -1.0, q), want.astype(np.float64))
    e, _ = st.series()
    assert e[-1] == chain.total_energy()


@pytest.mark.parametrize("builder,args,q,beta", [("triangular", (12, 12), 3, 0.9), ("triangular", (12, 12), 3, 120.0),
                                                 ("honeycomb", (8, 8), 4, 2.0)])
def test_graph_antiferro_heat_bath(sc, orc, builder, args, q, beta):
    seed, n = 23, 4
    print(f"case graph {builder}{args} q={q} J=-1 beta={beta}", flush=True)
    lat = getattr(sc.Lattice, builder)(*args)
    olat = getattr(orc.Lattice, builder)(*args)
    st = sc.State(lat, sc.model(sc.POTTS, -1.0, q), 1, seed)
    st.init(True)
    chain = orc.Chain(olat, orc.model(orc.POTTS, -1.0, q), st.download_f64())
    st.set_beta(beta)
    st.sweep(sc.HEAT_BATH, n)
    key = orc.replica_key(seed, 0)
    for t in range(n):
        chain.heat_bath(beta, key, t)
    assert np.array_equal(st.download_f64(), chain.spins)
