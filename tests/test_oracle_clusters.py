"""Pins of the oracle's cluster updates (orc_cluster_steps: Swendsen-Wang
and Wolff for Ising and Potts, CMC.rs/src/algorithms/swendsen_wang.rs:73-135,
wolff.rs:72-134, ClusterModel rules in lattice/models.rs:113-156, 295-338).

* the draws are Philox words: checked against an independent Python Philox;
* the bond threshold K = ceil(p 2^32) is the exact boundary of u < p;
* beta = 0 (no bond can be active): Swendsen-Wang sets every site from its
  own target draw, Wolff moves the seed site alone;
* p = 1 (every aligned bond active): the clusters are the aligned domains,
  counted here with scipy's connected components;
* both updates sample the Boltzmann distribution: <E> and <m^2> from long
  chains on small lattices agree with exact enumeration (Ising on a torus
  and on a weighted graph with a self-loop and a parallel bond, Potts q=3).
"""
import math
import os
import sys

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import exact  # noqa: E402
from test_oracle_rng import _philox_py  # noqa: E402


def test_cluster_word_is_philox(orc):
    key = orc.replica_key(5, 2)
    for idx, step, stream in [(0, 0, 11), (7, 3, 12), (2**34 + 5, 9, 11), (123456789, 2**33 + 1, 13)]:
        ctr = [(idx >> 2) & 0xFFFFFFFF, step & 0xFFFFFFFF,
               ((step >> 32) & 0xFFFF) | (stream << 24), (idx >> 34) & 0xFFFFFFFF]
        out = _philox_py(ctr, [key & 0xFFFFFFFF, key >> 32])
        assert orc.lib().orc_cluster_word(key, idx, step, stream) == out[idx & 3]


@pytest.mark.parametrize("kind,beta,w", [(0, 0.44, 1.0), (0, 0.1, 2.0), (1, 1.005, 1.0), (1, 0.3, 3.0),
                                         (0, 0.0, 1.0), (0, 400.0, 1.0), (1, 1e-9, 1.0)])
def test_threshold_is_the_exact_boundary(orc, kind, beta, w):
    K = orc.lib().orc_cluster_threshold(kind, 1.0, w, beta)
    p = 1.0 - math.exp((-2.0 if kind == 0 else -1.0) * beta * w)
    p = min(max(p, 0.0), 1.0)
    assert 0 <= K <= 2**32
    if K > 0:
        assert (K - 1) * 2.0**-32 < p
    if K < 2**32:
        assert not (K * 2.0**-32 < p)
    assert orc.lib().orc_cluster_threshold(kind, 0.0, w, beta) == 0  # no coupling, no bond


def _ising_start(orc, lat, seed):
    m = orc.model(orc.ISING)
    return m, orc.random_start(lat, m, orc.replica_key(seed, 0))


def test_beta_zero(orc):
    lat = orc.Lattice.square(6, 4)
    m, s = _ising_start(orc, lat, 1)
    key = orc.replica_key(1, 0)
    sw = s.copy()
    n = orc.cluster_steps(lat, m, sw, 0.0, key, 5, 1, False)
    assert n == lat.n_sites  # every site its own cluster
    want = [1.0 if orc.lib().orc_cluster_word(key, i, 5, 12) < 2**31 else -1.0 for i in range(lat.n_sites)]
    assert np.array_equal(sw, want)
    wf = s.copy()
    assert orc.cluster_steps(lat, m, wf, 0.0, key, 5, 1, True) == 1
    seed = (orc.lib().orc_cluster_word(key, 0, 5, 13) * lat.n_sites) >> 32
    flipped = np.nonzero(wf != s)[0]
    assert list(flipped) == [seed] and wf[seed] == -s[seed]


def _domains(lat, s):
    g = lat.csr()
    same = s[g["src"]] == s[g["dst"]]
    a = coo_matrix((np.ones(same.sum()), (g["src"][same], g["dst"][same])), shape=(lat.n_sites,) * 2)
    return connected_components(a, directed=False)


@pytest.mark.parametrize("potts", [False, True])
def test_full_bonds_give_aligned_domains(orc, potts):
    lat = orc.Lattice.hypercubic([6, 5, 4], True)
    m = orc.model(orc.POTTS, 1.0, 3) if potts else orc.model(orc.ISING)
    key = orc.replica_key(3, 0)
    for step in range(4):
        s = orc.random_start(lat, m, orc.replica_key(10 + step, 0))
        n_dom, label = _domains(lat, s)
        sw = s.copy()
        assert orc.cluster_steps(lat, m, sw, 1e4, key, step, 1, False) == n_dom
        for d in range(n_dom):  # every domain takes one state
            assert len(set(sw[label == d])) == 1
        wf = s.copy()
        size = orc.cluster_steps(lat, m, wf, 1e4, key, step, 1, True)
        seed = (orc.lib().orc_cluster_word(key, 0, step, 13) * lat.n_sites) >> 32
        dom = label == label[seed]
        assert size == dom.sum() and np.all(wf[~dom] == s[~dom]) and len(set(wf[dom])) == 1
        assert wf[seed] != s[seed]


def _chain_stats(orc, lat, m, beta, wolff, seeds, steps, burn=500):
    """z-scores of <E> and <m^2> from cluster chains against `want`."""
    out = []
    for seed in seeds:
        key = orc.replica_key(seed, 0)
        s = orc.random_start(lat, m, key)
        es, ms = [], []
        for t in range(steps):
            orc.cluster_steps(lat, m, s, beta, key, t, 1, wolff)
            if t >= burn:
                es.append(orc.lib().orc_total_energy(orc.C.byref(m), lat.ptr, orc._p(s)))
                ms.append(orc.lib().orc_magnetization(orc.C.byref(m), orc._p(s), lat.n_sites) ** 2)
        out.append((np.array(es), np.array(ms)))
    return out


def _zs(orc, chains, want_e, want_m2, binsize=200):
    zs = []
    for es, ms in chains:
        for series, want in ((es, want_e), (ms, want_m2)):
            est = orc.binned_estimate(series, binsize)
            zs.append((est.mean - want) / est.stderr)
    return np.array(zs)


@pytest.mark.parametrize("wolff", [False, True])
def test_ising_torus_samples_boltzmann(orc, wolff):
    L, beta = 4, 0.44
    src, dst = [], []
    for y in range(L):
        for x in range(L):
            src += [x + L * y, x + L * y]
            dst += [(x + 1) % L + L * y, x + L * ((y + 1) % L)]
    ex = exact.enumerate_ising(np.array(src), np.array(dst), L * L, beta)
    lat = orc.Lattice.square(L, L)
    chains = _chain_stats(orc, lat, orc.model(orc.ISING), beta, wolff, range(6), 12000)
    zs = _zs(orc, chains, ex["E"], ex["M2"])
    assert np.all(np.abs(zs) < 4.5), zs


@pytest.mark.parametrize("wolff", [False, True])
def test_ising_weighted_graph_samples_boltzmann(orc, wolff):
    # a self-loop (site 2), a parallel bond (0-1 twice) and weights 1, 2, 0.5
    src = [0, 0, 1, 2, 2, 3, 4, 0]
    dst = [1, 1, 2, 3, 2, 4, 0, 3]
    w = [1.0, 0.5, 2.0, 1.0, 1.0, 0.5, 1.0, 2.0]
    n, beta = 5, 0.4
    lat = orc.Lattice.from_edges(n, src, dst, w)
    states = ((np.arange(2**n)[:, None] >> np.arange(n)[None, :]) & 1) * 2 - 1
    e = -(np.array(w) * states[:, src] * states[:, dst]).sum(axis=1)
    m2 = (states.sum(axis=1) / n) ** 2
    p = np.exp(-beta * (e - e.min()))
    p /= p.sum()
    chains = _chain_stats(orc, lat, orc.model(orc.ISING), beta, wolff, range(4), 20000)
    zs = _zs(orc, chains, float((p * e).sum()), float((p * m2).sum()))
    assert np.all(np.abs(zs) < 4.5), zs


@pytest.mark.parametrize("wolff", [False, True])
def test_potts_samples_boltzmann(orc, wolff):
    L, q, beta = 3, 3, 0.8
    src, dst = [], []
    for y in range(L):
        for x in range(L):
            src += [x + L * y, x + L * y]
            dst += [(x + 1) % L + L * y, x + L * ((y + 1) % L)]
    ex = exact.enumerate_potts(np.array(src), np.array(dst), L * L, q, beta)
    lat = orc.Lattice.square(L, L)
    chains = _chain_stats(orc, lat, orc.model(orc.POTTS, 1.0, q), beta, wolff, range(4), 15000)
    zs = _zs(orc, chains, ex["E"], ex["M2"])
    assert np.all(np.abs(zs) < 4.5), zs
