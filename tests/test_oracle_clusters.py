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

O(3) embedding clusters (ClusterModel for ONModel<D>, models.rs:516-577):
* the direction draw equals an independent Python restatement bit for bit
  and is uniform on the sphere (first and second moments);
* for spins collinear with the direction the bond probability is the
  embedded Ising bond, K = ceil(p 2^32) of the Ising threshold; it is zero
  for opposite projections, for J = 0 and for a spin perpendicular to r;
* beta = 0 gives singleton clusters: Swendsen-Wang reflects each site on
  its own coin, Wolff the seed alone;
* chains reproduce exact energies: the periodic ring from its transfer
  spectrum 4 pi i_l(K) (modified spherical Bessel functions, checked to
  reduce to the Langevin sum of the open chain), and a weighted tree from
  the Langevin function of every bond.
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


# ------------------------------------------------ O(3) embedding clusters
# (ClusterModel for ONModel<D>, models.rs:516-577)

def _word_py(key, idx, step, stream):
    ctr = [(idx >> 2) & 0xFFFFFFFF, step & 0xFFFFFFFF, ((step >> 32) & 0xFFFF) | (stream << 24), (idx >> 34) & 0xFFFFFFFF]
    return _philox_py(ctr, [key & 0xFFFFFFFF, key >> 32])[idx & 3]


def _direction_py(key, step):
    """The rejection draw restated: try k on seed-stream words 4 + 3k ..."""
    for k in range(64):
        x = [2.0 * (_word_py(key, 4 + 3 * k + c, step, 13) * 2.0**-32) - 1.0 for c in range(3)]
        n2 = (x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]
        if 2.0**-40 < n2 <= 1.0:
            n = math.sqrt(n2)
            return [v / n for v in x]
    return [0.0, 0.0, 1.0]


def _direction(orc, key, step):
    r = np.zeros(3)
    orc.lib().orc_on3_direction(key, step, orc._p(r))
    return r


def test_on3_direction_is_the_rejection_draw_and_uniform(orc):
    key = orc.replica_key(7, 1)
    for step in [0, 1, 5, 2**33 + 3]:
        assert list(_direction(orc, key, step)) == _direction_py(key, step)  # bit for bit
    rs = np.array([_direction(orc, key, t) for t in range(20000)])
    assert np.allclose(np.linalg.norm(rs, axis=1), 1.0, atol=2e-16)
    n = len(rs)
    assert np.all(np.abs(rs.mean(axis=0)) < 4.5 * math.sqrt(1 / 3 / n))  # uniform on S^2: <r> = 0
    # <r_c^2> = 1/3 with variance 1/5 - 1/9
    assert np.all(np.abs((rs**2).mean(axis=0) - 1 / 3) < 4.5 * math.sqrt((1 / 5 - 1 / 9) / n))


def _bond_p(orc, j, w, beta, a, b, r):
    a, b, r = (np.ascontiguousarray(v, np.float64) for v in (a, b, r))
    return orc.lib().orc_on3_bond_probability(j, w, beta, orc._p(a), orc._p(b), orc._p(r))


def test_on3_bond_probability_reduces_to_the_ising_embedding(orc):
    rng = np.random.default_rng(3)
    for _ in range(50):
        r = rng.normal(size=3)
        r /= np.linalg.norm(r)
        beta, w = rng.uniform(0.05, 2.0), rng.uniform(0.2, 3.0)
        # spins collinear with r: the embedded Ising bond, 1 - exp(-2 beta J w)
        p = _bond_p(orc, 1.0, w, beta, r, r, r)
        K = orc.lib().orc_cluster_threshold(orc.ISING, 1.0, w, beta)
        assert K == math.ceil(p * 2**32)
        assert _bond_p(orc, 1.0, w, beta, r, -r, r) == 0.0  # anti-aligned projections: no bond
        a, b = rng.normal(size=3), rng.normal(size=3)
        a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
        assert _bond_p(orc, 1.0, w, beta, a, b, r) == _bond_p(orc, 1.0, w, beta, b, a, r)
        assert _bond_p(orc, 0.0, w, beta, a, b, r) == 0.0
        # a spin perpendicular to r never bonds
        perp = np.cross(r, a)
        perp /= np.linalg.norm(perp)
        assert _bond_p(orc, 1.0, w, beta, perp, b, r) <= 1e-15


def _reflect_py(s, r):
    proj = (s[0] * r[0] + s[1] * r[1]) + s[2] * r[2]
    v = [s[c] - 2.0 * proj * r[c] for c in range(3)]
    n = math.sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2])
    return [float(np.float32(x / n)) for x in v]


def test_on3_beta_zero(orc):
    """No bond at beta = 0: Swendsen-Wang reflects each site on its own coin,
    Wolff reflects the seed alone; a reflection keeps s.r up to its sign."""
    lat = orc.Lattice.hypercubic([4, 4, 2], True)
    m = orc.model(orc.ON, 1.0, 2, 3)
    key = orc.replica_key(2, 0)
    s = orc.random_start(lat, m, key).astype(np.float32).astype(np.float64)
    r = _direction_py(key, 3)
    sw = s.copy()
    assert orc.cluster_steps(lat, m, sw, 0.0, key, 3, 1, False) == lat.n_sites
    for site in range(lat.n_sites):
        old = s[3 * site:3 * site + 3]
        if _word_py(key, site, 3, 12) < 2**31:
            assert list(sw[3 * site:3 * site + 3]) == _reflect_py(old, r)
            assert abs(np.dot(sw[3 * site:3 * site + 3], r) + np.dot(old, r)) < 1e-6
        else:
            assert np.array_equal(sw[3 * site:3 * site + 3], old)
    wf = s.copy()
    assert orc.cluster_steps(lat, m, wf, 0.0, key, 3, 1, True) == 1
    seed = (_word_py(key, 0, 3, 13) * lat.n_sites) >> 32
    moved = np.nonzero(np.any((wf != s).reshape(-1, 3), axis=1))[0]
    assert list(moved) == [seed]


def _ring_energy(n, K):
    """<E>/J of the periodic classical Heisenberg ring of n sites at K = beta
    J: the transfer kernel exp(K s.s') has eigenvalues 4 pi i_l(K) (2l + 1
    times), so Z = sum (2l+1) (4 pi i_l)^n and <E> = -d ln Z / dK."""
    from scipy.special import spherical_in
    ls = np.arange(60)
    il = spherical_in(ls, K)
    dil = spherical_in(ls, K, derivative=True)
    return -n * np.sum((2 * ls + 1) * il ** (n - 1) * dil) / np.sum((2 * ls + 1) * il**n)


def _langevin(x):
    return 1.0 / math.tanh(x) - 1.0 / x


def _on3_chain_energy(orc, lat, beta, wolff, seeds, steps, burn=500):
    m = orc.model(orc.ON, 1.0, 2, 3)
    out = []
    for seed in seeds:
        key = orc.replica_key(seed, 0)
        s = orc.random_start(lat, m, key)
        es = []
        for t in range(steps):
            orc.cluster_steps(lat, m, s, beta, key, t, 1, wolff)
            if t >= burn:
                es.append(orc.lib().orc_total_energy(orc.C.byref(m), lat.ptr, orc._p(s)))
        out.append(np.array(es))
    return out


@pytest.mark.parametrize("wolff", [False, True])
def test_on3_ring_samples_boltzmann(orc, wolff):
    n, beta = 4, 0.9
    want = _ring_energy(n, beta)
    # the transfer-matrix sum reduces to the open chain's Langevin sum as n grows
    assert abs(_ring_energy(40, beta) / 40 + _langevin(beta)) < 1e-6
    chains = _on3_chain_energy(orc, orc.Lattice.hypercubic([n], True), beta, wolff, range(4), 12000)
    zs = [(orc.binned_estimate(e, 200).mean - want) / orc.binned_estimate(e, 200).stderr for e in chains]
    assert np.all(np.abs(zs) < 4.5), (want, zs)


@pytest.mark.parametrize("wolff", [False, True])
def test_on3_weighted_tree_samples_boltzmann(orc, wolff):
    """On a tree every bond is independent: <E> = -sum_b J w_b L(beta J w_b)."""
    src, dst, w = [0, 1, 1, 3], [1, 2, 3, 4], [1.0, 2.0, 0.5, 1.5]
    beta = 0.7
    lat = orc.Lattice.from_edges(5, src, dst, w)
    want = -sum(wb * _langevin(beta * wb) for wb in w)
    chains = _on3_chain_energy(orc, lat, beta, wolff, range(4), 12000)
    zs = [(orc.binned_estimate(e, 200).mean - want) / orc.binned_estimate(e, 200).stderr for e in chains]
    assert np.all(np.abs(zs) < 4.5), (want, zs)


# ------------------------------------------------------- O(2) embedding clusters

def _direction2_py(key, step):
    """The rejection draw from the square: try k on seed-stream words 4 + 2k, 5 + 2k."""
    for k in range(64):
        x = [2.0 * (_word_py(key, 4 + 2 * k + c, step, 13) * 2.0**-32) - 1.0 for c in range(2)]
        n2 = x[0] * x[0] + x[1] * x[1]
        if 2.0**-40 < n2 <= 1.0:
            n = math.sqrt(n2)
            return [v / n for v in x]
    return [0.0, 1.0]


def test_on2_direction_is_the_rejection_draw_and_uniform(orc):
    key = orc.replica_key(7, 2)

    def draw(step):
        r = np.zeros(2)
        orc.lib().orc_on_direction(2, key, step, orc._p(r))
        return r

    for step in [0, 1, 5, 2**33 + 3]:
        assert list(draw(step)) == _direction2_py(key, step)  # bit for bit
    rs = np.array([draw(t) for t in range(20000)])
    n = len(rs)
    assert np.allclose(np.linalg.norm(rs, axis=1), 1.0, atol=2e-16)
    assert np.all(np.abs(rs.mean(axis=0)) < 4.5 * math.sqrt(0.5 / n))  # uniform on S^1: <r> = 0
    # <cos^2> = 1/2 with variance 3/8 - 1/4; <cos 4 phi> = 0 (no square anisotropy)
    assert np.all(np.abs((rs**2).mean(axis=0) - 0.5) < 4.5 * math.sqrt(1 / 8 / n))
    c4 = np.cos(4 * np.arctan2(rs[:, 1], rs[:, 0]))
    assert abs(c4.mean()) < 4.5 * math.sqrt(0.5 / n)
    # the three-component draw is unchanged by the generalisation
    r3 = np.zeros(3)
    orc.lib().orc_on_direction(3, key, 9, orc._p(r3))
    assert list(r3) == _direction_py(key, 9)


def test_on2_bond_probability_reduces_to_the_ising_embedding(orc):
    rng = np.random.default_rng(5)
    for _ in range(50):
        phi = rng.uniform(0, 2 * math.pi)
        r = np.array([math.cos(phi), math.sin(phi)])
        beta, w = rng.uniform(0.05, 2.0), rng.uniform(0.2, 3.0)
        bp = lambda a, b, j=1.0: orc.lib().orc_on_bond_probability(  # noqa: E731
            2, j, w, beta, orc._p(np.ascontiguousarray(a)), orc._p(np.ascontiguousarray(b)), orc._p(r))
        assert orc.lib().orc_cluster_threshold(orc.ISING, 1.0, w, beta) == math.ceil(bp(r, r) * 2**32)
        assert bp(r, -r) == 0.0
        perp = np.array([-r[1], r[0]])
        assert bp(perp, r) <= 1e-15
        a, b = rng.normal(size=2), rng.normal(size=2)
        a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
        assert bp(a, b) == bp(b, a) and bp(a, b, 0.0) == 0.0
        # the product of projections: p = 1 - exp(-2 beta w (a.r)(b.r))
        pa, pb = a @ r, b @ r
        want = 1 - math.exp(-2 * beta * w * pa * pb) if pa * pb > 0 else 0.0
        assert abs(bp(a, b) - want) < 1e-14


def _xy_ring_energy(n, k):
    from scipy.special import iv
    m = np.arange(-60, 61)
    i, di = iv(m, k), 0.5 * (iv(m - 1, k) + iv(m + 1, k))
    return -n * np.sum(i ** (n - 1) * di) / np.sum(i**n)


def _on2_chain_energy(orc, lat, beta, wolff, seeds, steps, burn=500):
    m = orc.model(orc.ON, 1.0, 2, 2)
    out = []
    for seed in seeds:
        key = orc.replica_key(seed, 0)
        s = orc.random_start(lat, m, key)
        es = []
        for t in range(steps):
            orc.cluster_steps(lat, m, s, beta, key, t, 1, wolff)
            if t >= burn:
                es.append(orc.lib().orc_total_energy(orc.C.byref(m), lat.ptr, orc._p(s)))
        out.append(np.array(es))
    return out


@pytest.mark.parametrize("wolff", [False, True])
def test_on2_ring_samples_boltzmann(orc, wolff):
    n, beta = 5, 1.0
    want = _xy_ring_energy(n, beta)
    chains = _on2_chain_energy(orc, orc.Lattice.hypercubic([n], True), beta, wolff, range(4), 12000)
    zs = [(orc.binned_estimate(e, 200).mean - want) / orc.binned_estimate(e, 200).stderr for e in chains]
    assert np.all(np.abs(zs) < 4.5), (want, zs)


@pytest.mark.parametrize("wolff", [False, True])
def test_on2_weighted_tree_samples_boltzmann(orc, wolff):
    """On a tree every bond is independent: <E> = -sum_b J w_b I1/I0(beta J w_b)."""
    from scipy.special import iv
    src, dst, w = [0, 1, 1, 3], [1, 2, 3, 4], [1.0, 2.0, 0.5, 1.5]
    beta = 0.7
    lat = orc.Lattice.from_edges(5, src, dst, w)
    want = -sum(wb * iv(1, beta * wb) / iv(0, beta * wb) for wb in w)
    chains = _on2_chain_energy(orc, lat, beta, wolff, range(4), 12000)
    zs = [(orc.binned_estimate(e, 200).mean - want) / orc.binned_estimate(e, 200).stderr for e in chains]
    assert np.all(np.abs(zs) < 4.5), (want, zs)
