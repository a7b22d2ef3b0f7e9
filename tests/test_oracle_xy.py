"""Pins of the oracle's O(2) (XY) heat bath (ContinuousHeatBathable for
ONModel<2>, CMC.rs/src/lattice/models.rs:596-614, with the Best-Fisher
sampler sample_von_mises, models.rs:666-689, on counter-based draws).

* orc_von_mises equals an independent Python restatement of the rejection
  loop on the same Philox words, bit for bit (Python's math is the same libm);
* its samples have the von Mises moments <cos t> = I1/I0, <cos 2t> = I2/I0,
  <sin t> = 0;
* heat-bath chains on a periodic XY ring reproduce the exact energy from the
  transfer spectrum I_m(K) (m in Z): Z = sum_m I_m(K)^N.
"""
import math
import os
import sys

import numpy as np
import pytest
from scipy.special import iv

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle_rng import _philox_py  # noqa: E402


def _vm_block(key, j, sweep, colour, k):
    ctr = [j & 0xFFFFFFFF, sweep & 0xFFFFFFFF, ((sweep >> 32) & 0xFFFF) | (colour << 16) | (14 << 24),
           ((j >> 32) & 0xFFFFFFFF) | (k << 20)]
    return _philox_py(ctr, [key & 0xFFFFFFFF, key >> 32])


def _von_mises_py(key, j, sweep, colour, kappa):
    if kappa < 1e-8:
        return _vm_block(key, j, sweep, colour, 0)[0] * 2.0**-32 * (2 * math.pi)
    tau = 1.0 + math.sqrt(1.0 + 4.0 * kappa * kappa)
    rho = (tau - math.sqrt(2.0 * tau)) / (2.0 * kappa)
    r = (1.0 + rho * rho) / (2.0 * rho)
    for k in range(64):
        w = _vm_block(key, j, sweep, colour, k)
        z = math.cos(math.pi * (w[0] * 2.0**-32))
        value = (1.0 + r * z) / (r + z)
        c = kappa * (r - value)
        u = max(w[1] * 2.0**-32, 2.2250738585072014e-308)
        if c * (2.0 - c) > u or math.log(c / u) + 1.0 >= c:
            break
    a = math.acos(min(max(value, -1.0), 1.0))
    return a if w[2] < 2**31 else -a


def test_von_mises_is_the_restated_rejection_loop(orc):
    key = orc.replica_key(4, 1)
    for kappa in (1e-9, 0.05, 0.7, 3.0, 40.0):
        for j, sweep, colour in ((0, 0, 0), (17, 5, 1), (2**33 + 3, 2**32 + 9, 0)):
            assert orc.lib().orc_von_mises(key, j, sweep, colour, kappa) == _von_mises_py(key, j, sweep, colour, kappa)


@pytest.mark.parametrize("kappa", [0.3, 1.0, 4.0, 15.0])
def test_von_mises_moments(orc, kappa):
    key = orc.replica_key(9, 0)
    n = 60000
    t = np.array([orc.lib().orc_von_mises(key, j, 3, 1, kappa) for j in range(n)])
    for moment, want in ((np.cos(t), iv(1, kappa) / iv(0, kappa)), (np.cos(2 * t), iv(2, kappa) / iv(0, kappa)),
                         (np.sin(t), 0.0)):
        z = (moment.mean() - want) / (moment.std() / math.sqrt(n))
        assert abs(z) < 4.5, (kappa, moment.mean(), want, z)


def _xy_ring_energy(n, k):
    """<E>/J of the periodic XY ring: transfer eigenvalues I_m(K), m in Z."""
    m = np.arange(-60, 61)
    i = iv(m, k)
    di = 0.5 * (iv(m - 1, k) + iv(m + 1, k))
    return -n * np.sum(i ** (n - 1) * di) / np.sum(i**n)


def test_xy_ring_samples_boltzmann(orc):
    n, beta = 6, 1.2
    want = _xy_ring_energy(n, beta)
    # the ring sum reduces to the open chain's n I1/I0 per bond as n grows
    assert abs(_xy_ring_energy(60, beta) / 60 + iv(1, beta) / iv(0, beta)) < 1e-9
    lat = orc.Lattice.hypercubic([n], True)
    m = orc.model(orc.ON, 1.0, 2, 2)
    zs = []
    for seed in range(4):
        key = orc.replica_key(seed, 0)
        ch = orc.Chain(lat, m, orc.random_start(lat, m, key))
        es = []
        for t in range(12000):
            ch.on2_heat_bath(beta, key, t)
            if t >= 200:
                es.append(ch.total_energy())
        est = orc.binned_estimate(np.array(es), 200)
        zs.append((est.mean - want) / est.stderr)
    assert np.all(np.abs(zs) < 4.5), (want, zs)
