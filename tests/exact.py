"""Exact reference values used by the statistical tests.

- enumeration of all 2^N Ising / q^N Potts states of a small graph
  (the reference's tests/physics/common.rs:51-83, potts_exact.rs:47-113);
- row transfer matrix of the L x M periodic Ising model
  (tests/physics/zscore_extended.rs:60-135);
- Kaufman's closed form for the partition function of the finite periodic
  square lattice (B. Kaufman, Phys. Rev. 76, 1232 (1949); in the form of
  A. E. Ferdinand and M. E. Fisher, Phys. Rev. 185, 832 (1969)), pinned
  against enumeration and the transfer matrix in test_exact.py;
- Onsager's thermodynamic-limit energy (tests/physics/zscore_extended.rs:
  140-171) and spontaneous magnetisation (tests/exact/onsager.rs:82-84).
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def enumerate_ising(src, dst, n, beta, j=1.0):
    """<E>, <E^2>, <|m|>, <m^2>, <m^4> over all 2^n states."""
    states = ((np.arange(2**n)[:, None] >> np.arange(n)[None, :]) & 1) * 2 - 1
    e = -j * (states[:, src] * states[:, dst]).sum(axis=1)
    m = np.abs(states.sum(axis=1)) / n
    w = np.exp(-beta * (e - e.min()))
    z = w.sum()
    avg = lambda x: float((w * x).sum() / z)
    return dict(E=avg(e), E2=avg(e * e), M=avg(m), M2=avg(m * m), M4=avg(m**4))


def enumerate_potts(src, dst, n, q, beta, j=1.0):
    best = None
    es, ms = [], []
    for cfg in itertools.product(range(q), repeat=n):
        c = np.array(cfg)
        e = -j * float((c[src] == c[dst]).sum())
        counts = np.bincount(c, minlength=q)
        m = (q * counts.max() - n) / (n * (q - 1))
        es.append(e)
        ms.append(m)
    e = np.array(es)
    m = np.array(ms)
    w = np.exp(-beta * (e - e.min()))
    z = w.sum()
    avg = lambda x: float((w * x).sum() / z)
    return dict(E=avg(e), E2=avg(e * e), M=avg(m), M2=avg(m * m))


def transfer_matrix_energy(lx, ly, beta, j=1.0):
    """<E> of the lx x ly periodic Ising model from Tr T^ly."""
    n = 1 << lx
    s = ((np.arange(n)[:, None] >> np.arange(lx)[None, :]) & 1) * 2 - 1
    h = (s * np.roll(s, -1, axis=1)).sum(axis=1).astype(float)
    v = s @ s.T

    def lnz(b):
        # symmetric T = exp(b J (v + (h_s + h_s')/2)); Z = sum of eigenvalues^ly
        t = b * j * (v + 0.5 * (h[:, None] + h[None, :]))
        shift = t.max()
        ev = np.linalg.eigvalsh(np.exp(t - shift))
        big = np.abs(ev).max()
        return ly * (shift + math.log(big)) + math.log(np.sum((ev / big) ** ly))

    eps = 1e-5
    return -(lnz(beta + eps) - lnz(beta - eps)) / (2 * eps)


def _log_kaufman(k, m, n):
    """ln Z of the m x n periodic square Ising lattice at coupling K = beta J."""
    def gamma(r):
        if r == 0:
            return 2 * k + math.log(math.tanh(k))
        c = math.cosh(2 * k) / math.tanh(2 * k) - math.cos(math.pi * r / n)
        return math.acosh(c)

    terms = []  # (sign, log|value|) of Z1..Z4
    for odd, use_cosh in [(1, True), (1, False), (0, True), (0, False)]:
        sign, logv = 1.0, 0.0
        for r in range(n):
            g = gamma(2 * r + odd)
            x = 0.5 * m * g
            if use_cosh:
                logv += math.log(2.0) + abs(x) + math.log1p(math.exp(-2 * abs(x)))
            elif x == 0.0:
                sign = 0.0
                break
            else:
                sign *= 1.0 if x > 0 else -1.0
                logv += math.log(2.0) + abs(x) + math.log(-math.expm1(-2 * abs(x)))
        if sign != 0.0:
            terms.append((sign, logv))
    return _finish(k, m, n, terms)


def _finish(k, m, n, terms):
    top = max(t[1] for t in terms)
    total = sum(s * math.exp(lv - top) for s, lv in terms)
    return math.log(0.5) + 0.5 * m * n * math.log(2 * math.sinh(2 * k)) + top + math.log(total)


def kaufman_energy(l, m, beta, j=1.0):
    """<E> (total) of the l x m periodic Ising model."""
    eps = 1e-6 * max(1.0, beta)
    k = beta * j
    d = (_log_kaufman(k + eps * j, l, m) - _log_kaufman(k - eps * j, l, m)) / (2 * eps)
    return -d


def kaufman_specific_heat(l, m, beta, j=1.0):
    """<E^2> - <E>^2 from the second derivative of ln Z."""
    eps = 1e-4
    k = beta * j
    f = lambda x: _log_kaufman(x, l, m)
    d2 = (f(k + eps * j) - 2 * f(k) + f(k - eps * j)) / (eps * eps)
    return d2


def onsager_energy_per_site(beta, j=1.0):
    """Thermodynamic-limit energy per site, -J coth(2K)[1 + 2/pi (2 tanh^2 2K - 1) K1(k)]."""
    k = beta * j
    kk = 2 * math.sinh(2 * k) / math.cosh(2 * k) ** 2
    # complete elliptic integral of the first kind via the AGM
    a, b = 1.0, math.sqrt(1 - kk * kk)
    for _ in range(60):
        a, b = 0.5 * (a + b), math.sqrt(a * b)
    k1 = math.pi / (2 * a)
    return -j / math.tanh(2 * k) * (1 + 2 / math.pi * (2 * math.tanh(2 * k) ** 2 - 1) * k1)


def onsager_magnetisation(beta, j=1.0):
    k = beta * j
    x = 1 - math.sinh(2 * k) ** -4
    return x ** 0.125 if x > 0 else 0.0
