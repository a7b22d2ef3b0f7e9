"""The library's Accumulator/Estimate (host C++) against the reference's
known answers (Carlo.rs/tests/unit/accumulator.rs) and the oracle."""
import numpy as np
import pytest


def test_known_answers(sc):
    e = sc.binned_estimate(np.arange(1, 11), 10)
    assert abs(e.mean - 5.5) < 1e-10
    e = sc.binned_estimate(np.arange(1, 16), 5)
    assert abs(e.mean - 8.0) < 1e-10 and e.stderr > 0 and e.n_bins == 3
    assert sc.binned_estimate(np.arange(1, 16), 10).n_bins == 2
    e = sc.binned_estimate(np.zeros(0), 10)
    assert e.n_bins == 0 and e.mean == 0.0 and e.autocorr_time == 1.0
    e = sc.binned_estimate(np.arange(-3, 1), 3)
    assert abs(e.mean + 1.0) < 1e-10
    with pytest.raises(sc.ScamperError):
        sc.binned_estimate(np.ones(3), 0)


@pytest.mark.parametrize("n,binsize", [(1000, 10), (997, 100), (5, 7), (4096, 1), (12345, 64)])
def test_equals_oracle_bit_for_bit(sc, orc, n, binsize):
    rng = np.random.default_rng(n)
    x = np.cumsum(rng.normal(size=n)) * 0.1 + rng.normal(size=n)
    a = sc.binned_estimate(x, binsize)
    b = orc.binned_estimate(x, binsize)
    assert (a.mean, a.stderr, a.autocorr_time, a.n_bins) == (b.mean, b.stderr, b.autocorr_time, b.n_bins)
