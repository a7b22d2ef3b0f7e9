"""The L=64 periodic Ising workload end to end on the device:
Scheduler::run_one (Carlo.rs/src/scheduler.rs:94-155) with cold start,
base seed 42, T = 2.269185, 10^4 + 10^5 sweeps, binsize 100, driven through
scamper_run_one.  Checked against the exact finite-torus energy (Kaufman)
and against the reference's plain path as restated by the oracle
(tests/golden/c1_faithful.json, made by make_c1_golden.py).

Errors for the z-tests come from blocks of 2000 sweeps on both sides (the
binsize-100 errors the estimates report ignore the autocorrelation near
T_c, as the reference's own estimates do)."""
import json
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import exact  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_faithful.json")


def _blocks(x, block):
    b = np.asarray(x).reshape(-1, block).mean(axis=1)
    return b.mean(), b.std(ddof=1) / math.sqrt(b.size), b


def _binder(m2_blocks, m4_blocks):
    """Binder cumulant 1 - <M4>/(3 <M2>^2) (CMC.rs/src/postprocess.rs:70-77)
    with a jackknife error over blocks."""
    n = m2_blocks.size
    full = 1 - m4_blocks.mean() / (3 * m2_blocks.mean() ** 2)
    jk = np.array([1 - np.delete(m4_blocks, i).mean() / (3 * np.delete(m2_blocks, i).mean() ** 2) for i in range(n)])
    return full, math.sqrt((n - 1) / n * ((jk - jk.mean()) ** 2).sum())


def test_run_one_c1(sc):
    g = json.load(open(GOLDEN))
    cfg = g["config"]
    lat = sc.Lattice.square(cfg["L"], cfg["L"], True)
    res, e, m = sc.run_one(lat, sc.model(sc.ISING), sc.METROPOLIS, 1.0 / cfg["T"], False, cfg["therm"],
                           cfg["meas"], cfg["binsize"], cfg["seed"], series=True)
    # estimates are the Accumulator's over the returned series
    assert res["Energy"].n_bins == 1000
    assert res["Energy"].mean == pytest.approx(e.mean(), rel=1e-12)
    block = g["block"]
    n = g["n_sites"]
    exact_e = exact.kaufman_energy(cfg["L"], cfg["L"], 1.0 / cfg["T"])
    e_mean, e_se, _ = _blocks(e, block)
    z_exact = (e_mean - exact_e) / e_se
    print(f"E/N device {e_mean / n:.5f} +- {e_se / n:.5f}, exact {exact_e / n:.5f}, z = {z_exact:.2f}")
    assert abs(z_exact) < 3
    series = {"Energy": e, "E2": e * e, "Magnetization": m, "M2": m * m, "M4": (m * m) * (m * m)}
    blocks = {}
    for name, x in series.items():
        mu, se, blocks[name] = _blocks(x, block)
        gb = np.array(g["block_means"][name])
        z = (mu - gb.mean()) / math.sqrt(se**2 + (gb.std(ddof=1) / math.sqrt(gb.size)) ** 2)
        print(f"{name}: device {mu:.6g} vs faithful CPU {gb.mean():.6g}, z = {z:.2f}")
        assert abs(z) < 3
    u_dev, su_dev = _binder(blocks["M2"], blocks["M4"])
    u_cpu, su_cpu = _binder(np.array(g["block_means"]["M2"]), np.array(g["block_means"]["M4"]))
    z = (u_dev - u_cpu) / math.sqrt(su_dev**2 + su_cpu**2)
    print(f"Binder: device {u_dev:.4f}+-{su_dev:.4f} CPU {u_cpu:.4f}+-{su_cpu:.4f} z = {z:.2f}")
    assert abs(z) < 3
