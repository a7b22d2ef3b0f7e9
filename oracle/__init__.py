"""ctypes loader for the oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: the CPU restatement of the reference's classical
lattice sweep that tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg use as the checker.  The product library never imports it.

Parity status per function is recorded in oracle.c's comments and in the
tests that pin each function (tests/test_oracle_*.py).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

ISING, POTTS, ON = 0, 1, 2


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


class Estimate(C.Structure):
    _fields_ = [("mean", C.c_double), ("stderr", C.c_double),
                ("autocorr_time", C.c_double), ("n_bins", C.c_uint64)]


class Model(C.Structure):
    _fields_ = [("kind", C.c_int), ("j", C.c_double), ("q", C.c_int), ("spin_dim", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


u64, u32, i32, dbl, vp = C.c_uint64, C.c_uint32, C.c_int, C.c_double, C.c_void_p


def _declare(L: C.CDLL) -> None:
    sig = {
        "orc_stream_seed": (u64, [u64] * 8),
        "orc_philox4x32_10": (None, [vp, vp, vp]),
        "orc_replica_key": (u64, [u64, u64]),
        "orc_exchange_key": (u64, [u64]),
        "orc_byte": (u32, [u64, u64, u64, u32, u32, u32]),
        "orc_low": (u32, [u64, u64, u64, u32, u32]),
        "orc_word": (u32, [u64, u64, u64, u32, u32]),
        "orc_cont_words": (None, [u64, u64, u64, u32, vp]),
        "orc_cont_init_words": (None, [u64, u64, vp]),
        "orc_exchange_word": (u32, [u64, u64, u64]),
        "orc_accept_log_probability": (i32, [dbl, dbl]),
        "orc_log_acceptance": (dbl, [dbl, dbl]),
        "orc_lattice_from_edges": (vp, [u64, vp, vp, vp, vp, u64]),
        "orc_lattice_hypercubic": (vp, [vp, i32, i32]),
        "orc_lattice_triangular": (vp, [u64, u64]),
        "orc_lattice_honeycomb": (vp, [u64, u64]),
        "orc_lattice_kagome": (vp, [u64, u64]),
        "orc_lattice_free": (None, [vp]),
        "orc_lattice_validate": (i32, [vp]),
        "orc_lattice_get": (None, [vp, vp, vp, vp, vp, vp, vp]),
        "orc_lattice_n_colours": (u32, [vp]),
        "orc_lattice_colours": (None, [vp, vp, vp]),
        "orc_local_energy": (dbl, [vp, vp, vp, u64, vp]),
        "orc_delta_energy": (dbl, [vp, vp, vp, u64, vp]),
        "orc_total_energy": (dbl, [vp, vp, vp]),
        "orc_magnetization": (dbl, [vp, vp, u64]),
        "orc_sweep_metropolis": (u64, [vp, vp, vp, vp, dbl, u64, u64]),
        "orc_sweep_heat_bath": (None, [vp, vp, vp, vp, dbl, u64, u64]),
        "orc_sweep_on3_heat_bath": (None, [vp, vp, vp, vp, dbl, u64, u64]),
        "orc_sweep_on2_heat_bath": (None, [vp, vp, vp, vp, dbl, u64, u64]),
        "orc_von_mises": (dbl, [u64, u64, u64, u32, dbl]),
        "orc_vm_block": (None, [u64, u64, u64, u32, u32, vp]),
        "orc_sweep_over_relaxation": (None, [vp, vp, vp, vp]),
        "orc_random_start": (None, [vp, vp, u64, vp]),
        "orc_ising2d_i8_sweeps": (u64, [vp, u64, u64, dbl, dbl, u64, u64, u64, i32]),
        "orc_stencil_i8_sweeps": (u64, [vp, vp, i32, i32, i32, dbl, dbl, i32, u64, u64, u64, i32]),
        "orc_multispin_sweep": (None, [vp, vp, vp, dbl, u64, u64, u64]),
        "orc_multispin_energy": (dbl, [vp, vp, i32, dbl]),
        "orc_pt_exchange": (i32, [i32, vp, vp, vp, u64, u64]),
        "orc_binned_estimate": (Estimate, [vp, u64, u64]),
        "orc_run_one_ising": (u64, [vp, dbl, dbl, i32, u64, u64, u64, u64, vp, vp, vp]),
        "orc_plain_sweeps": (u64, [vp, dbl, dbl, u64, u64]),
        "orc_plain_chains": (u64, [vp, dbl, dbl, u64, u64, C.c_int, C.c_int]),
        "orc_xoshiro_seq": (None, [u64, u64, vp]),
        "orc_cluster_word": (u32, [u64, u64, u64, u32]),
        "orc_cluster_threshold": (u64, [i32, dbl, dbl, dbl]),
        "orc_cluster_steps": (u64, [vp, vp, vp, dbl, u64, u64, u64, i32]),
        "orc_on3_direction": (None, [u64, u64, vp]),
        "orc_on_direction": (None, [C.c_int, u64, u64, vp]),
        "orc_on_bond_probability": (dbl, [C.c_int, dbl, dbl, dbl, vp, vp, vp]),
        "orc_on3_bond_probability": (dbl, [dbl, dbl, dbl, vp, vp, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


# ------------------------------------------------------------------ helpers

def stream_seed(base, task=0, run=0, chain=0, replica=0, thread=0, phase=0, substream=0) -> int:
    return lib().orc_stream_seed(base, task, run, chain, replica, thread, phase, substream)


def philox(ctr, key) -> list[int]:
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return [int(v) for v in out]


def replica_key(seed: int, replica: int) -> int:
    return lib().orc_replica_key(seed, replica)


def word(key, j, sweep, colour, stream=0) -> int:
    return lib().orc_word(key, j, sweep, colour, stream)


def byte(key, j, sweep, colour, stream=0, part=0) -> int:
    return lib().orc_byte(key, j, sweep, colour, stream, part)


def low(key, j, sweep, colour, stream=0) -> int:
    return lib().orc_low(key, j, sweep, colour, stream)


def accept_log_probability(log_a: float, u: float) -> bool:
    return bool(lib().orc_accept_log_probability(log_a, u))


def binned_estimate(samples, binsize: int) -> Estimate:
    s = np.ascontiguousarray(samples, dtype=np.float64)
    return lib().orc_binned_estimate(_p(s), s.size, binsize)


class Lattice:
    """Owns an orc_lattice*."""

    def __init__(self, ptr: int):
        if not ptr:
            raise ValueError("invalid lattice")
        self.ptr = ptr
        L = lib()
        # orc_lattice struct: n_sites, n_edges, n_bonds first three u64
        hdr = (C.c_uint64 * 3).from_address(ptr)
        self.n_sites, self.n_edges, self.n_bonds = int(hdr[0]), int(hdr[1]), int(hdr[2])
        self.n_colours = int(L.orc_lattice_n_colours(ptr))

    def __del__(self):
        try:
            lib().orc_lattice_free(self.ptr)
        except Exception:
            pass

    @staticmethod
    def hypercubic(dims, pbc=True) -> "Lattice":
        d = np.asarray(dims, dtype=np.uint64)
        return Lattice(lib().orc_lattice_hypercubic(_p(d), d.size, int(pbc)))

    @staticmethod
    def square(lx, ly, pbc=True) -> "Lattice":
        return Lattice.hypercubic([lx, ly], pbc)

    @staticmethod
    def chain(n, pbc=True) -> "Lattice":
        return Lattice.hypercubic([n], pbc)

    @staticmethod
    def triangular(lx, ly) -> "Lattice":
        return Lattice(lib().orc_lattice_triangular(lx, ly))

    @staticmethod
    def honeycomb(lx, ly) -> "Lattice":
        return Lattice(lib().orc_lattice_honeycomb(lx, ly))

    @staticmethod
    def kagome(lx, ly) -> "Lattice":
        return Lattice(lib().orc_lattice_kagome(lx, ly))

    @staticmethod
    def from_edges(n_sites, src, dst, weight=None) -> "Lattice":
        s = np.ascontiguousarray(src, dtype=np.uint64)
        t = np.ascontiguousarray(dst, dtype=np.uint64)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        return Lattice(lib().orc_lattice_from_edges(n_sites, _p(s), _p(t),
                                                    _p(w) if w is not None else None,
                                                    None, s.size))

    def csr(self):
        off = np.zeros(self.n_sites + 1, np.uint64)
        nb = np.zeros(max(self.n_bonds, 1), np.uint64)
        eid = np.zeros(max(self.n_bonds, 1), np.uint64)
        src = np.zeros(max(self.n_edges, 1), np.uint64)
        dst = np.zeros(max(self.n_edges, 1), np.uint64)
        w = np.zeros(max(self.n_edges, 1), np.float64)
        lib().orc_lattice_get(self.ptr, _p(off), _p(nb), _p(eid), _p(src), _p(dst), _p(w))
        return dict(offsets=off, neighbors=nb[:self.n_bonds], edge_ids=eid[:self.n_bonds],
                    src=src[:self.n_edges], dst=dst[:self.n_edges], weight=w[:self.n_edges])

    def colours(self):
        col = np.zeros(self.n_sites, np.uint32)
        idx = np.zeros(self.n_sites, np.uint64)
        lib().orc_lattice_colours(self.ptr, _p(col), _p(idx))
        return col, idx

    def degree(self, site) -> int:
        off = self.csr()["offsets"]
        return int(off[site + 1] - off[site])

    def validate(self) -> bool:
        return bool(lib().orc_lattice_validate(self.ptr))


def model(kind: int, j: float = 1.0, q: int = 2, spin_dim: int = 1) -> Model:
    return Model(kind, j, q, spin_dim)


class Chain:
    """One f64 configuration in the reference's layout with its cached energy."""

    def __init__(self, lat: Lattice, m: Model, spins: np.ndarray):
        self.lat, self.m = lat, m
        self.spins = np.ascontiguousarray(spins, dtype=np.float64).copy()
        self.energy = C.c_double(self.total_energy())

    def total_energy(self) -> float:
        return lib().orc_total_energy(C.byref(self.m), self.lat.ptr, _p(self.spins))

    def magnetization(self) -> float:
        return lib().orc_magnetization(C.byref(self.m), _p(self.spins), self.lat.n_sites)

    def delta_energy(self, site, proposed) -> float:
        p = np.ascontiguousarray(np.atleast_1d(proposed), dtype=np.float64)
        return lib().orc_delta_energy(C.byref(self.m), self.lat.ptr, _p(self.spins), site, _p(p))

    def metropolis(self, beta, key, sweep) -> int:
        return lib().orc_sweep_metropolis(C.byref(self.m), self.lat.ptr, _p(self.spins),
                                          C.byref(self.energy), beta, key, sweep)

    def heat_bath(self, beta, key, sweep) -> None:
        lib().orc_sweep_heat_bath(C.byref(self.m), self.lat.ptr, _p(self.spins),
                                  C.byref(self.energy), beta, key, sweep)

    def on3_heat_bath(self, beta, key, sweep) -> None:
        lib().orc_sweep_on3_heat_bath(C.byref(self.m), self.lat.ptr, _p(self.spins),
                                      C.byref(self.energy), beta, key, sweep)

    def on2_heat_bath(self, beta, key, sweep) -> None:
        lib().orc_sweep_on2_heat_bath(C.byref(self.m), self.lat.ptr, _p(self.spins),
                                      C.byref(self.energy), beta, key, sweep)

    def over_relaxation(self) -> None:
        lib().orc_sweep_over_relaxation(C.byref(self.m), self.lat.ptr, _p(self.spins),
                                        C.byref(self.energy))


def random_start(lat: Lattice, m: Model, key: int) -> np.ndarray:
    out = np.zeros(lat.n_sites * m.spin_dim, np.float64)
    lib().orc_random_start(C.byref(m), lat.ptr, key, _p(out))
    return out


def cluster_steps(lat: Lattice, m: Model, spins: np.ndarray, beta: float, key: int, step0: int,
                  n: int, wolff: bool) -> int:
    """orc_cluster_steps: n Swendsen-Wang or Wolff updates in place (f64 spins)."""
    assert spins.dtype == np.float64 and spins.flags.c_contiguous
    return lib().orc_cluster_steps(C.byref(m), lat.ptr, _p(spins), beta, key, step0, n, int(wolff))


def stencil_sweeps(spins_i8: np.ndarray, dims, kind, q, beta, j, heat_bath, key, sweep0,
                   n_sweeps, threads=0) -> int:
    assert spins_i8.dtype == np.int8 and spins_i8.flags.c_contiguous
    d = np.asarray(dims, dtype=np.uint64)
    return lib().orc_stencil_i8_sweeps(_p(spins_i8), _p(d), d.size, kind, q, beta, j,
                                       int(heat_bath), key, sweep0, n_sweeps, threads)


def multispin_sweep(lat: Lattice, words: np.ndarray, beta_bits, j, seed, replica0, sweep):
    b = np.ascontiguousarray(beta_bits, dtype=np.float64)
    assert words.dtype == np.uint64 and b.size == 64
    lib().orc_multispin_sweep(lat.ptr, _p(words), _p(b), j, seed, replica0, sweep)


def multispin_energy(lat: Lattice, words: np.ndarray, bit: int, j: float = 1.0) -> float:
    return lib().orc_multispin_energy(lat.ptr, _p(words), bit, j)


def pt_exchange(labels: np.ndarray, energies, betas, seed, step) -> int:
    assert labels.dtype == np.int32
    e = np.ascontiguousarray(energies, dtype=np.float64)
    b = np.ascontiguousarray(betas, dtype=np.float64)
    return lib().orc_pt_exchange(labels.size, _p(labels), _p(e), _p(b), seed, step)


def run_one_ising(lat: Lattice, j, beta, initial_state, base_seed, therm, meas, binsize,
                  want_series=False):
    est = (Estimate * 5)()
    se = np.zeros(meas, np.float64) if want_series else None
    sm = np.zeros(meas, np.float64) if want_series else None
    n = lib().orc_run_one_ising(lat.ptr, j, beta, initial_state, base_seed, therm, meas,
                                binsize, est, _p(se) if se is not None else None,
                                _p(sm) if sm is not None else None)
    names = ["Energy", "E2", "Magnetization", "M2", "M4"]
    out = {k: est[i] for i, k in enumerate(names)}
    return (out, n, se, sm) if want_series else (out, n)


def plain_sweeps(lat: Lattice, j, beta, base_seed, n_sweeps) -> int:
    return lib().orc_plain_sweeps(lat.ptr, j, beta, base_seed, n_sweeps)


def plain_chains(lat: Lattice, j, beta, base_seed, n_sweeps, n_chains, threads) -> int:
    return lib().orc_plain_chains(lat.ptr, j, beta, base_seed, n_sweeps, n_chains, threads)


def xoshiro_seq(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint64)
    lib().orc_xoshiro_seq(seed, n, _p(out))
    return out
