"""scamper_amd -- MI355X (gfx950) classical lattice sweep.

Python access to libscamper_hip.so through its C ABI (include/scamper_hip.h).
The library holds all of the work: HIP kernels, decision tables, lattice
construction and the sweep driver.  This module only marshals arguments and
turns status codes into exceptions; there is no fallback path, and importing
the bindings fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscamper_hip.so")

ISING, POTTS, ON = 0, 1, 2
METROPOLIS, HEAT_BATH, CONT_HEAT_BATH, OVER_RELAXATION, HYBRID_HB_OR = 0, 1, 2, 3, 4
SWENDSEN_WANG, WOLFF = 5, 6

STATUS = {0: "OK", 1: "EINVAL", 2: "ENOMEM", 3: "EDEVICE", 4: "EINVARIANT",
          5: "EUNSUPPORTED", 6: "EBUDGET"}


class ScamperError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS.get(code, code)}: {message}")
        self.code = code


class Estimate(C.Structure):
    _fields_ = [("mean", C.c_double), ("stderr", C.c_double),
                ("autocorr_time", C.c_double), ("n_bins", C.c_uint64)]


class RunConfig(C.Structure):
    _fields_ = [("thermalization_sweeps", C.c_uint64), ("measurement_sweeps", C.c_uint64),
                ("binsize", C.c_uint64), ("base_seed", C.c_uint64)]


class Model(C.Structure):
    _fields_ = [("kind", C.c_int32), ("q", C.c_int32), ("spin_dim", C.c_int32),
                ("reserved", C.c_int32), ("j", C.c_double)]


class PoolConfig(C.Structure):
    _fields_ = [("spin_budget_bytes", C.c_uint64), ("max_jobs", C.c_uint32),
                ("chunk_sweeps", C.c_uint32), ("grid", C.c_uint32), ("use_graphs", C.c_uint32),
                ("int8_only", C.c_uint32), ("reserved", C.c_uint32)]


class JobSpec(C.Structure):
    _fields_ = [("n_dims", C.c_uint32), ("reserved", C.c_uint32), ("dims", C.c_uint64 * 3),
                ("model", Model), ("beta", C.c_double), ("algorithm", C.c_int32),
                ("random_start", C.c_int32), ("thermalization_sweeps", C.c_uint64),
                ("measurement_sweeps", C.c_uint64), ("binsize", C.c_uint64), ("seed", C.c_uint64),
                ("keep_series", C.c_uint32), ("keep_spins", C.c_uint32)]


class PoolStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("queued", C.c_uint64), ("running", C.c_uint64),
                ("done", C.c_uint64), ("spin_bytes_in_use", C.c_uint64), ("site_updates", C.c_uint64),
                ("chunks", C.c_uint64), ("grid", C.c_uint32), ("tuned", C.c_uint32),
                ("lane16_from_mi", C.c_uint32), ("chunk_sweeps", C.c_uint32)]


_lib = None
u64, u32, i32, dbl, vp, sz = C.c_uint64, C.c_uint32, C.c_int, C.c_double, C.c_void_p, C.c_size_t
PP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every function the header declares.
SIGNATURES = {
    "scamper_last_error": (sz, [C.c_char_p, sz]),
    "scamper_version": (C.c_char_p, []),
    "scamper_lattice_hypercubic": (i32, [vp, i32, i32, PP]),
    "scamper_lattice_triangular": (i32, [u64, u64, PP]),
    "scamper_lattice_honeycomb": (i32, [u64, u64, PP]),
    "scamper_lattice_kagome": (i32, [u64, u64, PP]),
    "scamper_lattice_from_edges": (i32, [u64, vp, vp, vp, u64, PP]),
    "scamper_lattice_info": (i32, [vp, vp, vp, vp, vp]),
    "scamper_lattice_csr": (i32, [vp, vp, vp, vp]),
    "scamper_lattice_edges": (i32, [vp, vp, vp, vp, vp]),
    "scamper_lattice_colours": (i32, [vp, vp, vp]),
    "scamper_lattice_free": (None, [vp]),
    "scamper_state_create": (i32, [vp, C.POINTER(Model), u32, u64, PP]),
    "scamper_state_free": (None, [vp]),
    "scamper_state_set_beta": (i32, [vp, u32, dbl]),
    "scamper_state_get_beta": (i32, [vp, u32, vp]),
    "scamper_state_init": (i32, [vp, u32, i32]),
    "scamper_state_upload_f64": (i32, [vp, u32, vp]),
    "scamper_state_download_f64": (i32, [vp, u32, vp]),
    "scamper_state_upload_i8": (i32, [vp, u32, vp]),
    "scamper_state_download_i8": (i32, [vp, u32, vp]),
    "scamper_state_upload_f32": (i32, [vp, u32, vp]),
    "scamper_state_download_f32": (i32, [vp, u32, vp]),
    "scamper_state_sweeps": (i32, [vp, vp]),
    "scamper_state_set_sweeps": (i32, [vp, u64]),
    "scamper_state_set_ladder": (i32, [vp, vp]),
    "scamper_sweep_with_phase": (i32, [vp, i32, u64, i32, i32]),
    "scamper_ising_sweep_noalloc": (i32, [vp, vp, i32, dbl, vp, u32, u64, u32, u64, u64, vp, vp, vp]),
    "scamper_state_exchange": (i32, [vp, u64]),
    "scamper_state_pt_step": (i32, [vp, i32, u64, i32]),
    "scamper_state_pt_graph_launches": (i32, [vp, vp]),
    "scamper_state_labels": (i32, [vp, vp]),
    "scamper_state_exchange_energies": (i32, [vp, vp]),
    "scamper_state_exchange_stats": (i32, [vp, vp, vp, vp]),
    "scamper_sweep": (i32, [vp, i32, u64, i32]),
    "scamper_synchronize": (i32, [vp]),
    "scamper_sweep_timed": (i32, [vp, i32, u64, i32, vp]),
    "scamper_state_observables": (i32, [vp, u32, vp, vp]),
    "scamper_series_length": (i32, [vp, vp]),
    "scamper_series_get": (i32, [vp, u32, vp, vp, u64]),
    "scamper_series_clear": (i32, [vp]),
    "scamper_multispin_create": (i32, [vp, dbl, u32, u64, PP]),
    "scamper_multispin_free": (None, [vp]),
    "scamper_multispin_set_betas": (i32, [vp, vp]),
    "scamper_multispin_init": (i32, [vp, i32]),
    "scamper_multispin_upload": (i32, [vp, vp]),
    "scamper_multispin_download": (i32, [vp, vp]),
    "scamper_multispin_step": (i32, [vp, u64, i32, i32]),
    "scamper_multispin_step_timed": (i32, [vp, u64, i32, i32, vp]),
    "scamper_multispin_labels": (i32, [vp, vp]),
    "scamper_multispin_stats": (i32, [vp, vp, vp, vp]),
    "scamper_multispin_series": (i32, [vp, vp, vp, u64, vp]),
    "scamper_binned_estimate": (i32, [vp, u64, u64, C.POINTER(Estimate)]),
    "scamper_run_one": (i32, [vp, C.POINTER(Model), i32, dbl, i32, C.POINTER(RunConfig),
                              vp, vp, vp]),
    "scamper_pool_create": (i32, [C.POINTER(PoolConfig), PP]),
    "scamper_pool_free": (None, [vp]),
    "scamper_pool_submit": (i32, [vp, C.POINTER(JobSpec), vp]),
    "scamper_pool_step": (i32, [vp, vp]),
    "scamper_pool_run": (i32, [vp]),
    "scamper_pool_completed": (i32, [vp, vp, u64, vp]),
    "scamper_pool_job_state": (i32, [vp, u64, vp, vp]),
    "scamper_pool_result": (i32, [vp, u64, vp, vp, vp, vp]),
    "scamper_pool_release": (i32, [vp, u64]),
    "scamper_pool_stats_get": (i32, [vp, C.POINTER(PoolStats)]),
}


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m scamper_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _sized(a: np.ndarray, n: int, what: str) -> np.ndarray:
    """a, after checking it holds exactly n values: the library reads n values
    from the pointer, so a shorter array must never reach it."""
    if a.size != n:
        raise ValueError(f"expected {n} {what}, got {a.size}")
    return a


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().scamper_last_error(buf, 1024)
    return buf.value.decode()


def check(rc: int) -> None:
    if rc != 0:
        raise ScamperError(rc, last_error())


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


# BondType::as_label in declaration order (CMC.rs/src/lattice/graph.rs:12-50).
BOND_LABELS = ("generic", "chain_x", "square_x", "square_y", "square_z", "cubic_x", "cubic_y",
               "cubic_z", "tri_x", "tri_y", "tri_diag", "honey_x", "honey_y", "kagome")


class Lattice:
    """CsrLattice handle (CMC.rs/src/lattice/graph.rs:112-124)."""

    def __init__(self, handle: int):
        self.h = handle
        n_sites, n_edges, n_bonds, n_col = u64(), u64(), u64(), u32()
        check(lib().scamper_lattice_info(self.h, C.byref(n_sites), C.byref(n_edges),
                                         C.byref(n_bonds), C.byref(n_col)))
        self.n_sites, self.n_edges = n_sites.value, n_edges.value
        self.n_bonds, self.n_colours = n_bonds.value, n_col.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.scamper_lattice_free(self.h)
            self.h = None

    @staticmethod
    def _make(fn, *args) -> "Lattice":
        out = C.c_void_p()
        check(fn(*args, C.byref(out)))
        return Lattice(out.value)

    @classmethod
    def hypercubic(cls, dims, pbc: bool = True) -> "Lattice":
        d = np.ascontiguousarray(dims, dtype=np.uint64)
        lat = cls._make(lib().scamper_lattice_hypercubic, _p(d), d.size, int(pbc))
        lat.dims = tuple(int(x) for x in dims)
        return lat

    @classmethod
    def square(cls, lx: int, ly: int, pbc: bool = True) -> "Lattice":
        return cls.hypercubic([lx, ly], pbc)

    @classmethod
    def cubic(cls, l: int, pbc: bool = True) -> "Lattice":
        return cls.hypercubic([l, l, l], pbc)

    @classmethod
    def chain(cls, n: int, pbc: bool = True) -> "Lattice":
        return cls.hypercubic([n], pbc)

    @classmethod
    def triangular(cls, lx, ly) -> "Lattice":
        return cls._make(lib().scamper_lattice_triangular, lx, ly)

    @classmethod
    def honeycomb(cls, lx, ly) -> "Lattice":
        return cls._make(lib().scamper_lattice_honeycomb, lx, ly)

    @classmethod
    def kagome(cls, lx, ly) -> "Lattice":
        return cls._make(lib().scamper_lattice_kagome, lx, ly)

    @classmethod
    def from_edges(cls, n_sites, src, dst, weight=None) -> "Lattice":
        s = np.ascontiguousarray(src, dtype=np.uint64)
        t = np.ascontiguousarray(dst, dtype=np.uint64)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float64)
        return cls._make(lib().scamper_lattice_from_edges, n_sites, _p(s), _p(t), _p(w), s.size)

    def csr(self):
        off = np.zeros(self.n_sites + 1, np.uint64)
        nb = np.zeros(max(1, self.n_bonds), np.uint64)
        eid = np.zeros(max(1, self.n_bonds), np.uint64)
        check(lib().scamper_lattice_csr(self.h, _p(off), _p(nb), _p(eid)))
        return off, nb[: self.n_bonds], eid[: self.n_bonds]

    def edges(self):
        """Physical edges (CsrLattice::edges, graph.rs:112-184): source, target,
        BondType label (graph.rs:33-50) and weight per edge."""
        n = max(1, self.n_edges)
        src, dst = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
        kind, w = np.zeros(n, np.uint8), np.zeros(n, np.float64)
        check(lib().scamper_lattice_edges(self.h, _p(src), _p(dst), _p(kind), _p(w)))
        k = self.n_edges
        return src[:k], dst[:k], [BOND_LABELS[int(c)] for c in kind[:k]], w[:k]

    def colours(self):
        col = np.zeros(self.n_sites, np.uint32)
        idx = np.zeros(self.n_sites, np.uint64)
        check(lib().scamper_lattice_colours(self.h, _p(col), _p(idx)))
        return col, idx


def model(kind: int, j: float = 1.0, q: int = 2, spin_dim: int = 1) -> Model:
    return Model(kind, q, spin_dim, 0, j)


class State:
    """Device-resident batch of System values (CMC.rs/src/lattice/state.rs:19-28)."""

    def __init__(self, lattice: Lattice, m: Model, n_replicas: int = 1, seed: int = 0):
        out = C.c_void_p()
        check(lib().scamper_state_create(lattice.h, C.byref(m), n_replicas, seed, C.byref(out)))
        self.h = out.value
        self.lattice, self.model, self.n_replicas, self.seed = lattice, m, n_replicas, seed

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.scamper_state_free(self.h)
            self.h = None

    __del__ = close

    def set_beta(self, beta: float, replica: int = 0):
        check(lib().scamper_state_set_beta(self.h, replica, beta))

    def init(self, random: bool, replica: int = 0):
        check(lib().scamper_state_init(self.h, replica, int(random)))

    def upload_i8(self, spins, replica: int = 0):
        a = _sized(np.ascontiguousarray(spins, dtype=np.int8).ravel(), self.lattice.n_sites, "spins")
        check(lib().scamper_state_upload_i8(self.h, replica, _p(a)))

    def download_i8(self, replica: int = 0) -> np.ndarray:
        a = np.zeros(self.lattice.n_sites, np.int8)
        check(lib().scamper_state_download_i8(self.h, replica, _p(a)))
        return a

    def _components(self) -> int:
        return self.lattice.n_sites * max(1, self.model.spin_dim)

    def upload_f64(self, spins, replica: int = 0):
        a = np.ascontiguousarray(spins, dtype=np.float64).ravel()
        if a.size != self._components():
            raise ValueError(f"expected {self._components()} values, got {a.size}")
        check(lib().scamper_state_upload_f64(self.h, replica, _p(a)))

    def download_f64(self, replica: int = 0) -> np.ndarray:
        a = np.zeros(self.lattice.n_sites * max(1, self.model.spin_dim), np.float64)
        check(lib().scamper_state_download_f64(self.h, replica, _p(a)))
        return a

    def upload_f32(self, spins, replica: int = 0):
        """O(N) states: N f32 components per site, site-major."""
        a = np.ascontiguousarray(spins, dtype=np.float32).ravel()
        if a.size != self._components():
            raise ValueError(f"expected {self._components()} values, got {a.size}")
        check(lib().scamper_state_upload_f32(self.h, replica, _p(a)))

    def download_f32(self, replica: int = 0) -> np.ndarray:
        a = np.zeros(self._components(), np.float32)
        check(lib().scamper_state_download_f32(self.h, replica, _p(a)))
        return a

    @property
    def sweeps(self) -> int:
        v = u64()
        check(lib().scamper_state_sweeps(self.h, C.byref(v)))
        return v.value

    def set_sweeps(self, sweeps: int):
        check(lib().scamper_state_set_sweeps(self.h, int(sweeps)))

    def set_ladder(self, betas):
        """Temperature label l gets beta betas[l]; replica r holds label r
        (the parameter ladder of parallel tempering)."""
        b = np.ascontiguousarray(betas, dtype=np.float64)
        if b.size != self.n_replicas:
            raise ValueError(f"expected {self.n_replicas} betas, got {b.size}")
        check(lib().scamper_state_set_ladder(self.h, _p(b)))

    def exchange(self, step: int):
        """One replica-exchange round (PtExchange::try_exchange,
        Carlo.rs/src/parallel_tempering.rs:362-419), queued on the device."""
        check(lib().scamper_state_exchange(self.h, int(step)))

    def pt_step(self, algorithm: int, n_steps: int, measure: bool = False):
        """n_steps of (sweep, exchange round numbered by the sweep count)."""
        check(lib().scamper_state_pt_step(self.h, algorithm, int(n_steps), int(measure)))

    def pt_graph_launches(self) -> int:
        n = u64()
        check(lib().scamper_state_pt_graph_launches(self.h, C.byref(n)))
        return n.value

    def labels(self) -> np.ndarray:
        out = np.zeros(self.n_replicas, np.int32)
        check(lib().scamper_state_labels(self.h, _p(out)))
        return out

    def exchange_energies(self) -> np.ndarray:
        out = np.zeros(self.n_replicas, np.float64)
        check(lib().scamper_state_exchange_energies(self.h, _p(out)))
        return out

    def exchange_stats(self):
        a, t, r = u64(), u64(), u64()
        check(lib().scamper_state_exchange_stats(self.h, C.byref(a), C.byref(t), C.byref(r)))
        return a.value, t.value, r.value

    def get_beta(self, replica: int = 0) -> float:
        b = dbl()
        check(lib().scamper_state_get_beta(self.h, replica, C.byref(b)))
        return b.value

    def save_snapshot(self, replica: int = 0, topology: bool = True) -> dict:
        """ClassicalMC::save_snapshot (CMC.rs/src/classical_mc.rs:65-92) of one
        replica; see scamper_amd.snapshot."""
        from .snapshot import save_snapshot
        return save_snapshot(self, replica, topology)

    def load_snapshot(self, snapshot: dict, replica: int = 0):
        """ClassicalMC::load_snapshot (classical_mc.rs:94-165) into one replica;
        see scamper_amd.snapshot."""
        from .snapshot import load_snapshot
        load_snapshot(self, snapshot, replica)

    def sweep(self, algorithm: int = METROPOLIS, n_sweeps: int = 1, measure: bool = False):
        check(lib().scamper_sweep(self.h, algorithm, n_sweeps, int(measure)))

    def sweep_with_phase(self, algorithm: int, n_sweeps: int, measure: bool, phase: int):
        """Algorithm::sweep_with_phase; phase 0 = Thermalization, 1 = Measurement."""
        check(lib().scamper_sweep_with_phase(self.h, algorithm, n_sweeps, int(measure), phase))

    def sweep_timed(self, algorithm: int = METROPOLIS, n_sweeps: int = 1,
                    measure: bool = False) -> float:
        """Device milliseconds of n_sweeps, from HIP events on the state's stream."""
        ms = C.c_float()
        check(lib().scamper_sweep_timed(self.h, algorithm, n_sweeps, int(measure), C.byref(ms)))
        return ms.value

    def synchronize(self):
        check(lib().scamper_synchronize(self.h))

    def observables(self, replica: int = 0):
        e, m = dbl(), dbl()
        check(lib().scamper_state_observables(self.h, replica, C.byref(e), C.byref(m)))
        return e.value, m.value

    def series(self, replica: int = 0):
        n = u64()
        check(lib().scamper_series_length(self.h, C.byref(n)))
        e = np.zeros(n.value, np.float64)
        m = np.zeros(n.value, np.float64)
        check(lib().scamper_series_get(self.h, replica, _p(e), _p(m), n.value))
        return e, m

    def clear_series(self):
        check(lib().scamper_series_clear(self.h))


OBSERVABLES = ("Energy", "E2", "Magnetization", "M2", "M4")


def binned_estimate(samples, binsize: int) -> Estimate:
    """Accumulator + Estimate (Carlo.rs/src/measurements.rs:60-280)."""
    a = np.ascontiguousarray(samples, dtype=np.float64)
    out = Estimate()
    check(lib().scamper_binned_estimate(_p(a), a.size, binsize, C.byref(out)))
    return out


def run_one(lattice: Lattice, m: Model, algorithm: int, beta: float, random_start: bool,
            thermalization_sweeps: int, measurement_sweeps: int, binsize: int = 100,
            base_seed: int = 0, series: bool = False):
    """Scheduler::run_one (Carlo.rs/src/scheduler.rs:94-155) on the device."""
    cfg = RunConfig(thermalization_sweeps, measurement_sweeps, binsize, base_seed)
    est = (Estimate * 5)()
    e = np.zeros(measurement_sweeps) if series else None
    mm = np.zeros(measurement_sweeps) if series else None
    check(lib().scamper_run_one(lattice.h, C.byref(m), algorithm, beta, int(random_start),
                                C.byref(cfg), est, _p(e), _p(mm)))
    res = {k: est[i] for i, k in enumerate(OBSERVABLES)}
    return (res, e, mm) if series else res


class MultiSpin:
    """64 * n_bundles Ising replicas, one bit each, with device-side replica
    exchange (MultiSpinIsing, CMC.rs/src/multi_spin.rs:28-376)."""

    def __init__(self, lattice: Lattice, j: float = 1.0, n_bundles: int = 1, seed: int = 0):
        out = C.c_void_p()
        check(lib().scamper_multispin_create(lattice.h, j, n_bundles, seed, C.byref(out)))
        self.h = out.value
        self.lattice, self.n = lattice, 64 * n_bundles
        self.n_bundles = n_bundles

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.scamper_multispin_free(self.h)
            self.h = None

    __del__ = close

    def set_betas(self, betas):
        b = _sized(np.ascontiguousarray(betas, dtype=np.float64).ravel(), self.n, "betas")
        check(lib().scamper_multispin_set_betas(self.h, _p(b)))

    def init(self, random: bool):
        check(lib().scamper_multispin_init(self.h, int(random)))

    def upload(self, words):
        w = _sized(np.ascontiguousarray(words, dtype=np.uint64).ravel(), self.n_bundles * self.lattice.n_sites,
                   "site words")
        check(lib().scamper_multispin_upload(self.h, _p(w)))

    def download(self) -> np.ndarray:
        w = np.zeros(self.n_bundles * self.lattice.n_sites, np.uint64)
        check(lib().scamper_multispin_download(self.h, _p(w)))
        return w

    def step(self, n_steps: int = 1, exchange: bool = False, measure: bool = False):
        check(lib().scamper_multispin_step(self.h, n_steps, int(exchange), int(measure)))

    def step_timed(self, n_steps: int = 1, exchange: bool = False, measure: bool = False) -> float:
        ms = C.c_float()
        check(lib().scamper_multispin_step_timed(self.h, n_steps, int(exchange), int(measure), C.byref(ms)))
        return ms.value

    def labels(self) -> np.ndarray:
        a = np.zeros(self.n, np.int32)
        check(lib().scamper_multispin_labels(self.h, _p(a)))
        return a

    def stats(self):
        a, b, c = u64(), u64(), u64()
        check(lib().scamper_multispin_stats(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def series(self):
        n = u64()
        check(lib().scamper_multispin_series(self.h, None, None, 0, C.byref(n)))
        e = np.zeros((n.value, self.n))
        m = np.zeros((n.value, self.n))
        check(lib().scamper_multispin_series(self.h, _p(e), _p(m), n.value, None))
        return e, m


JOB_QUEUED, JOB_RUNNING, JOB_DONE = 0, 1, 2


def job_spec(dims, m: Model, beta: float, thermalization_sweeps: int, measurement_sweeps: int,
             binsize: int = 100, seed: int = 0, random_start: bool = True, keep_series: bool = False,
             keep_spins: bool = False, algorithm: int = METROPOLIS) -> JobSpec:
    dims = [int(d) for d in dims]
    s = JobSpec()
    s.n_dims = len(dims)
    for i, d in enumerate(dims):
        s.dims[i] = d
    s.model = m
    s.beta = beta
    s.algorithm = algorithm
    s.random_start = int(random_start)
    s.thermalization_sweeps = thermalization_sweeps
    s.measurement_sweeps = measurement_sweeps
    s.binsize = binsize
    s.seed = seed
    s.keep_series = int(keep_series)
    s.keep_spins = int(keep_spins)
    return s


class Pool:
    """Continuous-batching job pool: independent run_one jobs (Ising or Potts
    Metropolis on periodic hypercubic lattices) sharing the GPU
    (Scheduler::run_parallel, Carlo.rs/src/scheduler.rs:345-405)."""

    def __init__(self, spin_budget_bytes: int = 1 << 30, max_jobs: int = 256, chunk_sweeps: int = 32,
                 grid: int = 0, use_graphs: bool = True, bit_planes: bool = True):
        cfg = PoolConfig(spin_budget_bytes, max_jobs, chunk_sweeps, grid, int(use_graphs), int(not bit_planes), 0)
        h = C.c_void_p()
        check(lib().scamper_pool_create(C.byref(cfg), C.byref(h)))
        self.h = h.value
        self._specs = {}

    def __del__(self):
        if getattr(self, "h", None):
            lib().scamper_pool_free(self.h)
            self.h = None

    def submit(self, spec: JobSpec) -> int:
        jid = u64()
        check(lib().scamper_pool_submit(self.h, C.byref(spec), C.byref(jid)))
        self._specs[jid.value] = spec
        return jid.value

    def step(self) -> int:
        n = u32()
        check(lib().scamper_pool_step(self.h, C.byref(n)))
        return n.value

    def run(self) -> None:
        check(lib().scamper_pool_run(self.h))

    def completed(self) -> list[int]:
        ids = np.zeros(4096, dtype=np.uint64)
        out = []
        while True:
            n = u64()
            check(lib().scamper_pool_completed(self.h, _p(ids), ids.size, C.byref(n)))
            out.extend(int(x) for x in ids[: n.value])
            if n.value < ids.size:
                return out

    def state(self, job_id: int):
        st, done = C.c_int32(), u64()
        check(lib().scamper_pool_job_state(self.h, job_id, C.byref(st), C.byref(done)))
        return st.value, done.value

    def result(self, job_id: int):
        """(estimates dict, energy series or None, |M| series or None, spins or None)."""
        spec = self._specs[job_id]
        est = (Estimate * 5)()
        n = spec.measurement_sweeps
        e = np.zeros(n) if spec.keep_series else None
        m = np.zeros(n) if spec.keep_series else None
        n_sites = int(np.prod([spec.dims[i] for i in range(spec.n_dims)]))
        sp = np.zeros(n_sites, dtype=np.int8) if spec.keep_spins else None
        check(lib().scamper_pool_result(self.h, job_id, est, _p(e), _p(m), _p(sp)))
        return {k: est[i] for i, k in enumerate(OBSERVABLES)}, e, m, sp

    def release(self, job_id: int) -> None:
        check(lib().scamper_pool_release(self.h, job_id))
        self._specs.pop(job_id, None)

    def stats(self) -> PoolStats:
        out = PoolStats()
        check(lib().scamper_pool_stats_get(self.h, C.byref(out)))
        return out


from . import snapshot  # noqa: E402  (cmc-rs-snapshot-v2 export/import)
from . import params  # noqa: E402  (FromParams / build_lattice_from_params)
