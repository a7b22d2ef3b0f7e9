"""`cmc-rs-snapshot-v2` export and import for device-resident state.

The format is ClassicalMC::save_snapshot / load_snapshot
(CMC.rs/src/classical_mc.rs:65-165, helpers 231-290): a versioned JSON
document holding
    format    "cmc-rs-snapshot-v2"
    spins     site-major f64, spin_dim values per site
    beta, spin_dim, n_sites, n_edges
    offsets, neighbors, edge_ids    the CSR arrays of the lattice
    edges     [{source, target, kind, weight}] with BondType labels.

Loading checks the following, and raises CheckpointCorrupted ("field:
reason") on a failure, leaving the state unchanged:
* the format tag;
* the spins: an array of finite numbers of the right length;
* every topology field present against the current lattice.  A field may be
  absent, as in version-1 documents, and is then not checked.  Edge weights
  are compared bit for bit;
* beta: finite and non-negative.
Only then does it install the spins and beta.  Hamiltonian::
validate_configuration runs as part of the upload.  The energy is never
read from the document: the device state computes it from the spins.

The random-number state of a device replica is (seed, replica, sweeps), so
the document carries it under one extra key, "scamper": {"seed",
"replica", "sweeps"}.  The reference's loader reads only its own keys.
Loading into a state of the same seed and the same replica index restores
the sweep counter, and the replica then continues the original trajectory
bit for bit, as the reference's split-run checkpoint test requires
(CMC.rs/tests/integration/checkpoint.rs:27-73, 154-193).  Under another
seed or replica index the counter is left alone (with a warning).  The
replicas of one state share the counter, so a load that would move it for
the other replicas of a state that has swept is refused.

O(N) spins live on the device as FP32 unit vectors, with norms within ~1e-7
of 1.  The reference accepts an O(N) spin only when |norm^2 - 1| <= 1e-10 D
(models.rs:454-473).  So "spins" holds every vector renormalised in f64,
which the reference loads, and "scamper"."spins_f32" holds the device's
exact components.  This loader checks "spins" with the reference's rule,
requires the two copies to agree within 1e-6, and restores the exact
components.
"""
from __future__ import annotations

import math
import struct
import warnings

import numpy as np

from . import ScamperError

FORMAT = "cmc-rs-snapshot-v2"


class CheckpointCorrupted(ScamperError):
    """CarloError::CheckpointCorrupted { detail } (classical_mc.rs:286-290)."""

    def __init__(self, field: str, reason: str):
        super().__init__(1, f"{field}: {reason}")
        self.detail = f"{field}: {reason}"


def _spin_dim(state) -> int:
    return max(1, int(state.model.spin_dim))


def _sweep_counter(state, ext, replica: int):
    """The sweep counter to restore from the document's "scamper" record, or
    None to leave the state's counter alone.

    The counter is restored only for the same (seed, replica): under another
    seed or replica index the document's counter names no position in this
    replica's stream, and the spins and beta load as the reference's loader
    loads them.  A state of several replicas shares one counter, so changing
    it would rewind or advance the other replicas onto draws they have
    already used: that is refused unless no replica has swept yet (a fresh
    state) or the counter already has the document's value."""
    if not isinstance(ext, dict) or _as_u64(ext.get("sweeps")) is None:
        return None
    sweeps = _as_u64(ext["sweeps"])
    seed, rep = _as_u64(ext.get("seed")), _as_u64(ext.get("replica"))
    if (seed is not None and seed != int(state.seed)) or (rep is not None and rep != int(replica)):
        warnings.warn("snapshot from another seed or replica index: its sweep counter is not restored, "
                      "so the continuation draws from this replica's own stream")
        return None
    if getattr(state, "n_replicas", 1) > 1 and state.sweeps not in (0, sweeps):
        raise ValueError(f"the state's {state.n_replicas} replicas share one sweep counter ({state.sweeps}); "
                         f"setting it to the snapshot's {sweeps} would move the other replicas onto draws "
                         "they have used: load into a fresh state, or from snapshots with the same sweep count")
    return sweeps


def save_snapshot(state, replica: int = 0, topology: bool = True) -> dict:
    """The snapshot of one replica (classical_mc.rs:65-92).  With topology=False
    the CSR arrays and edge list are left out, which the loader accepts."""
    lat = state.lattice
    dim = _spin_dim(state)
    exact = None
    if dim > 1:
        exact = state.download_f32(replica).astype(np.float64)
        v = exact.reshape(-1, dim)
        spins = (v / np.sqrt((v * v).sum(axis=1))[:, None]).ravel()
    else:
        spins = state.download_f64(replica)
    doc = {
        "format": FORMAT,
        "spins": [float(x) for x in spins],
        "beta": state.get_beta(replica),
        "spin_dim": _spin_dim(state),
        "n_sites": int(lat.n_sites),
        "n_edges": int(lat.n_edges),
    }
    if topology:
        off, nb, eid = lat.csr()
        src, dst, kinds, w = lat.edges()
        doc["offsets"] = [int(x) for x in off]
        doc["neighbors"] = [int(x) for x in nb]
        doc["edge_ids"] = [int(x) for x in eid]
        doc["edges"] = [{"source": int(s), "target": int(t), "kind": k, "weight": float(x)}
                        for s, t, k, x in zip(src, dst, kinds, w)]
    doc["scamper"] = {"seed": int(state.seed), "replica": int(replica), "sweeps": int(state.sweeps)}
    if exact is not None:
        doc["scamper"]["spins_f32"] = [float(x) for x in exact]
    return doc


def _is_number(v) -> bool:  # serde_json's as_f64: any JSON number
    return isinstance(v, (int, float)) and not isinstance(v, bool)


def _as_u64(v):  # serde_json's as_u64: a non-negative integer, else None
    if isinstance(v, int) and not isinstance(v, bool) and v >= 0:
        return v
    return None


def _check_usize_array(snapshot: dict, name: str, expected) -> None:
    """validate_snapshot_usize_array (classical_mc.rs:231-255)."""
    values = snapshot.get(name)
    if not isinstance(values, list):
        return  # version-1 documents carry no topology
    if len(values) != len(expected):
        raise CheckpointCorrupted(f"snapshot.{name}", "topology length mismatch")
    for i, (v, e) in enumerate(zip(values, expected)):
        if _as_u64(v) != int(e):
            raise CheckpointCorrupted(f"snapshot.{name}[{i}]", "topology mismatch")


def _bits(x: float) -> int:
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def _check_edges(snapshot: dict, lattice) -> None:
    """validate_snapshot_edges (classical_mc.rs:257-284)."""
    values = snapshot.get("edges")
    if not isinstance(values, list):
        return
    src, dst, kinds, w = lattice.edges()
    if len(values) != len(kinds):
        raise CheckpointCorrupted("snapshot.edges", "topology length mismatch")
    for i, v in enumerate(values):
        ok = isinstance(v, dict)
        if ok:
            weight = v.get("weight")
            ok = (_as_u64(v.get("source")) == int(src[i]) and _as_u64(v.get("target")) == int(dst[i])
                  and isinstance(v.get("kind"), str) and v["kind"] == kinds[i]
                  and _is_number(weight) and _bits(float(weight)) == _bits(float(w[i])))
        if not ok:
            raise CheckpointCorrupted(f"snapshot.edges[{i}]", "physical edge mismatch")


def load_snapshot(state, snapshot: dict, replica: int = 0) -> None:
    """Restore one replica from a snapshot (classical_mc.rs:94-165)."""
    fmt = snapshot.get("format") if isinstance(snapshot, dict) else None
    if not isinstance(fmt, str) or fmt != FORMAT:
        got = fmt if isinstance(fmt, str) else ""
        raise CheckpointCorrupted("snapshot", f"unknown snapshot format: expected '{FORMAT}', got '{got}'")
    values = snapshot.get("spins")
    if not isinstance(values, list):
        raise CheckpointCorrupted("snapshot.spins", "missing or invalid array")
    for v in values:
        if not _is_number(v):
            raise CheckpointCorrupted("snapshot.spins", "contains a non-number")
        if not math.isfinite(float(v)):
            raise CheckpointCorrupted("snapshot.spins", "contains a non-finite value")
    lat = state.lattice
    expected = lat.n_sites * _spin_dim(state)
    if len(values) != expected:
        raise CheckpointCorrupted("snapshot.spins", f"length mismatch: expected {expected}, got {len(values)}")
    n_sites = _as_u64(snapshot.get("n_sites"))
    if n_sites is not None and n_sites != lat.n_sites:
        raise CheckpointCorrupted("snapshot.n_sites", "topology mismatch")
    spin_dim = _as_u64(snapshot.get("spin_dim"))
    if spin_dim is not None and spin_dim != _spin_dim(state):
        raise CheckpointCorrupted("snapshot.spin_dim", "model mismatch")
    n_edges = _as_u64(snapshot.get("n_edges"))
    if n_edges is not None and n_edges != lat.n_edges:
        raise CheckpointCorrupted("snapshot.n_edges", "topology mismatch")
    if any(isinstance(snapshot.get(k), list) for k in ("offsets", "neighbors", "edge_ids")):
        off, nb, eid = lat.csr()
        _check_usize_array(snapshot, "offsets", off)
        _check_usize_array(snapshot, "neighbors", nb)
        _check_usize_array(snapshot, "edge_ids", eid)
    _check_edges(snapshot, lat)
    beta = snapshot.get("beta")
    if not _is_number(beta):
        raise CheckpointCorrupted("snapshot.beta", "missing or invalid")
    beta = float(beta)
    if not math.isfinite(beta) or beta < 0.0:  # System::set_beta (lattice/state.rs:67-73)
        raise CheckpointCorrupted("snapshot.beta", "beta must be finite and non-negative")
    ext = snapshot.get("scamper")
    spins = np.asarray(values, dtype=np.float64)
    dim = _spin_dim(state)
    exact = None
    if dim > 1:
        # Hamiltonian::validate_configuration for O(N) (models.rs:454-473)
        norm2 = (spins.reshape(-1, dim) ** 2).sum(axis=1)
        bad = np.nonzero(np.abs(norm2 - 1.0) > 1e-10 * dim)[0]
        if bad.size:
            raise CheckpointCorrupted("snapshot.spins", f"invalid spin at site {int(bad[0])}: "
                                      f"O(N) spin must have unit norm: norm^2={float(norm2[bad[0]])}")
        f32 = ext.get("spins_f32") if isinstance(ext, dict) else None
        if f32 is not None:
            if (not isinstance(f32, list) or len(f32) != spins.size or not all(_is_number(x) for x in f32)):
                raise CheckpointCorrupted("snapshot.scamper.spins_f32", "missing or invalid array")
            exact = np.asarray(f32, dtype=np.float64)
            if not np.all(np.isfinite(exact)) or np.max(np.abs(exact - spins), initial=0.0) > 1e-6:
                raise CheckpointCorrupted("snapshot.scamper.spins_f32", "does not match snapshot.spins")
    sweeps = _sweep_counter(state, ext, replica)
    try:  # validate_configuration on upload; nothing is written when it fails
        if exact is not None:
            state.upload_f32(exact.astype(np.float32), replica)
        else:
            state.upload_f64(spins, replica)
    except ScamperError as err:
        raise CheckpointCorrupted("snapshot.spins", str(err)) from None
    state.set_beta(beta, replica)
    if sweeps is not None:
        state.set_sweeps(sweeps)
