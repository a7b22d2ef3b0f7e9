"""cmc-rs-snapshot-v2 on the host: edge labels, topology arrays, and the
loader's validation (CMC.rs/src/classical_mc.rs:65-165, 231-290), checked
against a stand-in for the device state (no GPU needed).

The rejection cases mirror the reference's checkpoint tests
(CMC.rs/tests/integration/checkpoint.rs: unknown format C.2, bad edge kind
C.3, topology mismatch C.4); edge order and CSR arrays are pinned to the
oracle's lattice builders, edge labels to the builders' rules
(graph.rs:402-543, classical_mc.rs:361-383).
"""
import copy
import json

import numpy as np
import pytest


class StubState:
    """The parts of scamper_amd.State the snapshot code uses, on the host."""

    def __init__(self, sc, lattice, spin_dim=1, seed=7):
        self.lattice, self.model, self.seed = lattice, sc.model(sc.ISING, 1.0, 2, spin_dim), seed
        self.spins = np.ones(lattice.n_sites * spin_dim)
        self.beta, self.sweeps = 0.5, 0

    def download_f64(self, replica=0):
        return self.spins.copy()

    def upload_f64(self, spins, replica=0):
        self.spins = np.asarray(spins, dtype=np.float64).copy()

    def get_beta(self, replica=0):
        return self.beta

    def set_beta(self, beta, replica=0):
        self.beta = beta

    def set_sweeps(self, n):
        self.sweeps = n


def test_edge_labels_follow_the_builders(sc):
    src, dst, kinds, w = sc.Lattice.square(4, 4).edges()
    assert kinds == ["square_x", "square_y"] * 16 and np.all(w == 1.0)
    assert sc.Lattice.chain(6).edges()[2] == ["chain_x"] * 6
    assert sc.Lattice.hypercubic([4, 4, 4]).edges()[2] == ["cubic_x", "cubic_y", "cubic_z"] * 64
    assert set(sc.Lattice.hypercubic([3, 3, 3, 3]).edges()[2]) == {"generic"}
    assert sc.Lattice.triangular(3, 4).edges()[2] == ["tri_x", "tri_y", "tri_diag"] * 12
    lx, ly = 4, 3
    s, t, k, _ = sc.Lattice.honeycomb(lx, ly).edges()
    assert k == ["honey_x" if a // lx == b // lx else "honey_y" for a, b in zip(s, t)]
    assert k.count("honey_x") == lx * ly and k.count("honey_y") == lx * ly // 2
    assert set(sc.Lattice.kagome(3, 2).edges()[2]) == {"kagome"}
    assert sc.Lattice.from_edges(3, [0, 1], [1, 2], [1.0, 2.5]).edges()[2] == ["generic", "generic"]
    # the non-periodic square (CSR path) labels the same way
    assert set(sc.Lattice.square(3, 5, pbc=False).edges()[2]) == {"square_x", "square_y"}


@pytest.mark.parametrize("make", [
    lambda m: m.Lattice.square(4, 4),
    lambda m: m.Lattice.hypercubic([6, 4, 2]),
    lambda m: m.Lattice.triangular(3, 4),
    lambda m: m.Lattice.honeycomb(4, 2),
    lambda m: m.Lattice.kagome(2, 3),
    lambda m: m.Lattice.from_edges(5, [0, 1, 2, 3, 0], [1, 2, 3, 3, 1], [1.0, -2.0, 0.5, 1.0, 3.0]),
])
def test_edges_and_csr_equal_the_oracle(sc, orc, make):
    mine, ref = make(sc), make(orc).csr()
    src, dst, _, w = mine.edges()
    off, nb, eid = mine.csr()
    assert np.array_equal(src, ref["src"]) and np.array_equal(dst, ref["dst"])
    assert np.array_equal(w, ref["weight"])
    assert np.array_equal(off, ref["offsets"]) and np.array_equal(nb, ref["neighbors"])
    assert np.array_equal(eid, ref["edge_ids"])


def _stub(sc):
    st = StubState(sc, sc.Lattice.square(4, 4))
    st.spins = np.where(np.arange(16) % 3 == 0, -1.0, 1.0)
    st.sweeps = 200
    return st


def test_snapshot_document(sc):
    st = _stub(sc)
    snap = sc.snapshot.save_snapshot(st)
    assert snap["format"] == "cmc-rs-snapshot-v2"
    assert snap["spins"] == list(st.spins) and snap["beta"] == 0.5
    assert snap["spin_dim"] == 1 and snap["n_sites"] == 16 and snap["n_edges"] == 32
    off, nb, eid = st.lattice.csr()
    assert snap["offsets"] == list(map(int, off)) and snap["neighbors"] == list(map(int, nb))
    assert snap["edge_ids"] == list(map(int, eid))
    assert snap["edges"][:2] == [{"source": 0, "target": 1, "kind": "square_x", "weight": 1.0},
                                 {"source": 0, "target": 4, "kind": "square_y", "weight": 1.0}]
    assert snap["scamper"] == {"seed": 7, "replica": 0, "sweeps": 200}
    # JSON round trip, then load into a fresh stand-in
    fresh = StubState(sc, sc.Lattice.square(4, 4))
    sc.snapshot.load_snapshot(fresh, json.loads(json.dumps(snap)))
    assert np.array_equal(fresh.spins, st.spins) and fresh.beta == 0.5 and fresh.sweeps == 200
    # version-1 style: no topology arrays, no extension
    v1 = {k: snap[k] for k in ("format", "spins", "beta")}
    sc.snapshot.load_snapshot(StubState(sc, sc.Lattice.square(4, 4)), v1)
    assert "offsets" not in sc.snapshot.save_snapshot(st, topology=False)


def _corrupted(sc, snap, field_substr):
    st = StubState(sc, sc.Lattice.square(4, 4))
    before = (st.spins.copy(), st.beta, st.sweeps)
    with pytest.raises(sc.snapshot.CheckpointCorrupted) as err:
        sc.snapshot.load_snapshot(st, snap)
    assert field_substr in err.value.detail
    assert np.array_equal(st.spins, before[0]) and (st.beta, st.sweeps) == before[1:], "state changed"


def test_loader_rejections(sc):
    snap = sc.snapshot.save_snapshot(_stub(sc))
    cases = []
    bad = copy.deepcopy(snap); bad["format"] = "cmc-rs-snapshot-v1"; cases.append((bad, "unknown snapshot format"))
    bad = copy.deepcopy(snap); del bad["format"]; cases.append((bad, "unknown snapshot format"))
    bad = copy.deepcopy(snap); bad["edges"][0]["kind"] = "invalid_kind"; cases.append((bad, "snapshot.edges[0]"))
    bad = copy.deepcopy(snap); bad["n_sites"] = 999; cases.append((bad, "snapshot.n_sites"))
    bad = copy.deepcopy(snap); bad["spin_dim"] = 3; cases.append((bad, "snapshot.spin_dim"))
    bad = copy.deepcopy(snap); bad["n_edges"] = 31; cases.append((bad, "snapshot.n_edges"))
    bad = copy.deepcopy(snap); bad["spins"] = bad["spins"][:-1]; cases.append((bad, "length mismatch"))
    bad = copy.deepcopy(snap); bad["spins"][3] = "up"; cases.append((bad, "non-number"))
    bad = copy.deepcopy(snap); bad["spins"][3] = True; cases.append((bad, "non-number"))
    bad = copy.deepcopy(snap); bad["spins"][3] = float("nan"); cases.append((bad, "non-finite"))
    bad = copy.deepcopy(snap); del bad["spins"]; cases.append((bad, "missing or invalid array"))
    bad = copy.deepcopy(snap); bad["neighbors"][5] += 1; cases.append((bad, "snapshot.neighbors[5]"))
    bad = copy.deepcopy(snap); bad["offsets"] = bad["offsets"][:-1]; cases.append((bad, "snapshot.offsets"))
    bad = copy.deepcopy(snap); bad["edge_ids"][0] = 1.0; cases.append((bad, "snapshot.edge_ids[0]"))
    bad = copy.deepcopy(snap); bad["edges"][7]["weight"] = 1.0 + 2.0 ** -52; cases.append((bad, "snapshot.edges[7]"))
    bad = copy.deepcopy(snap); bad["edges"][2]["target"] = 0; cases.append((bad, "snapshot.edges[2]"))
    bad = copy.deepcopy(snap); bad["edges"] = bad["edges"][:5]; cases.append((bad, "snapshot.edges"))
    bad = copy.deepcopy(snap); del bad["beta"]; cases.append((bad, "snapshot.beta"))
    bad = copy.deepcopy(snap); bad["beta"] = -0.1; cases.append((bad, "snapshot.beta"))
    for doc, where in cases:
        _corrupted(sc, doc, where)
    # the reference skips a count that is not a non-negative integer
    ok = copy.deepcopy(snap); ok["n_sites"] = 16.0
    sc.snapshot.load_snapshot(StubState(sc, sc.Lattice.square(4, 4)), ok)
