"""cmc-rs-snapshot-v2 with device state: a split run (sweep, save, restore
into a fresh state, sweep on) equals the continuous run bit for bit, the
checks of CMC.rs/tests/integration/checkpoint.rs:27-73 (C.1), 120-146 (C.5)
and 148-193 (C.6), here on the device's counter-based draws, for Ising on
the stencil and on graph lattices, Potts and O(3).  Invalid spins are
refused by the upload's validate_configuration and leave the state as it
was."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(sc, lat, m, beta, seed):
    st = sc.State(lat, m, 1, seed)
    st.set_beta(beta)
    return st


CASES = [
    # lattice, model, algorithm, beta, seed, first, second
    (lambda sc: sc.Lattice.square(4, 4), lambda sc: sc.model(sc.ISING, 1.0), 0, 0.5, 42, 200, 200),
    (lambda sc: sc.Lattice.square(4, 4), lambda sc: sc.model(sc.ISING, 1.0), 0, 0.44, 999, 400, 600),
    (lambda sc: sc.Lattice.square(64, 64), lambda sc: sc.model(sc.ISING, 1.0), 0, 0.4406868, 5, 30, 50),
    (lambda sc: sc.Lattice.square(16, 16), lambda sc: sc.model(sc.POTTS, 1.0, 3), 0, 1.005052, 11, 40, 60),
    (lambda sc: sc.Lattice.square(16, 16), lambda sc: sc.model(sc.POTTS, 1.0, 3), 1, 1.005052, 12, 20, 30),
    (lambda sc: sc.Lattice.triangular(6, 6), lambda sc: sc.model(sc.ISING, 1.0), 0, 0.3, 13, 50, 70),
    (lambda sc: sc.Lattice.cubic(4), lambda sc: sc.model(sc.ON, 1.0, 2, 3), 4, 0.7, 14, 10, 15),
    (lambda sc: sc.Lattice.square(8, 8), lambda sc: sc.model(sc.ON, 1.0, 2, 2), 4, 1.1, 15, 10, 15),
    (lambda sc: sc.Lattice.square(8, 8), lambda sc: sc.model(sc.ON, 1.0, 2, 2), 5, 1.1, 16, 10, 15),
    # graph lattices: O(3) antiferromagnet, XY clusters, real-weight Ising
    (lambda sc: sc.Lattice.triangular(4, 4), lambda sc: sc.model(sc.ON, -1.0, 2, 3), 4, 0.9, 17, 10, 15),
    (lambda sc: sc.Lattice.honeycomb(4, 4), lambda sc: sc.model(sc.ON, 1.0, 2, 2), 6, 1.0, 18, 10, 15),
    (lambda sc: sc.Lattice.from_edges(5, [0, 1, 2, 3, 4, 1], [1, 2, 3, 4, 0, 1], [0.5, 1.75, -0.25, 2.0, 1.0, 0.5]),
     lambda sc: sc.model(sc.ISING, 1.0), 0, 0.6, 19, 30, 40),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_split_run_equals_continuous_run(sc, case):
    make_lat, make_model, alg, beta, seed, first, second = CASES[case]
    lat, m = make_lat(sc), make_model(sc)
    print(f"case {case}: n_sites={lat.n_sites} alg={alg} sweeps {first}+{second}", flush=True)
    cont = _state(sc, lat, m, beta, seed)
    cont.init(True)
    cont.sweep(alg, first + second)
    split = _state(sc, lat, m, beta, seed)
    split.init(True)
    split.sweep(alg, first)
    snap = json.loads(json.dumps(split.save_snapshot()))
    split.close()
    fresh = _state(sc, lat, m, 1.0, seed)  # another beta: the snapshot restores it
    fresh.load_snapshot(snap)
    assert fresh.sweeps == first
    fresh.sweep(alg, second)
    assert np.array_equal(fresh.download_f64(), cont.download_f64()), "spins differ"
    assert fresh.observables() == cont.observables()
    assert fresh.sweeps == cont.sweeps


def test_roundtrip_recomputes_energy(sc):
    lat = sc.Lattice.square(4, 4)
    st = _state(sc, lat, sc.model(sc.ISING, 1.0), 0.5, 3)
    st.init(True)
    st.sweep(0, 7)
    before = st.observables()
    snap = st.save_snapshot()
    st.init(False)  # a different configuration in between
    assert st.observables() != before
    st.load_snapshot(snap)
    assert st.observables() == before


def test_invalid_spins_rejected_and_state_kept(sc):
    lat = sc.Lattice.square(4, 4)
    st = _state(sc, lat, sc.model(sc.POTTS, 1.0, 3), 0.8, 4)
    st.init(True)
    keep = st.download_f64()
    snap = st.save_snapshot()
    snap["spins"][5] = 5.0  # not a state of q = 3
    snap["beta"] = 0.3
    with pytest.raises(sc.snapshot.CheckpointCorrupted) as err:
        st.load_snapshot(snap)
    assert "snapshot.spins" in err.value.detail
    assert np.array_equal(st.download_f64(), keep) and st.get_beta() == 0.8


@pytest.mark.parametrize("n_comp", [3, 2])
def test_o3_snapshot_meets_the_reference_norm_rule(sc, n_comp):
    """Exported O(N) spins (N = 3, 2) pass Hamiltonian::validate_configuration's
    |norm^2 - 1| <= 1e-10 D (models.rs:454-473); the exact FP32 copy agrees,
    and a document without it (as the reference writes one) loads too."""
    lat, m = sc.Lattice.cubic(4), sc.model(sc.ON, 1.0, 2, n_comp)
    st = _state(sc, lat, m, 0.7, 21)
    st.init(True)
    st.sweep(4, 5)
    snap = st.save_snapshot()
    assert snap["spin_dim"] == n_comp
    v = np.asarray(snap["spins"]).reshape(-1, n_comp)
    assert np.all(np.abs((v * v).sum(axis=1) - 1.0) <= 3e-10)
    f = np.asarray(snap["scamper"]["spins_f32"])
    assert np.array_equal(f.astype(np.float32), st.download_f32())
    assert np.max(np.abs(f - v.ravel())) < 1e-6
    plain = {k: val for k, val in snap.items() if k != "scamper"}
    fresh = _state(sc, lat, m, 1.0, 21)
    fresh.load_snapshot(plain)
    assert np.max(np.abs(fresh.download_f32() - f)) < 1e-6 and fresh.get_beta() == 0.7


def test_two_replica_state_keeps_its_shared_counter(sc):
    """The replicas of one state share a sweep counter: a snapshot loads into
    a fresh two-replica state (the counter moves, no replica has swept), and
    each replica then continues its own original trajectory; a load that
    would move the counter of a state that has swept is refused and changes
    nothing; a document from another seed leaves the counter alone."""
    lat = sc.Lattice.square(16, 16)
    m = sc.model(sc.ISING, 1.0)
    ref = sc.State(lat, m, 2, 7)
    for r in range(2):
        ref.init(True, r)
        ref.set_beta(0.45 + 0.1 * r, r)
    ref.sweep(sc.METROPOLIS, 20)
    docs = [ref.save_snapshot(r) for r in range(2)]
    ref.sweep(sc.METROPOLIS, 15)
    fresh = sc.State(lat, m, 2, 7)
    for r in range(2):
        fresh.load_snapshot(json.loads(json.dumps(docs[r])), r)
    assert fresh.sweeps == 20
    fresh.sweep(sc.METROPOLIS, 15)
    for r in range(2):
        assert np.array_equal(fresh.download_i8(r), ref.download_i8(r)), f"replica {r}"
    # a state that has swept: moving its shared counter is refused
    before = fresh.download_i8(1)
    with pytest.raises(ValueError, match="share one sweep counter"):
        fresh.load_snapshot(json.loads(json.dumps(docs[1])), 1)
    assert np.array_equal(fresh.download_i8(1), before) and fresh.sweeps == 35
    # another seed: spins and beta load, the counter stays
    other = sc.State(lat, m, 1, 8)
    other.sweep(sc.METROPOLIS, 3)
    with pytest.warns(UserWarning, match="another seed"):
        other.load_snapshot(json.loads(json.dumps(docs[0])), 0)
    assert other.sweeps == 3
    assert np.array_equal(other.download_f64(0), np.asarray(docs[0]["spins"]))
