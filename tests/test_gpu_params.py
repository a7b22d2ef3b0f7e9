"""States built from parameters (scamper_amd.params.state_from_params,
FromParams::from_params, CMC.rs/src/classical_mc.rs:443-491) equal states
built directly: same lattice, model, beta and start, so the same sweeps give
the same spins; a cold start has every bond satisfied."""
import numpy as np
import pytest

import scamper_amd as sc
from scamper_amd.params import state_from_params

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("params,kind,direct", [
    ({"lattice_type": "triangular", "Lx": 6, "Ly": 6, "beta": 0.3}, sc.ISING,
     lambda: (sc.Lattice.triangular(6, 6), sc.model(sc.ISING, 1.0))),
    ({"lattice_type": "kagome", "Lx": 4, "beta": "0.8", "q": 4, "J": "1.5"}, sc.POTTS,
     lambda: (sc.Lattice.kagome(4, 4), sc.model(sc.POTTS, 1.5, 4))),
    ({"Lx": 8, "Ly": 8, "Lz": 4, "beta": 0.7}, sc.ON,
     lambda: (sc.Lattice.hypercubic([8, 8, 4]), sc.model(sc.ON, 1.0, 2, 3))),
    ({"lattice_type": "square", "Lx": 16, "beta": 1.1, "J": 0.5}, -2,
     lambda: (sc.Lattice.hypercubic([16, 16]), sc.model(sc.ON, 0.5, 2, 2))),
])
def test_state_from_params_equals_direct_state(params, kind, direct):
    n_comp = 2 if kind == -2 else 3
    kind = sc.ON if kind == -2 else kind
    alg = sc.CONT_HEAT_BATH if kind == sc.ON else sc.METROPOLIS
    a = state_from_params(params, kind, seed=9, n_replicas=2, spin_dim=n_comp)
    lat, m = direct()
    b = sc.State(lat, m, 2, 9)
    for r in range(2):
        b.init(True, r)
        b.set_beta(float(params["beta"]), r)
    a.sweep(alg, 5)
    b.sweep(alg, 5)
    for r in range(2):
        assert np.array_equal(a.download_f64(r), b.download_f64(r))
        assert a.get_beta(r) == b.get_beta(r)


def test_cold_start_satisfies_every_bond():
    st = state_from_params({"lattice_type": "honeycomb", "Lx": 6, "Ly": 4, "beta": 0.5, "J": 2,
                            "initial_state": "ordered"}, sc.ISING, seed=1)
    lat = sc.Lattice.honeycomb(6, 4)
    e, m = st.observables()
    assert e == -2.0 * lat.n_edges and m == 1.0
