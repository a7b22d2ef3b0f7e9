"""Parameter-driven construction (scamper_amd.params), the checks of
CMC.rs/tests/physics/input_validation.rs:57-113 on the reference's
`FromParams for ClassicalMC` (classical_mc.rs:299-529): every malformed
configuration raises InvalidConfig naming the field, and the lattices the
parameters select are the library's builders with the reference's defaults.
Runs without a GPU (lattices are host objects; no state is created)."""
import numpy as np
import pytest

import scamper_amd as sc
from scamper_amd.params import InvalidConfig, build_lattice_from_params, parse_bool, parse_param, validate_params

ISING_BASE = {"L": "4", "beta": "0.5"}
POTTS_BASE = {"L": "4", "beta": "0.5", "q": "3"}


def _invalid(base, field, value, kind=sc.ISING, want_field=None):
    params = dict(base)
    params[field] = value
    with pytest.raises(InvalidConfig) as err:
        validate_params(params, kind)
    assert err.value.field == (want_field or field)
    assert err.value.code == 1  # SCAMPER_EINVAL


@pytest.mark.parametrize("value", ["-0.5", "NaN", "not-a-number"])
@pytest.mark.parametrize("kind", [sc.ISING, sc.POTTS, sc.ON])
def test_rejects_invalid_temperature(value, kind):
    _invalid(POTTS_BASE if kind == sc.POTTS else ISING_BASE, "beta", value, kind)


def test_rejects_unknown_lattices_and_bad_dimensions():
    for kind, base in ((sc.ISING, ISING_BASE), (sc.POTTS, POTTS_BASE)):
        _invalid(base, "lattice_type", "penrose-tiles", kind)
    for value in ["0", "-3", "four"]:
        _invalid(ISING_BASE, "L", value)
    # open-boundary triangles are unsupported geometry, not silently wrong
    with pytest.raises(InvalidConfig) as err:
        validate_params({"lattice_type": "triangular", "Lx": 4, "Ly": 4, "pbc": "false"}, sc.ISING)
    assert err.value.field == "pbc"
    for lt, lx, ly in (("triangular", 1, 4), ("honeycomb", 3, 4), ("honeycomb", 4, 1), ("kagome", 2, 1)):
        with pytest.raises(InvalidConfig) as err:
            validate_params({"lattice_type": lt, "Lx": lx, "Ly": ly}, sc.ISING)
        assert err.value.field == "Lx/Ly"
    _invalid(ISING_BASE, "pbc", "maybe")


def test_rejects_invalid_models():
    _invalid(POTTS_BASE, "q", "1", sc.POTTS)
    _invalid(POTTS_BASE, "q", "zero", sc.POTTS)
    _invalid(ISING_BASE, "J", "NaN")
    _invalid(POTTS_BASE, "J", "inf", sc.POTTS)
    # q is read before J, as the reference reads it
    _invalid(dict(POTTS_BASE, J="NaN"), "q", "1", sc.POTTS)


def test_parsers_follow_rust_str_parse():
    assert parse_param({"x": "+7"}, "x", int) == 7
    for bad in ["7.0", " 7", "0x7", "", "1_000", str(2**64)]:
        with pytest.raises(InvalidConfig):
            parse_param({"x": bad}, "x", int)
    for good, want in [("1e3", 1000.0), (".5", 0.5), ("5.", 5.0), ("-inf", -np.inf), ("Infinity", np.inf)]:
        assert parse_param({"x": good}, "x", float) == want
    assert np.isnan(parse_param({"x": "nan"}, "x", float))
    for bad in ["1_0", " 1", "0x1p3", "e5", "1e"]:
        with pytest.raises(InvalidConfig):
            parse_param({"x": bad}, "x", float)
    assert parse_param({}, "x", float) is None
    for raw, want in [("TRUE", True), ("on", True), ("1", True), ("No", False), ("off", False), (False, False)]:
        assert parse_bool({"b": raw}, "b", not want) is want
    assert parse_bool({}, "b", True) is True


def _same_lattice(a, b):
    assert (a.n_sites, a.n_edges) == (b.n_sites, b.n_edges)
    (sa, ta, ka, wa), (sb, tb, kb, wb) = a.edges(), b.edges()
    assert np.array_equal(sa, sb) and np.array_equal(ta, tb) and np.array_equal(wa, wb) and ka == kb


@pytest.mark.parametrize("params,direct", [
    ({}, lambda: sc.Lattice.chain(10)),                                       # inferred chain, default L
    ({"L": 6, "pbc": "off"}, lambda: sc.Lattice.chain(6, False)),
    ({"Lx": 6}, lambda: sc.Lattice.square(6, 6)),                             # inferred hypercubic
    ({"Lx": 6, "Ly": 4, "Lz": 2}, lambda: sc.Lattice.hypercubic([6, 4, 2])),
    ({"lattice_type": "Square", "Lx": 8, "Ly": 2}, lambda: sc.Lattice.square(8, 2)),
    ({"lattice_type": "square"}, lambda: sc.Lattice.square(4, 4)),
    ({"lattice_type": "cubic", "Lx": 4, "Lz": 6}, lambda: sc.Lattice.hypercubic([4, 4, 6])),
    ({"lattice_type": "chain", "Lx": 8}, lambda: sc.Lattice.chain(8)),
    ({"lattice_type": "triangular"}, lambda: sc.Lattice.triangular(4, 4)),
    ({"lattice_type": "honeycomb", "Lx": 4, "Ly": 2}, lambda: sc.Lattice.honeycomb(4, 2)),
    ({"lattice_type": "kagome"}, lambda: sc.Lattice.kagome(2, 2)),
])
def test_lattice_selection_and_defaults(params, direct):
    pbc = parse_bool(params, "pbc", True)
    _same_lattice(build_lattice_from_params(params, pbc), direct())


def test_rejects_unknown_initial_state_before_touching_the_device():
    from scamper_amd.params import state_from_params
    with pytest.raises(InvalidConfig) as err:
        state_from_params({"L": 4, "beta": 0.5, "initial_state": "lukewarm"}, sc.ISING, seed=1)
    assert err.value.field == "initial_state"
    validate_params({"L": 4, "beta": 0.5, "initial_state": "lukewarm"}, sc.ISING)  # not a validate_params check
