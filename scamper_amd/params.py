"""Parameter-driven construction of a device state, mirroring the
reference's `FromParams for ClassicalMC` and `build_lattice_from_params`
(CMC.rs/src/classical_mc.rs:299-491) and `FromHamiltonianParams`
(classical_mc.rs:493-529).

Parameters are a mapping of names to values, read as strings the way
carlo_rs `Params` stores them, so `{"L": 8}` and `{"L": "8"}` mean the same.
Every rejection the reference makes raises `InvalidConfig(field, reason)`,
the counterpart of `CarloError::InvalidConfig`, before any device memory is
touched:

* beta: finite and non-negative (default 1.0);
* pbc: true/1/yes/on or false/0/no/off, any case (default true);
* lattice_type: chain, square, cubic, hypercubic (inferred from Lx when
  absent), triangular, honeycomb or kagome, the last three periodic only,
  with the reference's dimension rules;
* J finite (default 1.0); Potts q >= 2 (default 3);
* initial_state: hot/random or cold/ordered (default hot), checked by
  `state_from_params` only, as the reference checks it in from_params.

Usage:

    st = state_from_params({"lattice_type": "square", "Lx": 64, "beta": 0.44},
                           ISING, seed=42)
    st.sweep(METROPOLIS, 1000)
"""
from __future__ import annotations

import math
import re
from typing import Mapping

from . import ISING, ON, POTTS, Lattice, ScamperError, State, model

__all__ = ["InvalidConfig", "parse_param", "parse_bool", "build_lattice_from_params", "model_from_params",
           "validate_params", "state_from_params"]


class InvalidConfig(ScamperError):
    """CarloError::InvalidConfig { field, reason } (classical_mc.rs:292-297)."""

    def __init__(self, field: str, reason: str):
        super().__init__(1, f"invalid parameter `{field}`: {reason}")  # SCAMPER_EINVAL
        self.field = field
        self.reason = reason


# str::parse::<f64> and ::<usize> accept exactly these forms
_F64 = re.compile(r"[+-]?(?:inf|infinity|nan|(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?)", re.IGNORECASE)
_USIZE = re.compile(r"\+?\d+")
_USIZE_MAX = 2**64 - 1


def _raw(params: Mapping, name: str) -> str:
    value = params[name]
    if isinstance(value, bool):
        return "true" if value else "false"
    return str(value)


def parse_param(params: Mapping, name: str, kind: type):
    """parse_param (classical_mc.rs:299-313): None when absent, the parsed
    value, or InvalidConfig when the string does not parse as `kind`
    (float or int, the reference's f64 and usize)."""
    if name not in params:
        return None
    raw = _raw(params, name)
    if kind is float:
        if not _F64.fullmatch(raw):
            raise InvalidConfig(name, f"cannot parse `{raw}`: invalid float literal")
        return float(raw)
    if kind is int:
        if not _USIZE.fullmatch(raw) or int(raw) > _USIZE_MAX:
            raise InvalidConfig(name, f"cannot parse `{raw}`: invalid digit found in string")
        return int(raw)
    raise TypeError(kind)


def parse_bool(params: Mapping, name: str, default: bool) -> bool:
    """parse_bool (classical_mc.rs:315-324)."""
    if name not in params:
        return default
    value = _raw(params, name)
    low = value.lower()
    if low in ("true", "1", "yes", "on"):
        return True
    if low in ("false", "0", "no", "off"):
        return False
    raise InvalidConfig(name, f"expected boolean, got `{value}`")


def _positive_dimension(params: Mapping, name: str, default: int) -> int:
    """positive_dimension (classical_mc.rs:326-333)."""
    value = parse_param(params, name, int)
    value = default if value is None else value
    if value == 0:
        raise InvalidConfig(name, "dimension must be positive")
    return value


def _lattice_spec(params: Mapping, pbc: bool):
    """The builder and its arguments that build_lattice_from_params
    (classical_mc.rs:336-425) picks, after every check it makes."""
    inferred = "hypercubic" if "Lx" in params else "chain"
    kind = _raw(params, "lattice_type").lower() if "lattice_type" in params else inferred
    if kind == "chain":
        length = _positive_dimension(params, "L", 10) if "L" in params else _positive_dimension(params, "Lx", 10)
        return "hypercubic", ([length], pbc)
    if kind == "square":
        lx = _positive_dimension(params, "Lx", 4)
        return "hypercubic", ([lx, _positive_dimension(params, "Ly", lx)], pbc)
    if kind == "cubic":
        lx = _positive_dimension(params, "Lx", 4)
        ly = _positive_dimension(params, "Ly", lx)
        return "hypercubic", ([lx, ly, _positive_dimension(params, "Lz", lx)], pbc)
    if kind == "hypercubic":
        if "Lx" in params:
            lx = _positive_dimension(params, "Lx", 4)
            ly = _positive_dimension(params, "Ly", lx)
            if "Lz" in params:
                return "hypercubic", ([lx, ly, _positive_dimension(params, "Lz", lx)], pbc)
            return "hypercubic", ([lx, ly], pbc)
        return "hypercubic", ([_positive_dimension(params, "L", 10)], pbc)
    if kind in ("triangular", "honeycomb", "kagome"):
        if not pbc:
            raise InvalidConfig("pbc", f"lattice_type `{kind}` currently requires periodic boundaries")
        default = 2 if kind == "kagome" else 4
        lx = _positive_dimension(params, "Lx", default)
        ly = _positive_dimension(params, "Ly", lx)
        if kind == "triangular" and (lx < 2 or ly < 2):
            raise InvalidConfig("Lx/Ly", "triangular lattice requires both dimensions >= 2")
        if kind == "honeycomb" and (lx < 2 or ly < 2 or lx % 2):
            raise InvalidConfig("Lx/Ly", "honeycomb lattice requires even Lx >= 2 and Ly >= 2")
        if kind == "kagome" and (lx < 2 or ly < 2):
            raise InvalidConfig("Lx/Ly", "kagome lattice requires both dimensions >= 2")
        return kind, (lx, ly)
    raise InvalidConfig("lattice_type", f"unknown lattice `{kind}`; expected chain, square, cubic, hypercubic, "
                                        "triangular, honeycomb, or kagome")


def build_lattice_from_params(params: Mapping, pbc: bool) -> Lattice:
    """build_lattice_from_params (classical_mc.rs:336-425) over the
    library's builders (same sites, bonds and bond order)."""
    builder, args = _lattice_spec(params, pbc)
    return getattr(Lattice, builder)(*args)


def _finite_coupling(params: Mapping) -> float:
    """finite_coupling (classical_mc.rs:497-504)."""
    coupling = parse_param(params, "J", float)
    coupling = 1.0 if coupling is None else coupling
    if not math.isfinite(coupling):
        raise InvalidConfig("J", "must be finite")
    return coupling


def model_from_params(kind: int, params: Mapping, spin_dim: int = 3):
    """FromHamiltonianParams (classical_mc.rs:506-529) for ISING, POTTS (q
    before J, as the reference reads them) and ON (spin_dim >= 2; the
    device runs N = 2 and N = 3 and refuses others when the state is
    created)."""
    if kind == ISING:
        return model(ISING, _finite_coupling(params))
    if kind == POTTS:
        q = parse_param(params, "q", int)
        q = 3 if q is None else q
        if q < 2:
            raise InvalidConfig("q", "Potts q must be >= 2")
        return model(POTTS, _finite_coupling(params), q)
    if kind == ON:
        if spin_dim < 2:
            raise InvalidConfig("spin_dim", "O(N) dimension must be >= 2")
        return model(ON, _finite_coupling(params), 2, spin_dim)
    raise ValueError(f"unknown model kind {kind}")


def validate_params(params: Mapping, kind: int, spin_dim: int = 3) -> None:
    """FromParams::validate_params (classical_mc.rs:432-441): beta, pbc, the
    lattice and the model, in that order."""
    beta = parse_param(params, "beta", float)
    beta = 1.0 if beta is None else beta
    if not math.isfinite(beta) or beta < 0.0:
        raise InvalidConfig("beta", "must be finite and non-negative")
    _lattice_spec(params, parse_bool(params, "pbc", True))
    model_from_params(kind, params, spin_dim)


def state_from_params(params: Mapping, kind: int, seed: int, n_replicas: int = 1, spin_dim: int = 3) -> State:
    """FromParams::from_params (classical_mc.rs:443-491) as a device state:
    every replica at the parameters' beta, started hot (random) or cold
    (ordered).  The reference draws the hot start from its RNG; here it is
    the counter-based start of (seed, replica)."""
    validate_params(params, kind, spin_dim)
    initial = _raw(params, "initial_state").lower() if "initial_state" in params else "hot"
    if initial not in ("hot", "random", "cold", "ordered"):
        raise InvalidConfig("initial_state", "expected hot/random or cold/ordered")
    lat = build_lattice_from_params(params, parse_bool(params, "pbc", True))
    beta = parse_param(params, "beta", float)
    st = State(lat, model_from_params(kind, params, spin_dim), n_replicas, seed)
    for r in range(n_replicas):
        st.init(initial in ("hot", "random"), r)
        st.set_beta(1.0 if beta is None else beta, r)
    return st
