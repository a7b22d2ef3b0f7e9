"""The reference-side binding, call for call: a device-backed MetropolisCore
that honours System (CMC.rs/src/lattice/state.rs:19-28).

GpuMetropolisCore below is the mirror core of the integration notes, written
against the Python bindings with the same calls in the same order: upload
the host's spins when the host changed them, set beta when it changed, sweep
with the phase, download the spins, and set System::energy (a plain field)
from the device's exact energy.  After every sweep the reference's own
readers of System must see what the reference's CPU core would leave there:
the spins, the cached energy that audit_lattice_cache (CMC.rs/src/audit.rs:
41-52) compares with a recount, and the magnetisation that
ClassicalMC::measure computes from system.spins (observables/
magnetization.rs:14-15).  The CPU core is the oracle's restatement of
MetropolisCore with the same draws.  Mid-run the host writes spins (as
load_snapshot does, classical_mc.rs:150-165) and changes beta (as
change_parameter does, classical_mc.rs:216-228).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THERMALIZATION, MEASUREMENT = 0, 1


class System:
    """The reference's System: plain fields."""

    def __init__(self, spins, energy, beta):
        self.spins, self.energy, self.beta = spins, energy, beta


class GpuMetropolisCore:
    def __init__(self, sc, lattice, model, seed):
        self.sc = sc
        self.state = sc.State(lattice, model, 1, seed)
        self.synced = None
        self.beta = math.nan

    def sweep_with_phase(self, system, phase):
        if self.synced is None or not np.array_equal(system.spins, self.synced):
            self.state.upload_f64(system.spins)
        if np.float64(system.beta).tobytes() != np.float64(self.beta).tobytes():
            self.state.set_beta(system.beta)
            self.beta = system.beta
        self.state.sweep_with_phase(self.sc.METROPOLIS, 1, False, phase)
        system.spins = self.state.download_f64()
        self.synced = system.spins.copy()
        system.energy, _ = self.state.observables()


CASES = [
    ("hypercubic", [[16, 16]], 0, 2, 1.0),
    ("hypercubic", [[12, 12]], 1, 3, 1.0),
    ("triangular", [10, 10], 0, 2, -1.0),
]


@pytest.mark.parametrize("builder,args,kind,q,j", CASES)
def test_mirror_core_keeps_system_exact(sc, orc, builder, args, kind, q, j):
    seed, beta0, beta1 = 8, 0.35, 0.6
    print(f"case binding {builder}{args} kind={kind} q={q}", flush=True)
    olat = getattr(orc.Lattice, builder)(*args)
    om = orc.model(kind, j, q)
    key = orc.replica_key(seed, 0)
    start = orc.random_start(olat, om, orc.replica_key(99, 0))
    chain = orc.Chain(olat, om, start.copy())
    system = System(start.copy(), chain.total_energy(), beta0)
    core = GpuMetropolisCore(sc, getattr(sc.Lattice, builder)(*args), sc.model(kind, j, q), seed)
    rng = np.random.default_rng(1)
    for t in range(16):
        if t == 6:
            # the host writes spins (a snapshot load) and recomputes the energy
            sites = rng.choice(olat.n_sites, 5, replace=False)
            new = system.spins.copy()
            new[sites] = (-new[sites]) if kind == 0 else (new[sites] + 1) % q
            chain = orc.Chain(olat, om, new.copy())
            system.spins, system.energy = new, chain.total_energy()
        if t == 10:
            system.beta = beta1  # change_parameter
        core.sweep_with_phase(system, THERMALIZATION if t < 8 else MEASUREMENT)
        chain.metropolis(system.beta, key, t)
        assert np.array_equal(system.spins, chain.spins), f"sweep {t}: spins"
        assert system.energy == chain.energy.value, f"sweep {t}: cached energy"
        # audit_lattice_cache on the reference's side
        exact = orc.Chain(olat, om, system.spins.copy()).total_energy()
        assert abs(system.energy - exact) <= 1e-10 * (1.0 + abs(exact))
        # ClassicalMC::measure's magnetisation from system.spins
        assert orc.Chain(olat, om, system.spins.copy()).magnetization() == chain.magnetization()


def test_phase_is_validated(sc):
    st = sc.State(sc.Lattice.square(8, 8), sc.model(sc.ISING), 1, 1)
    with pytest.raises(sc.ScamperError):
        st.sweep_with_phase(sc.METROPOLIS, 1, False, 2)
