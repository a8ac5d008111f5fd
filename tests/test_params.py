"""Run parameters in the reference's parameters.toml layout (host logic):
keys and derived values of the config-0 fixture, and fail-fast errors that
name the offending key (the reference aborts on them)."""
import os

import pytest

from ohhelppic3d_amd import params

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "parameters_config0.toml")


def fixture_text():
    with open(FIXTURE) as f:
        return f.read()


def test_config0_parameters():
    spec = params.load_parameters(FIXTURE)
    c = spec.cfg
    assert tuple(c.n) == (32, 32, 32) and c.dt == 0.1 and spec.nstep == 100
    assert tuple(c.qm) == (-1.0, 0.01)
    # charge_per_macro_particle (parameters.f90:140-149): wp^2 / (q/m) / ppc
    assert c.q_macro[0] == pytest.approx(-1.0 / 16, rel=1e-15)
    assert c.q_macro[1] == pytest.approx(1.0 / 16, rel=1e-12)
    assert spec.counts == [16 * 32 ** 3] * 2 and list(c.capacity) == spec.counts
    assert spec.vth == [1.0, 0.1] and spec.seed == (42, 52)
    assert tuple(c.b0) == (0.0, 0.0, 0.5)   # omega_ce / |q/m| of species 1
    assert spec.name == "config0"
    assert "plasma.injection" in spec.ignored and "system.mpi" in spec.ignored


@pytest.mark.parametrize("old,new,key", [
    ("boundary_communication = ['periodic', 'periodic', 'periodic']",
     "boundary_communication = ['periodic', 'periodic', 'no_periodic']", "boundary_communication"),
    ("['periodic', 'periodic', 'periodic'],\n]", "['periodic', 'periodic', 'vanish'],\n]",
     "boundary_type_for_particle"),
    ("particle_mover = 'boris'", "particle_mover = 'buneman'", "particle_mover"),
    ("thermal_velocity_perp = [1.0, 0.1]", "thermal_velocity_perp = [2.0, 0.1]",
     "thermal_velocity_perp"),
    ("flow_velocity = [0.0, 0.0]", "flow_velocity = [-13.3, 0.0]", "flow_velocity"),
    ("nz = 32\n", "", "system.nz"),
    ("plasma_initialization = ['random-uniform', 'random-uniform']",
     "plasma_initialization = ['random-uniform', 'maxwell-beam']", "initialisation"),
    ("nspecies = 2", "nspecies = 3", "nspecies is 3"),
])
def test_unsupported_or_missing_keys_fail_fast(old, new, key):
    text = fixture_text()
    assert old in text
    with pytest.raises(params.ParameterError) as e:
        params.parse_parameters(text.replace(old, new, 1))
    assert key in str(e.value)


def test_species_without_initial_particles():
    text = fixture_text().replace("plasma_initialization = ['random-uniform', 'random-uniform']",
                                  "plasma_initialization = ['random-uniform', 'none']")
    spec = params.parse_parameters(text)
    assert spec.counts == [16 * 32 ** 3, 0]


def test_malformed_toml():
    with pytest.raises(params.ParameterError):
        params.parse_parameters("[system\nnx = 3")
