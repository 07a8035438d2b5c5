"""The params.json loader and Validate checks (host code; no GPU needed).

Each expectation is the reference's behaviour read from its source: the keys and defaults of
physics_params.{h,cc}, interaction_params.{h,cc} and stochastic_params.{h,cc}, nlohmann::json's
value()/get<size_t>() conversions, and the quirks the reference's driver depends on."""
import json
import math
import os

import pytest

from antimatter_amd import AmError
from antimatter_amd import params as P
from conftest import GOLDEN

M_E = 9.1093837015e-31
M_P = 1.67262192369e-27
E = 1.602176634e-19


def write(tmp_path, obj, name="params.json"):
    path = tmp_path / name
    path.write_text(obj if isinstance(obj, str) else json.dumps(obj))
    return str(path)


def shipped():
    return json.load(open(os.path.join(GOLDEN, "params_config1.json")))


def test_shipped_params_json():
    p = P.load_params(os.path.join(GOLDEN, "params_config1.json"))
    assert p.loaded == 1 and p.seed == 42 and p.b0 == 1.0 and p.trap_voltage == 10.0
    assert p.num_steps == 100  # "1e2" read through size_t
    assert list(p.populations) == [0, 100, 0, 100, 0, 0] and list(p.population_listed) == [0, 1, 0, 1, 0, 0]
    assert list(p.temperatures) == [0, 1000.0, 0, 1000.0, 0, 0]
    assert list(p.initial_cloud_radii) == [1e-4] * 3
    assert p.bath_temp == 4.0 and p.cooling_scale == 1.0
    assert p.coulomb_enabled == 1 and p.coulomb_method == b"standard" and p.energy_method == b"standard"
    assert p.cutoff_radius == 1e-5 and p.barnes_hut_theta == 0.7 and p.fast_multipole_order == 4.0
    assert p.annihilation_enabled == 1 and p.contact_radius == 5e-6 and p.annihilation_probability == 1.0
    # dt = pi min(1e-6, m/|q|) / (2 b0): the electron's m/e, 8.930967e-12 s as the notebook states
    assert p.dt == (M_E / E) * math.pi / 2.0
    assert float("%.6e" % p.dt) == 8.930967e-12
    P.validate(p)


def test_missing_file_keeps_defaults_and_fails_validate(tmp_path):
    p = P.load_params(str(tmp_path / "absent.json"))
    assert p.loaded == 0 and p.dt == 1e-9 and p.num_steps == 5000 and p.cooling_scale == 10.0
    assert list(p.populations) == [0, 10, 0, 0, 0, 0]
    assert list(p.temperatures) == [1000.0, 4.0, 0, 4.0, 0, 0]
    with pytest.raises(AmError, match="cooling_scale"):  # the default 10 > 1 (physics_params.cc:139)
        P.validate(p)


def test_temperatures_and_counts_go_through_size_t():
    cfg = shipped()
    cfg["particle_temperatures"] = {"positron": 1000.7, "electron": 3.99}
    cfg["num_steps"] = 250.9
    p = P.parse_params(cfg)
    assert p.temperatures[1] == 1000.0 and p.temperatures[3] == 3.0 and p.num_steps == 250


@pytest.mark.parametrize("key, value", [("num_steps", 1e300), ("num_steps", -1.5), ("seed", 1e20),
                                        ("particle_temperatures", {"electron": -2.5}),
                                        ("particle_populations", {"positron": 1e30})])
def test_out_of_range_integers_are_rejected(key, value):
    # nlohmann's get<size_t>/get<int> would static_cast these (undefined behaviour); the loader
    # rejects them instead
    cfg = shipped()
    cfg[key] = value
    with pytest.raises(AmError, match="outside the range"):
        P.parse_params(cfg)


def test_dt_rule():
    cfg = shipped()
    cfg["particle_populations"] = {"antiproton": 10, "antihydrogen": 5}
    cfg["b0"] = 3.0
    p = P.parse_params(cfg)
    assert p.dt == ((M_P / E) * math.pi) / (3.0 * 2)  # pbar's m/e; antihydrogen is neutral (m/0 = inf)
    cfg = {"b0": 2.0}  # no populations: the default ratio still sets dt (physics_params.cc:74)
    p = P.parse_params(cfg)
    assert p.dt == 1e-6 * math.pi / 4.0 and list(p.populations) == [0, 10, 0, 0, 0, 0]


def test_unknown_species_and_malformed_json(tmp_path):
    cfg = shipped()
    cfg["particle_populations"] = {"antihelium": 3}  # not in ParseSpecies (physics_params.cc:17-25)
    with pytest.raises(AmError, match="antihelium"):
        P.parse_params(cfg)
    with pytest.raises(AmError):
        P.load_params(write(tmp_path, '{"b0": 1, '))
    with pytest.raises(AmError):
        P.parse_params({"num_steps": "many"})


@pytest.mark.parametrize("key,value,ok", [
    ("coulomb_computation_method", "barnes_hut", False),  # the README's spelling is rejected
    ("coulomb_computation_method", "barnes-hut", True),
    ("energy_computation_method", "fast_multipole", False),
    ("b0", 0.0, False), ("z0", -1.0, False), ("num_steps", 0, False), ("bath_temperature", -1.0, False),
    ("cooling_scale", 1.0000001, False), ("cooling_scale", 0.0, True), ("cutoff_radius", 0.0, False),
    ("initial_cloud_radii", {"radius_x": 0.01}, False), ("initial_cloud_radii", {"radius_z": 0.0099}, True),
    ("annihilation_probability", 1.5, False), ("contact_radius", 0.0, False),
])
def test_validate_rules(key, value, ok):
    cfg = shipped()
    cfg[key] = value
    p = P.parse_params(cfg)
    if ok:
        P.validate(p)
    else:
        with pytest.raises(AmError):
            P.validate(p)


def test_validate_quirks():
    # The energy half of InteractionParams::Validate uses the Coulomb method (interaction_params.cc:74):
    # Barnes-Hut forces with a "standard" energy and no cutoff pass.
    cfg = shipped()
    cfg.update({"coulomb_computation_method": "barnes-hut", "energy_computation_method": "standard",
                "cutoff_radius": 0.0})
    P.validate(P.parse_params(cfg))
    # ... and with Coulomb forces off, the "standard" method still needs a cutoff.
    cfg = shipped()
    cfg.update({"coulomb_forces_enabled": False, "cutoff_radius": 0.0})
    with pytest.raises(AmError, match="cutoff"):
        P.validate(P.parse_params(cfg))
    # An FMM order above 8 fails, a fractional one inside [1, 8] passes (it is truncated at the call).
    cfg = shipped()
    cfg.update({"coulomb_computation_method": "fast-multipole", "fast_multipole_order": 8.5})
    with pytest.raises(AmError):
        P.validate(P.parse_params(cfg))
    cfg["fast_multipole_order"] = 4.9
    P.validate(P.parse_params(cfg))
    cfg.update({"fast_multipole_order": 4, "fast_multipole_theta": 1.0})
    with pytest.raises(AmError):
        P.validate(P.parse_params(cfg))
    # Annihilation settings are checked only with annihilation on.
    cfg = shipped()
    cfg.update({"annihilation_enabled": False, "annihilation_probability": 7.0})
    P.validate(P.parse_params(cfg))


def test_method_codes():
    assert P.method_code("standard") == 0 and P.method_code("barnes-hut") == 1
    assert P.method_code("fast-multipole") == 2
    with pytest.raises(AmError):
        P.method_code("fmm")
