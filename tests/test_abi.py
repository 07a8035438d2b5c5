"""The C-ABI library loads and exports every symbol include/*.h declares; its host-only entry
points (initial conditions, dt rule, field, species tables) agree with the oracle.  No GPU."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import antimatter_amd as am
import oracle
from conftest import ROOT


def declared_symbols():
    syms = set()
    for h in glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(am_\w+)\s*\(", text, re.M):
            syms.add(m.group(1))
    return syms


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(am.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) > 20
    missing = [s for s in sorted(syms) if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding knows every declared entry point
    from antimatter_amd import params
    known = set(am.SIGNATURES) | set(params.SIGNATURES)
    assert not (syms - known), sorted(syms - known)


def test_initial_conditions_bit_identical_to_oracle():
    pops = {"antiproton": 300, "positron": 257, "electron": 129, "proton": 3}
    temps = {"antiproton": 10.0, "positron": 100.0, "electron": 4.0, "proton": 0.0}
    radii = (1e-3, 2e-3, 4e-3)
    pos, vel, sp = am.initial_conditions(pops, temps, radii, 17)
    np.testing.assert_array_equal(pos, oracle.set_particle_positions(pops, radii, 17))
    np.testing.assert_array_equal(vel, oracle.set_particle_velocities(pops, temps, 17))
    np.testing.assert_array_equal(sp, oracle.species_array(pops))
    assert np.all(vel[-3:] == 0)  # T = 0: at rest


def test_dt_rule_and_field():
    pops = {"positron": 100, "electron": 100}
    assert am.derive_dt(pops, 1.0) == oracle.derive_dt(pops, 1.0)
    assert am.derive_dt(pops, 1.0) == pytest.approx(8.930967e-12, rel=1e-6)
    assert am.derive_dt({"antiproton": 1, "positron": 1}, 3.0) == pytest.approx(2.977e-12, rel=1e-3)
    assert am.derive_dt({"antihydrogen": 5}, 1.0) == pytest.approx(1e-6 * np.pi / 2)  # ratio default 1e-6
    t = am.make_trap(1.0, 10.0, 0.01, 0.01)
    assert t.d == oracle.characteristic_dimension(0.01, 0.01)
    with pytest.raises(am.AmError):
        am.make_trap(1.0, 10.0, 0.0, 0.0)


def test_species_tables_and_pairs():
    for s in range(6):
        assert am.species_properties(s) == oracle.species_props(s)
    pairs = {(a, b) for a in range(6) for b in range(6) if am.is_antiparticle_pair(a, b)}
    assert pairs == {(0, 4), (4, 0), (1, 3), (3, 1)}
    m, q = am.species_properties(3)
    assert am.cooling_time(abs(q), m, 1.0) == oracle.cooling_time(abs(q), m, 1.0)
    assert am.cooling_time(0.0, m, 1.0) == float("inf")


def test_invalid_initial_conditions_raise():
    with pytest.raises(am.AmError):
        am.initial_conditions({"positron": 2}, {"positron": 1.0}, (0.0, 1e-3, 1e-3), 1)
    with pytest.raises(am.AmError):
        am.initial_conditions({"positron": 0}, {}, (1e-3, 1e-3, 1e-3), 1)
