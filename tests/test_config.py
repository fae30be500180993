"""SimulationConfig mirror (config.rs:283-465): defaults, JSON round trip,
error kinds, and the sf_params mapping of DrainageSimulator::new."""
import json
import os

import pytest

from soap_bubble_simulation_amd import _lib as L
from soap_bubble_simulation_amd.config import ConfigError, SimulationConfig

EPS = 2.220446049250313e-16

REFERENCE_DEFAULT_JSON = {       # the layout of config/default.json
    "bubble": {"diameter": 0.05, "film_thickness_nm": 500.0, "critical_thickness_nm": 10.0,
               "refractive_index": 1.33},
    "fluid": {"viscosity": 0.001, "surface_tension": 0.025, "density": 1000.0,
              "surfactant_diffusion": 1e-9, "surfactant_concentration": 0.5},
    "environment": {"gravity": 9.81, "temperature": 293.15, "pressure": 101325.0},
    "dt": 0.001,
    "resolution": 128,
}


def test_reference_config_asserts():
    c = SimulationConfig()
    assert abs(c.bubble.diameter - 0.05) < EPS                      # config.rs:432-437
    assert abs(c.bubble.film_thickness_nm - 500.0) < EPS
    assert abs(c.fluid.viscosity - 0.001) < EPS
    assert abs(c.environment.gravity - 9.81) < EPS
    assert abs(c.bubble_radius() - 0.025) < EPS                     # 440-443
    assert abs(c.internal_pressure_difference() - 4.0) < 1e-10     # 446-450
    assert abs(c.film_thickness_meters() - 500e-9) < EPS           # 453-456


def test_roundtrip(tmp_path):
    c = SimulationConfig()
    c.bubble.diameter = 0.031
    c.resolution = 64
    path = tmp_path / "cfg.json"
    c.to_file(str(path))
    d = SimulationConfig.from_file(str(path))
    assert abs(d.bubble.diameter - 0.031) < EPS and d.resolution == 64   # 459-464


def test_reference_default_json(tmp_path):
    path = tmp_path / "default.json"
    path.write_text(json.dumps(REFERENCE_DEFAULT_JSON))
    c = SimulationConfig.from_file(str(path))
    assert c == SimulationConfig()          # missing wind/buoyancy/foam take serde defaults


def test_errors(tmp_path):
    with pytest.raises(ConfigError) as e:
        SimulationConfig.from_file(str(tmp_path / "missing.json"))
    assert e.value.kind == "Io"
    bad = tmp_path / "bad.json"
    bad.write_text("{not json")
    with pytest.raises(ConfigError) as e:
        SimulationConfig.from_file(str(bad))
    assert e.value.kind == "Parse"
    partial = dict(REFERENCE_DEFAULT_JSON)
    del partial["resolution"]
    bad.write_text(json.dumps(partial))
    with pytest.raises(ConfigError) as e:
        SimulationConfig.from_file(str(bad))
    assert e.value.kind == "Parse"


def test_to_params():
    c = SimulationConfig()
    c.fluid.viscosity = 0.004
    c.bubble.critical_thickness_nm = 20.0
    c.bubble.diameter = 0.02
    p = c.to_params()
    assert p.viscosity == 0.004 and p.critical_thickness == 20.0 * 1e-9
    assert p.bubble_radius == 0.01 and p.diffusion_coeff == c.fluid.surfactant_diffusion
    assert p.flags & L.FLAG_TFE
