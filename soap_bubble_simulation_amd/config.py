"""Simulation configuration (mirror of ``src/config.rs``).

``SimulationConfig`` carries the same fields, defaults and JSON layout as the
reference's serde structs (config.rs:12-316, ``config/default.json``), so a
reference config file loads unchanged; ``to_params`` maps it onto the flat
``sf_params`` struct of the C ABI the way ``DrainageSimulator::new`` reads it
(drainage.rs:263-288).
"""
from __future__ import annotations

import json
from dataclasses import asdict, dataclass, field, fields
from typing import Any, Dict


class ConfigError(Exception):
    """Io / Parse / Serialize errors of config.rs:374-425."""

    def __init__(self, kind: str, path: str, error: Exception | str):
        self.kind, self.path, self.error = kind, path, error
        if kind == "Io":
            msg = "Failed to read/write config file '%s': %s" % (path, error)
        elif kind == "Parse":
            msg = "Failed to parse config file '%s': %s" % (path, error)
        else:
            msg = "Failed to serialize config: %s" % (error,)
        super().__init__(msg)


@dataclass
class BubbleParameters:                     # config.rs:12-35
    diameter: float = 0.05
    film_thickness_nm: float = 500.0
    critical_thickness_nm: float = 10.0
    refractive_index: float = 1.33


@dataclass
class FluidParameters:                      # config.rs:38-67
    viscosity: float = 0.001
    surface_tension: float = 0.025
    density: float = 1000.0
    surfactant_diffusion: float = 1e-9
    surfactant_concentration: float = 0.5


@dataclass
class EnvironmentParameters:                # config.rs:70-105
    gravity: float = 9.81
    temperature: float = 293.15
    pressure: float = 101325.0
    wind: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    buoyancy: float = 1.0


@dataclass
class FoamParameters:                       # config.rs:108-278
    enabled: bool = False
    max_bubbles: int = 64
    initial_count: int = 5
    van_der_waals_strength: float = 1e-12
    coalescence_enabled: bool = False
    plateau_stiffness: float = 1.0
    min_radius: float = 0.005
    max_radius: float = 0.1
    time_scale: float = 1.0
    positioning_mode: int = 0
    spacing: float = 0.05
    jitter: float = 0.1
    size_distribution: int = 0
    mean_radius: float = 0.022
    std_dev: float = 0.005
    sigma: float = 0.3
    pdi: float = 1.5
    bimodal_ratio: float = 0.5
    bimodal_mean2: float = 0.028
    bimodal_std2: float = 0.003


_REQUIRED = {
    BubbleParameters: ("diameter", "film_thickness_nm", "critical_thickness_nm", "refractive_index"),
    FluidParameters: ("viscosity", "surface_tension", "density", "surfactant_diffusion",
                      "surfactant_concentration"),
    EnvironmentParameters: ("gravity", "temperature", "pressure"),
    FoamParameters: (),
}


def _build(cls, obj: Any, where: str):
    if not isinstance(obj, dict):
        raise ValueError("%s: expected an object" % where)
    names = {f.name for f in fields(cls)}
    for k in _REQUIRED[cls]:
        if k not in obj:
            raise ValueError("%s: missing field `%s`" % (where, k))
    kw = {k: v for k, v in obj.items() if k in names}
    for k, v in kw.items():
        default = getattr(cls(), k)
        if isinstance(default, bool):
            if not isinstance(v, bool):
                raise ValueError("%s.%s: expected a boolean" % (where, k))
        elif isinstance(default, int) and not isinstance(default, bool):
            if not isinstance(v, int) or isinstance(v, bool):
                raise ValueError("%s.%s: expected an integer" % (where, k))
        elif isinstance(default, float):
            if not isinstance(v, (int, float)) or isinstance(v, bool):
                raise ValueError("%s.%s: expected a number" % (where, k))
            kw[k] = float(v)
    return cls(**kw)


@dataclass
class SimulationConfig:                     # config.rs:283-316
    bubble: BubbleParameters = field(default_factory=BubbleParameters)
    fluid: FluidParameters = field(default_factory=FluidParameters)
    environment: EnvironmentParameters = field(default_factory=EnvironmentParameters)
    foam: FoamParameters = field(default_factory=FoamParameters)
    dt: float = 0.001
    resolution: int = 128

    # config.rs:356-370
    def bubble_radius(self) -> float:
        return self.bubble.diameter / 2.0

    def internal_pressure_difference(self) -> float:
        return 4.0 * self.fluid.surface_tension / self.bubble_radius()

    def film_thickness_meters(self) -> float:
        return self.bubble.film_thickness_nm * 1e-9

    @classmethod
    def from_dict(cls, d: Dict[str, Any]) -> "SimulationConfig":
        for k in ("bubble", "fluid", "environment", "dt", "resolution"):
            if k not in d:
                raise ValueError("missing field `%s`" % k)
        res = d["resolution"]
        if not isinstance(res, int) or isinstance(res, bool) or res < 0 or res > 0xFFFFFFFF:
            raise ValueError("resolution: expected a u32")
        if not isinstance(d["dt"], (int, float)) or isinstance(d["dt"], bool):
            raise ValueError("dt: expected a number")
        return cls(
            bubble=_build(BubbleParameters, d["bubble"], "bubble"),
            fluid=_build(FluidParameters, d["fluid"], "fluid"),
            environment=_build(EnvironmentParameters, d["environment"], "environment"),
            foam=_build(FoamParameters, d.get("foam", {}), "foam"),
            dt=float(d["dt"]),
            resolution=res,
        )

    @classmethod
    def from_file(cls, path: str) -> "SimulationConfig":   # config.rs:328-337
        try:
            with open(path, "r") as f:
                text = f.read()
        except OSError as e:
            raise ConfigError("Io", path, e) from e
        try:
            return cls.from_dict(json.loads(text))
        except (ValueError, TypeError) as e:
            raise ConfigError("Parse", path, e) from e

    def to_file(self, path: str) -> None:                  # config.rs:346-354
        try:
            text = json.dumps(asdict(self), indent=2)
        except (TypeError, ValueError) as e:
            raise ConfigError("Serialize", path, e) from e
        try:
            with open(path, "w") as f:
                f.write(text)
        except OSError as e:
            raise ConfigError("Io", path, e) from e

    def to_params(self):
        """sf_params of DrainageSimulator::new(&self) (drainage.rs:263-288)."""
        from . import _lib as L
        p = L.Params()
        L.check(L.lib().sf_params_cpu_default(p))
        p.density = self.fluid.density
        p.gravity = self.environment.gravity
        p.viscosity = self.fluid.viscosity
        p.diffusion_coeff = self.fluid.surfactant_diffusion
        p.critical_thickness = self.bubble.critical_thickness_nm * 1e-9
        p.bubble_radius = self.bubble_radius()
        p.dt = self.dt
        return p
