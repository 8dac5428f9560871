"""Validated run inputs.

``ReferenceCase`` keeps the reference's legacy-fixture request and its validation
(``python/monki/input/__init__.py:8-57``); the HIP backends prepare it on the host with
``monki.preparation`` from the fixture data shipped under ``monki/data/upstream``.
``SceneCase`` is the same request for a scene whose transport tables are already prepared in
memory (``monki.scene.TransportScene``), such as the synthetic benchmark scenes.
"""

from dataclasses import dataclass
from math import isfinite
from pathlib import Path
from typing import Tuple


def _check_counts(case) -> None:
    # input/__init__.py:30-37 — the legacy count fields are 32-bit.
    for name in ("photons_per_column", "rounds"):
        value = getattr(case, name)

        if type(value) is not int or not 0 < value < 2**31:
            raise ValueError(f"{name} must be a positive 32-bit integer")


def _check_name(name: str) -> None:
    if not name or not name.replace("-", "").replace("_", "").isalnum():
        raise ValueError("case name must contain only letters, numbers, hyphens, underscores")


@dataclass(frozen=True)
class ReferenceCase:
    """Specify a supported legacy fixture plus explicit Monte Carlo and geometry settings."""

    name: str
    input_file: str
    photons_per_column: int = 200
    rounds: int = 32
    surface_albedo: float = 0.05
    solar_zenith: float = 45.0
    solar_azimuth: float = 0.0
    view_azimuth: float = 0.0
    view_zeniths: Tuple[float, ...] = (0.0, 30.0, 60.0)
    backward: bool = False

    def __post_init__(self):
        _check_name(self.name)
        path = Path(self.input_file)

        if path.is_absolute() or ".." in path.parts:
            raise ValueError("input_file must be relative to the vendored INPUT directory")

        _check_counts(self)

        if not isfinite(self.surface_albedo) or not 0 <= self.surface_albedo <= 1:
            raise ValueError("surface_albedo must be in [0, 1]")

        if not self.view_zeniths or len(self.view_zeniths) > 91:
            raise ValueError("supply 1 to 91 viewing zenith angles")

        for angle in (self.solar_zenith, *self.view_zeniths):
            if not isfinite(angle) or not 0 <= angle < 90:
                raise ValueError("reference harness currently supports zeniths in [0, 90)")

        for angle in (self.solar_azimuth, self.view_azimuth):
            if not isfinite(angle) or not 0 <= angle < 360:
                raise ValueError("azimuths must be in [0, 360)")

        if type(self.backward) is not bool:
            raise ValueError("backward must be a bool")


@dataclass(frozen=True)
class SceneCase:
    """A Monte Carlo request over prepared transport tables.

    ``scene`` is a ``monki.scene.TransportScene``; geometry (solar and view angles) and
    surface albedo live in the scene, so only the photon budget is set here.
    """

    name: str
    scene: object
    photons_per_column: int = 200
    rounds: int = 32
    backward: bool = False

    def __post_init__(self):
        _check_name(self.name)
        _check_counts(self)

        if type(self.backward) is not bool:
            raise ValueError("backward must be a bool")

        validate = getattr(self.scene, "validate", None)

        if validate is None:
            raise TypeError("SceneCase.scene must be a TransportScene")

        validate()
