"""Prepared transport tables for one monochromatic scene.

``TransportScene`` holds exactly what the reference's preparation hands to photon
transport, in float64 and Fortran axis order: the contents of the ``transport.bin`` v1
export written by ``kernels/cpu/worker.f90:1102-1120`` (geometry, layer thicknesses, view
angles, surface tables, extinction ``BETA``, normalized extinction per species, gas
single-scattering albedo, and per scattering species its albedo and tabulated phase
matrix with the bound CDF of ``main_routines_general_mod.f90:822-866``).

Several tabulated species are allowed (the reference Metal path accepts one,
``metal_tables.py:29-36``; the Fortran transport accepts ``n_types``). Layers are indexed
bottom-up and include the vacuum layer at the top (``main_routines_cartesian_mod.f90:82``).
"""

from dataclasses import dataclass, field
from pathlib import Path
from typing import List, Optional, Tuple

import numpy as np

VACUUM_THICKNESS_KM = 1e-9  # main_routines_cartesian_mod.f90:459 adds a 1e-6 m layer


@dataclass
class Species:
    """One tabulated scatterer: albedo and phase-matrix elements on ascending cosines.

    ``cosines`` ascend from -1 to 1 (``main_routines_general_mod.f90:775-777`` reverses the
    LUT's angle order). ``f11`` is normalized so that its mean over the sphere is one; the
    /4 convention of ``mie_scat_matrix`` is applied by transport, not here.
    """

    name: str
    ssalb: float
    cosines: np.ndarray
    f11: np.ndarray
    f22: np.ndarray
    f33: np.ndarray
    f12: np.ndarray
    f34: Optional[np.ndarray] = None
    f44: Optional[np.ndarray] = None
    cdf: Optional[np.ndarray] = None

    def __post_init__(self):
        self.cosines = np.asarray(self.cosines, dtype=np.float64)

        for name in ("f11", "f22", "f33", "f12"):
            setattr(self, name, np.asarray(getattr(self, name), dtype=np.float64))

        if self.f34 is None:
            self.f34 = np.zeros_like(self.f11)

        if self.f44 is None:
            self.f44 = self.f33.copy()

    def bound_cdf(self, nmat: int) -> np.ndarray:
        """Trapezoid CDF of the rejection bound F11 + |F12| (F11 alone for nmat = 1).

        Restates ``main_routines_general_mod.f90:822-866``: cumulative trapezoid integral
        over ascending cosines, divided by the total.
        """

        bound = self.f11 + (np.abs(self.f12) if nmat > 1 else 0.0)
        steps = np.diff(self.cosines) * (bound[1:] + bound[:-1]) / 2.0
        cdf = np.concatenate([[0.0], np.cumsum(steps)])
        return cdf / cdf[-1]


@dataclass
class TransportScene:
    """Float64 transport tables of one scene (``transport.bin`` v1 semantics)."""

    nx: int
    ny: int
    nz: int  # including the vacuum layer
    nmat: int
    xunit: float  # km
    yunit: float  # km
    zunit: np.ndarray  # (nz,) km, bottom-up, last entry is the vacuum layer
    beta: np.ndarray  # (nx, ny, nz) 1/km
    fractions: np.ndarray  # (nspecies + 1, nx, ny, nz), gas last
    ssalb_gas: np.ndarray  # (nz,)
    albedo: np.ndarray  # (nx, ny)
    depol: float
    solar_zenith: float  # degrees, as in ReferenceCase
    solar_azimuth: float
    view_azimuth: float
    view_zeniths: Tuple[float, ...]
    species: List[Species] = field(default_factory=list)
    save_path: bool = True
    # Single/multiple-scattering decomposition (calc_radiances_and_scats); excludes save_path.
    save_scats: bool = False
    # Detector and output grid: "toa" (views 0 <= zenith < 90, looking down from TOA) or
    # "surface" (views 90 < zenith <= 180: the direction toward a detector at the ground),
    # main.f90:711-745 with the result grid at the detector.
    detector: str = "toa"
    # Result grid: None (at the detector), "toa" or "surface" (main.f90:747-757); a grid away
    # from the detector registers each contribution at the column where the line of sight
    # through the event meets the grid's level (mc_routines_cartesian_mod.f90:1041-1105).
    register: Optional[str] = None
    wavelength_nm: float = 340.0
    altitude_m: Optional[np.ndarray] = None  # (nz,) level altitudes, bottom to TOA
    description: str = ""
    # Optional factored extinction of a scene with one tabulated species: float32 arrays
    # (species extinction (nx, ny, nz), gas extinction (nz,)) from which the float32 tables
    # follow exactly: beta = species + gas[k] and the species fraction = species / beta, both in
    # float32 arithmetic. Scenes of a batch that share the species field (a cloud seen at several
    # wavelengths) then share one device copy of it.
    extinction_parts: Optional[Tuple[np.ndarray, np.ndarray]] = None

    def __post_init__(self):
        self.zunit = np.asarray(self.zunit, dtype=np.float64)
        self.beta = np.asarray(self.beta, dtype=np.float64)
        self.fractions = np.asarray(self.fractions, dtype=np.float64)
        self.ssalb_gas = np.asarray(self.ssalb_gas, dtype=np.float64)
        self.albedo = np.asarray(self.albedo, dtype=np.float64)
        self.view_zeniths = tuple(float(v) for v in self.view_zeniths)

        if self.altitude_m is None:
            self.altitude_m = 1000.0 * np.concatenate([[0.0], np.cumsum(self.zunit[:-1])])

        self.altitude_m = np.asarray(self.altitude_m, dtype=np.float64)

    @property
    def angles(self) -> int:
        return len(self.view_zeniths)

    @property
    def propagation_zenith(self) -> float:
        """Zenith of the propagation direction of sunlight (``worker.f90:789``)."""

        return 180.0 - self.solar_zenith

    def validate(self) -> None:
        """Reject inconsistent tables before any device or native call."""

        nx, ny, nz = self.nx, self.ny, self.nz

        for name, value in (("nx", nx), ("ny", ny), ("nz", nz), ("nmat", self.nmat)):
            if type(value) is not int or value < 1:
                raise ValueError(f"{name} must be a positive integer")

        if nz < 2:
            raise ValueError("nz must include at least one physical layer and the vacuum layer")

        if self.nmat not in (1, 3, 4):
            raise ValueError("nmat must be 1 (I), 3 (I, Q, U) or 4 (I, Q, U, V)")

        shapes = {
            "zunit": (self.zunit, (nz,)),
            "beta": (self.beta, (nx, ny, nz)),
            "fractions": (self.fractions, (len(self.species) + 1, nx, ny, nz)),
            "ssalb_gas": (self.ssalb_gas, (nz,)),
            "albedo": (self.albedo, (nx, ny)),
            "altitude_m": (self.altitude_m, (nz,)),
        }

        for name, (value, shape) in shapes.items():
            if value.shape != shape:
                raise ValueError(f"{name} has shape {value.shape}, expected {shape}")

            if not np.all(np.isfinite(value)):
                raise ValueError(f"{name} is not finite")

        if np.any(self.zunit <= 0) or np.any(self.beta < 0) or np.any(self.fractions < 0):
            raise ValueError("Invalid layer thicknesses, extinction or species fractions")

        if np.any(self.beta[:, :, -1] != 0):
            raise ValueError("the vacuum layer must have zero extinction")

        total = self.fractions.sum(axis=0)
        scattering = self.beta > 1e-16

        if not np.allclose(total[scattering], 1.0, rtol=0, atol=1e-9):
            raise ValueError("species fractions must sum to one where extinction is positive")

        if np.any((self.ssalb_gas < 0) | (self.ssalb_gas > 1)):
            raise ValueError("gas single-scattering albedo must be in [0, 1]")

        if np.any((self.albedo < 0) | (self.albedo > 1)):
            raise ValueError("surface albedo must be in [0, 1]")

        if self.extinction_parts is not None:
            self.check_extinction_parts()

        if not 0 <= self.depol < 0.5:
            raise ValueError("depolarization factor must be in [0, 0.5)")

        if not self.view_zeniths or len(self.view_zeniths) > 3:
            raise ValueError("supply 1 to 3 viewing zenith angles")

        if self.detector not in ("toa", "surface"):
            raise ValueError('detector must be "toa" or "surface"')

        if self.register not in (None, "toa", "surface"):
            raise ValueError('register must be None, "toa" or "surface"')

        if not np.isfinite(self.solar_zenith) or not 0 <= self.solar_zenith < 90:
            raise ValueError("zeniths must be in [0, 90)")

        for angle in self.view_zeniths:
            if self.detector == "toa" and (not np.isfinite(angle) or not 0 <= angle < 90):
                raise ValueError("zeniths must be in [0, 90)")

            if self.detector == "surface" and (not np.isfinite(angle) or not 90 < angle <= 180):
                raise ValueError("view zeniths of a surface detector must be in (90, 180]")

        for angle in (self.solar_azimuth, self.view_azimuth):
            if not np.isfinite(angle) or not 0 <= angle < 360:
                raise ValueError("azimuths must be in [0, 360)")

        if self.xunit <= 0 or self.yunit <= 0:
            raise ValueError("cell sizes must be positive")

        for sp in self.species:
            n = sp.cosines.size

            if n < 2 or np.any(np.diff(sp.cosines) <= 0):
                raise ValueError(f"{sp.name}: cosines must ascend strictly")

            if abs(sp.cosines[0] + 1) > 1e-12 or abs(sp.cosines[-1] - 1) > 1e-12:
                raise ValueError(f"{sp.name}: cosines must span [-1, 1]")

            for name in ("f11", "f22", "f33", "f12", "f34", "f44"):
                if getattr(sp, name).shape != (n,):
                    raise ValueError(f"{sp.name}: {name} must match the cosine grid")

            if np.any(sp.f11 <= 0) or np.any(np.abs(sp.f12) > sp.f11 * (1 + 1e-9)):
                raise ValueError(f"{sp.name}: F11 must be positive and |F12| <= F11")

            if not 0 <= sp.ssalb <= 1:
                raise ValueError(f"{sp.name}: single-scattering albedo must be in [0, 1]")

    def check_extinction_parts(self) -> None:
        """The factored extinction must give back the float32 tables bit for bit."""

        if len(self.species) != 1:
            raise ValueError("factored extinction needs exactly one tabulated species")

        species, gas = (np.asarray(part) for part in self.extinction_parts)

        if species.dtype != np.float32 or gas.dtype != np.float32:
            raise ValueError("extinction parts must be float32 arrays")

        if species.shape != (self.nx, self.ny, self.nz) or gas.shape != (self.nz,):
            raise ValueError("extinction parts must be shaped (nx, ny, nz) and (nz,)")

        beta = species + gas[None, None, :]

        if not np.array_equal(beta, self.beta.astype(np.float32)):
            raise ValueError("extinction parts do not reproduce the extinction table in float32")

        positive = beta > 0

        with np.errstate(invalid="ignore", divide="ignore"):
            fraction = species / beta

        if not np.array_equal(fraction[positive], self.fractions[0].astype(np.float32)[positive]):
            raise ValueError("extinction parts do not reproduce the species fraction in float32")

    def varying_layers(self) -> np.ndarray:
        """Layers whose extinction or species mix changes across columns."""

        flat_beta = self.beta.reshape(self.nx * self.ny, self.nz)
        flat_frac = self.fractions.reshape(-1, self.nx * self.ny, self.nz)
        beta_varies = np.any(flat_beta != flat_beta[:1], axis=0)
        frac_varies = np.any(np.any(flat_frac != flat_frac[:, :1], axis=1), axis=0)
        return np.flatnonzero(beta_varies | frac_varies)

    def tally_shape(self) -> dict:
        return {
            "stokes": (self.nmat, self.nx, self.ny, self.angles),
            "paths": (self.nx, self.ny, self.nz, self.angles),
            "weights": (self.nx, self.ny, self.angles),
        }

    # -- transport.bin v1 -------------------------------------------------------------

    def transport_bin(self, backward: bool = False) -> np.ndarray:
        """Return the float64 stream ``worker.f90:1102-1120`` writes, in the same order.

        With nmat 4 each species' block continues with F34 and F44 (the worker's export stops
        at F12 because the reference GPU path takes nmat 1 and 3 only).
        """

        nz = self.nz
        header = [
            1,
            self.nx,
            self.ny,
            nz,
            self.nmat,
            self.angles,
            len(self.species),
            nz,  # k_source: TOA
            nz if self.detector == "toa" else 0,  # k_detector
            nz if (self.register or self.detector) == "toa" else 0,  # k_register
            int(self.save_path),
            int(backward),  # transport direction (mc.f90 backward_mode)
        ]
        parts = [
            np.asarray(header, dtype=np.float64),
            np.array(
                [
                    self.xunit,
                    self.yunit,
                    self.depol,
                    self.propagation_zenith,
                    self.solar_azimuth,
                    self.view_azimuth,
                ]
            ),
            self.zunit,
            np.asarray(self.view_zeniths, dtype=np.float64),
            np.full(self.nx * self.ny, -1.0),  # Lambertian surface type
            self.albedo.ravel(order="F"),
            np.ones(self.nx * self.ny),  # flat surface elevation index
            self.beta.ravel(order="F"),
            self.fractions.ravel(order="F"),
            self.ssalb_gas,
        ]

        for sp in self.species:
            cdf = sp.bound_cdf(self.nmat) if sp.cdf is None else sp.cdf
            parts += [
                np.array([sp.cosines.size, sp.ssalb], dtype=np.float64),
                sp.cosines,
                cdf,
                sp.f11,
                sp.f22,
                sp.f33,
                sp.f12,
            ]

            if self.nmat == 4:
                parts += [sp.f34, sp.f44]

        return np.concatenate(parts).astype(np.float64)

    def write_transport_bin(self, path, backward: bool = False) -> Path:
        path = Path(path)
        self.transport_bin(backward).tofile(path)
        return path

    @classmethod
    def from_transport_bin(cls, raw: np.ndarray, **metadata) -> "TransportScene":
        """Parse a ``transport.bin`` v1 stream (inverse of ``transport_bin``)."""

        raw = np.asarray(raw, dtype=np.float64)

        if raw.ndim != 1 or raw.size < 18 or raw[0] != 1:
            raise ValueError("Unsupported transport export version")

        nx, ny, nz, nmat, angles, nspecies = (int(v) for v in raw[1:7])

        if raw[11] != 0 or raw[7] != nz or raw[8] not in (0, nz) or raw[9] not in (0, nz):
            raise NotImplementedError("only forward exports with a TOA source are read")

        metadata.setdefault("detector", "toa" if raw[8] == nz else "surface")
        metadata.setdefault("register", None if raw[9] == raw[8] else ("toa" if raw[9] == nz else "surface"))

        offset = 18

        def take(count):
            nonlocal offset
            out = raw[offset : offset + count]
            offset += count

            if out.size != count:
                raise ValueError("Truncated transport export")

            return out

        xunit, yunit, depol, vartheta0, phi0, phi = raw[12:18]
        zunit = take(nz)
        views = take(angles)
        take(nx * ny)
        albedo = take(nx * ny).reshape((nx, ny), order="F")
        take(nx * ny)
        beta = take(nx * ny * nz).reshape((nx, ny, nz), order="F")
        fractions = take((nspecies + 1) * nx * ny * nz).reshape(
            (nspecies + 1, nx, ny, nz), order="F"
        )
        ssalb_gas = take(nz)
        species = []

        for index in range(nspecies):
            size, ssalb = take(2)
            size = int(size)
            cosines, cdf, f11, f22, f33, f12 = take(6 * size).reshape(6, size)
            f34, f44 = take(2 * size).reshape(2, size) if nmat == 4 else (None, None)
            species.append(
                Species(f"species{index + 1}", float(ssalb), cosines, f11, f22, f33, f12, f34, f44, cdf=cdf)
            )

        if offset != raw.size:
            raise ValueError("Trailing data in transport export")

        return cls(
            nx=nx,
            ny=ny,
            nz=nz,
            nmat=nmat,
            xunit=float(xunit),
            yunit=float(yunit),
            zunit=zunit.copy(),
            beta=beta.copy(),
            fractions=fractions.copy(),
            ssalb_gas=ssalb_gas.copy(),
            albedo=albedo.copy(),
            depol=float(depol),
            solar_zenith=180.0 - float(vartheta0),
            solar_azimuth=float(phi0),
            view_azimuth=float(phi),
            view_zeniths=tuple(views),
            species=species,
            save_path=bool(raw[10]),
            **metadata,
        )
