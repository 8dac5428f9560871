"""Optical preparation of the reference's positional fixtures, in float64 on the host.

The reference prepares a ``ReferenceCase`` by running its Fortran worker up to the export of
``transport.bin`` (``kernels/cpu/worker.f90:413-1120``; the Metal backend reuses that CPU
preparer, ``python/monki/execution/metal.py:98-104``). This module restates that preparation
so the HIP backend needs no Fortran:

- ``read_fixture``: ``rdsize`` and ``rdinput`` (``main_routines_cartesian_mod.f90:32-487``),
  the positional input reader, with its list-directed record semantics;
- ``molecular_optical_depths``: ``molecules_earth`` (``main_routines_general_mod.f90:33-321``)
  with the ozone cross sections of ``get_O3_crabs`` (``get_O3_crabs.f90:15-157``) and the
  linear interpolation ``linpol_dp_single`` (``interpol_sp.f:105-152``);
- ``scatterer_tables``: ``aerosols`` (``main_routines_general_mod.f90:629-951``), the Mie table
  reader and bound CDF;
- ``prepare_fixture``: the extinction, species fractions and gas albedo of
  ``kernels/cpu/worker.f90:933-1010``, returned as a ``TransportScene``.

Arithmetic follows the Fortran statement by statement (same operand order, scalar libm
``log``/``cos``), so the tables equal the reference's to the last bits on the same inputs.
"""

import math
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, List, Tuple

import numpy as np

from .scene import Species, TransportScene

# physical_parameters_mod.f90:16-21
PI = 3.141592653589793
DEG_TO_RAD = 1.7453292519943295e-2
AVOGADRO = 6.022137e23
LOSCHMIDT = 2.5468993e25
GAS_CONSTANT = 8.314510

EPS = 1.0e-16  # main_routines_cartesian_mod.f90:264, worker.f90:221

# main_routines_general_mod.f90:68-76: dry-air composition (volume percent) and molar masses.
FRACTION = {"N2": 78.084, "O2": 20.9476, "Ar": 0.934, "CO2": 0.0375}
MOLAR_MASS = {"N2": 28.0134, "O2": 32.00, "Ar": 39.948, "CO2": 44.01}
GRAVITY = 9.807

# python/monki/reference.py:27-37: positional markers of the supported testrun layout.
LAYOUT_MARKERS = {
    4: "number of columns in horizontal x",
    5: "number of columns in horizontal y",
    6: "number of layers",
    16: "number of photons per column per round",
    17: "maximum number of rounds",
    18: "stop simulation when result has converged",
    19: "maximum allowed reflectance standard deviation",
    24: "size of Stokes vector",
    38: "save the layer AMFs",
    39: "save the single",
}


class _Records:
    """Sequential records of a text file, read as Fortran ``READ(unit, *)`` does.

    A list-directed read starts at the next record and takes values separated by blanks or
    commas, continuing into further records until it has as many as it asks for; the rest
    of its last record is skipped. ``text(width)`` is a formatted ``(A<width>)`` read.
    """

    def __init__(self, path: Path):
        self.path = path
        self.lines = Path(path).read_text().splitlines()
        self.index = 0

    def _next(self) -> str:
        if self.index >= len(self.lines):
            raise ValueError(f"{self.path}: unexpected end of file")

        line = self.lines[self.index]
        self.index += 1
        return line

    def skip(self, count: int = 1) -> None:
        for _ in range(count):
            self._next()

    def values(self, count: int) -> List[str]:
        tokens: List[str] = []

        while len(tokens) < count:
            line = self._next().split("/", 1)[0]
            tokens += line.replace(",", " ").split()

        return tokens[:count]

    def integer(self) -> int:
        return _integer(self.values(1)[0], self.path)

    def real(self) -> float:
        return _real(self.values(1)[0], self.path)

    def logical(self) -> bool:
        return _logical(self.values(1)[0], self.path)

    def word(self) -> str:
        return self.values(1)[0]

    def text(self, width: int) -> str:
        return self._next()[:width].strip()


def _integer(token: str, path) -> int:
    try:
        return int(token)
    except ValueError:
        raise ValueError(f"{path}: ERROR reading an integer from {token!r}") from None


def _real(token: str, path) -> float:
    try:
        return float(token.replace("D", "E").replace("d", "e"))
    except ValueError:
        raise ValueError(f"{path}: ERROR reading a real from {token!r}") from None


def _logical(token: str, path) -> bool:
    # A list-directed logical value is T or F, optionally after a period (".true.", "F").
    letter = token.lstrip(".")[:1].upper()

    if letter not in ("T", "F"):
        raise ValueError(f"{path}: ERROR reading a logical from {token!r}")

    return letter == "T"


@dataclass(frozen=True)
class FixtureInput:
    """Everything ``rdsize`` and ``rdinput`` read from one positional input file.

    ``nz`` includes the vacuum layer that ``rdsize`` appends; level arrays have ``nz + 1``
    entries, the last one the synthetic vacuum top (``main_routines_cartesian_mod.f90:457-461``).
    """

    nx: int
    ny: int
    nz: int
    x_dim: float  # km
    y_dim: float  # km
    source_location: int
    detector_location: int
    register_location: int
    photons_per_column: int
    rounds: int
    stop_converged: bool
    tol_error: float
    nmat: int
    wave_type: str
    wave_min: float
    wave_max: float
    dwave: float
    read_bmsca_from_file: bool
    bmsca_file: str
    gas_col_type: int
    background: str
    fix_depol: bool
    depol_fixed: float
    read_bmabs_from_file: bool
    bmabs_file: str
    cloud_input_type: str
    save_path: bool
    save_scats: bool
    nscatt_types: int
    coefile: Tuple[str, ...]
    coelut: Tuple[str, ...]
    n_scat_angles: int
    ngases: int
    nclouds: int
    alt: np.ndarray  # (nz + 1,) m
    pres: np.ndarray  # (nz + 1,) Pa
    temp: np.ndarray  # (nz + 1,) K
    vmrgas: np.ndarray  # (ngases, nz + 1) ppm
    cloud_abundance: np.ndarray  # (nscatt_types, nx, ny, nz): COT (baer) or column density (ncol)

    @property
    def wavelengths(self) -> int:
        """``worker.f90:428``: INT((wave_max - wave_min) / dwave) + 1."""

        return int((self.wave_max - self.wave_min) / self.dwave) + 1

    @property
    def wavelength_nm(self) -> float:
        """First (only) wavelength in nm (``worker.f90:840-858``)."""

        return 1.0e7 / self.wave_min if self.wave_type == "wn" else self.wave_min


def check_layout(path) -> List[str]:
    """Reject files whose positional layout differs from the supported testrun fixtures.

    Restates ``python/monki/reference.py:22-37``: the parser reads fixed line positions, so
    an older input format would otherwise shift the physics silently.
    """

    path = Path(path)

    if path.parent.name != "testrun":
        raise ValueError("Only the supplied testrun input format is supported")

    lines = path.read_text().splitlines()

    for index, marker in LAYOUT_MARKERS.items():
        if index >= len(lines) or marker not in lines[index]:
            raise ValueError(f"Unsupported positional input at line {index + 1}: {path}")

    return lines


def read_fixture(path) -> FixtureInput:
    """``rdsize`` then ``rdinput`` (``main_routines_cartesian_mod.f90:32-487``)."""

    path = Path(path)
    r = _Records(path)

    # rdsize: grid block.
    r.skip(4)
    nx, ny, nz = r.integer(), r.integer(), r.integer() + 1  # :124 adds the vacuum layer
    x_dim, y_dim = r.real(), r.real()
    source_location, detector_location, register_location = r.integer(), r.integer(), r.integer()
    # Monte Carlo block (read by rdinput).
    r.skip(4)
    photons_per_column, rounds = r.integer(), r.integer()
    stop_converged, tol_error = r.logical(), r.real()

    if rounds < 30 and stop_converged:
        raise ValueError("stop_converged needs at least 30 rounds (main_routines_cartesian_mod.f90:321-328)")

    # General block.
    r.skip(4)
    nmat, wave_type = r.integer(), r.word()[:2]
    wave_min, wave_max, dwave = r.real(), r.real(), r.real()
    read_bmsca_from_file, bmsca_file = r.logical(), r.text(23)
    gas_col_type, background = r.integer(), r.word()[:5]
    fix_depol, depol_fixed = r.logical(), r.real()
    read_bmabs_from_file, bmabs_file = r.logical(), r.text(23)
    cloud_input_type = r.word()[:4]
    save_path, save_scats = r.logical(), r.logical()
    # Aerosol block.
    r.skip(4)
    nscatt_types = r.integer()
    coefile, coelut = [], []

    if nscatt_types < 1:
        raise ValueError("nscatt_types should be 1 or larger (main_routines_cartesian_mod.f90:186-191)")

    for _ in range(nscatt_types):
        coefile.append(r.text(23))
        coelut.append(r.text(23))

    n_scat_angles = r.integer()
    # Layer block.
    r.skip(4)
    ngases, nclouds = r.integer(), r.integer()

    if nmat not in (1, 3, 4):
        raise ValueError("rdsize: nmat should be 1, 3 or 4")

    if ngases != 1:
        raise ValueError("ngases should be 1 (main_routines_cartesian_mod.f90:179-184)")

    r.skip(3)
    levels = nz + 1
    alt, pres, temp = np.zeros(levels), np.zeros(levels), np.zeros(levels)
    vmrgas = np.zeros((ngases, levels))
    clouds = np.zeros((nscatt_types, nx, ny, nz))

    for k in range(nz):
        tokens = r.values(4 + ngases + 6 * nclouds)
        _integer(tokens[0], path)
        xalt, xpres, temp[k] = (_real(t, path) for t in tokens[1:4])
        vmrgas[:, k] = [_real(t, path) for t in tokens[4 : 4 + ngases]]

        for c in range(nclouds):
            field = tokens[4 + ngases + 6 * c : 10 + ngases + 6 * c]
            value = _real(field[0], path)
            kind, ixmin, ixmax, iymin, iymax = (_integer(t, path) for t in field[1:])

            if kind < 0 or kind > nscatt_types:
                kind = 0

            if value > EPS:
                if kind == 0:
                    raise ValueError(f"{path}: cloud at level {k + 1} names no scatterer type")

                for name, low, high, limit in (("x", ixmin, ixmax, nx), ("y", iymin, iymax, ny)):
                    if not (1 <= low <= limit and 1 <= high <= limit):
                        raise ValueError(f"{path}: cloud {name} range outside 1..n{name} at level {k + 1}")

                clouds[kind - 1, ixmin - 1 : ixmax, iymin - 1 : iymax, k] += value

        pres[k] = xpres * 1.0e5
        alt[k] = xalt * 1.0e3

        if pres[k] < EPS or temp[k] < EPS:
            raise ValueError(f"{path}: pressure and temperature must be positive (level {k + 1})")

    # main_routines_cartesian_mod.f90:457-461: vacuum top just above the last level.
    pres[nz] = 0.0
    alt[nz] = alt[nz - 1] + 0.000001
    temp[nz] = temp[nz - 1] - 100.0

    for line in r.lines[r.index :]:
        first = line.split("/", 1)[0].replace(",", " ").split()

        if first:
            try:
                int(first[0])
            except ValueError:
                break

            raise ValueError(f"{path}: more lines than levels (nz + 1)")

    return FixtureInput(
        nx=nx,
        ny=ny,
        nz=nz,
        x_dim=x_dim,
        y_dim=y_dim,
        source_location=source_location,
        detector_location=detector_location,
        register_location=register_location,
        photons_per_column=photons_per_column,
        rounds=rounds,
        stop_converged=stop_converged,
        tol_error=tol_error,
        nmat=nmat,
        wave_type=wave_type,
        wave_min=wave_min,
        wave_max=wave_max,
        dwave=dwave,
        read_bmsca_from_file=read_bmsca_from_file,
        bmsca_file=bmsca_file,
        gas_col_type=gas_col_type,
        background=background,
        fix_depol=fix_depol,
        depol_fixed=depol_fixed,
        read_bmabs_from_file=read_bmabs_from_file,
        bmabs_file=bmabs_file,
        cloud_input_type=cloud_input_type,
        save_path=save_path,
        save_scats=save_scats,
        nscatt_types=nscatt_types,
        coefile=tuple(coefile),
        coelut=tuple(coelut),
        n_scat_angles=n_scat_angles,
        ngases=ngases,
        nclouds=nclouds,
        alt=alt,
        pres=pres,
        temp=temp,
        vmrgas=vmrgas,
        cloud_abundance=clouds,
    )


def interpolate_linear(x: np.ndarray, y: np.ndarray, xp: float) -> float:
    """``linpol_dp_single`` (``interpol_sp.f:105-152``): linear interpolation with linear
    extrapolation beyond either end; an exact node returns its value."""

    n = len(x)

    if xp < x[0]:
        return (y[1] - y[0]) / (x[1] - x[0]) * (xp - x[0]) + y[0]

    if xp > x[n - 1]:
        return (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]) * (xp - x[n - 1]) + y[n - 1]

    if xp == x[n - 1]:
        return float(y[n - 1])

    for i in range(n - 1):
        if xp == x[i]:
            return float(y[i])

        if x[i] < xp < x[i + 1]:
            return (y[i + 1] - y[i]) / (x[i + 1] - x[i]) * (xp - x[i]) + y[i]

    raise ValueError("interpolation node not found: abscissae must ascend")


def ozone_cross_sections(table_path, wavelength_nm: float, temp: np.ndarray) -> np.ndarray:
    """O3 absorption cross sections (cm^2/molecule) at every level (``get_O3_crabs.f90``).

    The table holds a0, a1, a2 per wavelength; sigma = a0 + a1 T + a2 T^2 with T in Celsius.
    """

    r = _Records(Path(table_path))
    comments, rows = r.integer(), r.integer()
    r.skip(comments)
    table = np.array([[_real(t, table_path) for t in r.values(4)] for _ in range(rows)])
    coefficients = [interpolate_linear(table[:, 0], table[:, 1 + i], wavelength_nm) for i in range(3)]
    out = np.zeros(len(temp))

    for level, kelvin in enumerate(temp):
        celsius = float(kelvin) - 273.15
        value = 0.0

        for power, coefficient in enumerate(coefficients):
            value = value + coefficient * _power(celsius, power)

        out[level] = value

    return out


def _power(x: float, n: int) -> float:
    # Fortran real**integer: x**0 is one, higher powers by repeated multiplication.
    result = 1.0 if n == 0 else x

    for _ in range(n - 1):
        result = result * x

    return result


def molecular_optical_depths(fixture: FixtureInput, wavelength_nm: float, data_dir) -> Tuple[np.ndarray, np.ndarray, float]:
    """Rayleigh scattering and ozone absorption optical depth per layer, and the
    depolarization factor (``molecules_earth``, ``main_routines_general_mod.f90:33-321``).

    Returns ``(bmsca, bmabs, depol)``; both depths are zero in the vacuum layer.
    """

    nz, pres, temp, alt = fixture.nz, fixture.pres, fixture.temp, fixture.alt
    nmol = np.zeros(nz)

    if fixture.gas_col_type == 1:
        # :96-116 hydrostatic equilibrium with the mean molar mass of dry air (kg/molecule).
        weighted = (
            MOLAR_MASS["N2"] * FRACTION["N2"]
            + MOLAR_MASS["O2"] * FRACTION["O2"]
            + MOLAR_MASS["Ar"] * FRACTION["Ar"]
            + MOLAR_MASS["CO2"] * FRACTION["CO2"]
        )
        total = FRACTION["N2"] + FRACTION["O2"] + FRACTION["Ar"] + FRACTION["CO2"]
        mass = weighted / total * 1.0e-3 / AVOGADRO

        for k in range(nz - 1):
            nmol[k] = (pres[k] - pres[k + 1]) / (mass * GRAVITY)
    elif fixture.gas_col_type == 0:
        # :118-140 ideal-gas column with p/T varying exponentially across the layer.
        for k in range(nz - 1):
            pt1 = pres[k] / temp[k]
            pt2 = pres[k + 1] / temp[k + 1]
            dz = alt[k + 1] - alt[k]

            if abs(pt1 - pt2) < EPS:
                pt = pt1 * dz
            else:
                pt = (pt2 - pt1) * dz / (math.log(pt2) - math.log(pt1))

            nmol[k] = AVOGADRO / GAS_CONSTANT * pt
    else:
        raise ValueError("Wrong gas_col_type (0 or 1)")

    wavmic = wavelength_nm / 1000.0
    # :162-175 refractive index of dry air at STP (Peck and Reader 1972 form).
    w2 = 1.0 / (wavmic * wavmic)
    ri = 1.0 + 1.0e-8 * 8060.51 + 1.0e-8 * (2480990.0 / (132.274 - w2)) + 1.0e-8 * (17455.7 / (39.32957 - w2))
    rindex = (ri * ri - 1.0) * (ri * ri - 1.0) / ((ri * ri + 2.0) * (ri * ri + 2.0))
    # :199-218 King factor of dry air (Bodhaine et al. 1999) and depolarization factor.
    sig = 1.0e3 / wavelength_nm
    sig2 = sig * sig
    king = (
        FRACTION["N2"] * (1.034 + 3.17e-4 * sig2)
        + FRACTION["O2"] * (1.096 + 1.385e-3 * sig2 + 1.448e-4 * (sig2 * sig2))
        + FRACTION["Ar"] * 1.0
        + FRACTION["CO2"] * 1.15
    )
    dep = king / (FRACTION["N2"] + FRACTION["O2"] + FRACTION["Ar"] + FRACTION["CO2"])
    depol = (6.0 * dep - 6.0) / (3.0 + 7.0 * dep)
    # :223-233 molecular scattering cross section (m^2) and layer optical depth.
    ssca = (
        24.0 * PI * PI * PI * rindex * dep
        / (wavmic * wavmic * wavmic * wavmic * LOSCHMIDT * LOSCHMIDT * 1.0e-24)
    )  # fmt: skip
    bmsca = np.array([ssca * n for n in nmol])
    # :249-293 ozone: cross sections at the levels, column integral across each layer.
    sigma = ozone_cross_sections(Path(data_dir) / "abs_data" / "O3T_BP.dat", wavelength_nm, temp)
    sigma = sigma / 100.0 / 100.0
    bmabs = np.zeros(nz)

    for k in range(nz - 1):
        pt1 = pres[k] / temp[k]
        pt2 = pres[k + 1] / temp[k + 1]
        dz = alt[k + 1] - alt[k]
        sigmu1 = fixture.vmrgas[0, k] / 1000000.0 * sigma[k]
        sigmu2 = fixture.vmrgas[0, k + 1] / 1000000.0 * sigma[k + 1]

        if abs(pt1 - pt2) < EPS:
            ptsigmu = 0.5 * pt1 * (sigmu1 + sigmu2) * dz
        else:
            ratio = math.log(pt2 / pt1)
            ptsigmu = (
                (pt2 - pt1) * dz * (sigmu1 - sigmu2) / ratio**2
                + pt2 * dz * sigmu2 / ratio
                - pt1 * dz * sigmu1 / ratio
            )  # fmt: skip

        bmabs[k] = 0.0 + AVOGADRO / GAS_CONSTANT * ptsigmu

    return bmsca, bmabs, depol


def _wavelength_tag(wavelength_nm: float) -> str:
    # main_routines_general_mod.f90:689-705: '00' + F4.3 or '0' + F5.3 of the wavelength in um.
    if 1000.0 <= wavelength_nm < 10000.0 or 100.0 <= wavelength_nm < 1000.0:
        return "0" + f"{wavelength_nm / 1000.0:.3f}"

    raise ValueError("aerosols: wavelength outside [100, 10000) nm has no Mie table name")


def _read_lut(path: Path, count: int) -> np.ndarray:
    r = _Records(path)
    return np.array([[_real(t, path) for t in r.values(8)] for _ in range(count)])


def _read_coefficients(path: Path) -> Tuple[float, float, float]:
    # :860-872: single-scattering albedo, extinction efficiency and geometrical cross section
    # from fixed columns 33-57 of lines 8, 10 and 12.
    lines = Path(path).read_text().splitlines()

    try:
        return tuple(_real(lines[i][32:57].strip(), path) for i in (7, 9, 11))
    except IndexError:
        raise ValueError(f"{path}: unexpected end of file") from None


def scatterer_tables(fixture: FixtureInput, wavelength_nm: float, data_dir) -> Tuple[List[Species], np.ndarray]:
    """Tabulated scatterers and their optical depth per cell (``aerosols``,
    ``main_routines_general_mod.f90:629-951``).

    Reads each type's scattering-matrix table (ascending angles, stored in ascending cosine)
    and expansion-coefficient header; a wavelength between whole nanometres interpolates the
    two neighbouring tables linearly. Returns the species and ``baer`` of shape
    ``(nscatt_types, nx, ny, nz)``.
    """

    data_dir = Path(data_dir)
    interpolate = math.fmod(wavelength_nm, 1.0) >= EPS

    if interpolate:
        low = float(int(wavelength_nm))
        weight = wavelength_nm - low
        tags = (_wavelength_tag(low), _wavelength_tag(low + 1.0))
    else:
        weight = 0.0
        tags = (_wavelength_tag(wavelength_nm),)

    species, depths = [], np.zeros_like(fixture.cloud_abundance)
    count = fixture.n_scat_angles

    for index in range(fixture.nscatt_types):
        tables = [_read_lut(data_dir / "Mie_data" / (fixture.coelut[index] + tag), count) for tag in tags]
        headers = [_read_coefficients(data_dir / "Mie_data" / (fixture.coefile[index] + tag)) for tag in tags]

        if interpolate:
            lut = tables[0].copy()
            lut[:, 1:7] = (1.0 - weight) * tables[0][:, 1:7] + weight * tables[1][:, 1:7]
            lut[:, 0] = tables[1][:, 0]
            ssalb, q_ext, geometric = ((1.0 - weight) * a + weight * b for a, b in zip(*headers))
        else:
            lut = tables[0]
            ssalb, q_ext, geometric = headers[0]

        # :776-785: reverse into ascending cosine; columns angle F11 F22 F33 F44 F12 F34.
        lut = lut[::-1]
        cosines = np.array([math.cos(angle * DEG_TO_RAD) for angle in lut[:, 0]])
        f11, f22, f33, f44, f12, f34 = (lut[:, c].copy() for c in range(1, 7))
        # :790-818: trapezoid CDF of the rejection bound, normalized by its total.
        bound = f11 + np.abs(f12) if fixture.nmat > 1 else f11.copy()
        cdf = np.zeros(count)
        total = 0.0

        for i in range(1, count):
            total = total + (cosines[i] - cosines[i - 1]) * (bound[i] + bound[i - 1]) / 2.0
            cdf[i] = total

        cdf = cdf / total
        name = fixture.coelut[index].rstrip(".").replace("/", "-")
        species.append(Species(name, float(ssalb), cosines, f11, f22, f33, f12, f34, f44, cdf=cdf))

        if fixture.cloud_input_type == "ncol":
            depths[index] = fixture.cloud_abundance[index] * (q_ext * geometric)
        elif fixture.cloud_input_type == "baer":
            depths[index] = fixture.cloud_abundance[index]
        else:
            raise ValueError("cloud_input_type should be ncol or baer")

    return species, depths


def _depths_from_file(path: Path, wavelength_nm: float, nz: int) -> np.ndarray:
    # worker.f90:903-925: rows "wavelength(um) tau(1..nz-1)" after one header line; the row
    # whose wavelength equals wav/1000 within 1e-12 is taken, the vacuum layer is zero.
    r = _Records(path)
    r.skip(1)

    while True:
        tokens = r.values(nz)
        row = [_real(t, path) for t in tokens]

        if abs(row[0] - wavelength_nm / 1000.0) <= 1.0e-12:
            return np.array(row[1:] + [0.0])


@dataclass
class PreparedFixture:
    """A prepared fixture: transport tables plus the optical inputs the product reports."""

    scene: TransportScene
    fixture: FixtureInput
    bmsca: np.ndarray
    bmabs: np.ndarray
    baer: np.ndarray

    def product_variables(self) -> Dict[str, Tuple[np.ndarray, Tuple[str, ...], dict]]:
        """Optical inputs as ``Report_ncdf`` writes them (``worker.f90:1399-1440``,
        ``:1610-1622``, ``:1742-1767``), vacuum level and layer dropped, C axis order."""

        from .metadata import product_attributes

        f = self.fixture
        physical, levels = f.nz - 1, f.nz
        values = {
            "vmrgas": (f.vmrgas[:, :levels].T.copy(), ("nlev", "ngases")),
            "pres": (f.pres[:levels].copy(), ("nlev",)),
            "temp": (f.temp[:levels].copy(), ("nlev",)),
            "bmabs": (self.bmabs[:physical, None].copy(), ("nz", "nwavs")),
            "bmsca": (self.bmsca[:physical, None].copy(), ("nz", "nwavs")),
            "baer": (
                np.transpose(self.baer[:, :, :, :physical], (3, 2, 1, 0))[..., None].copy(),
                ("nz", "ny", "nx", "nscatt_types", "nwavs"),
            ),
        }
        return {name: (value, dims, product_attributes(name)) for name, (value, dims) in values.items()}


def prepare_fixture(
    path,
    data_dir,
    *,
    surface_albedo: float,
    solar_zenith: float,
    solar_azimuth: float,
    view_azimuth: float,
    view_zeniths,
    save_path: bool,
    description: str = "",
) -> PreparedFixture:
    """Transport tables of one positional fixture (``kernels/cpu/worker.f90:413-1010``).

    ``data_dir`` holds ``Mie_data/`` and ``abs_data/`` (the reference's ``DATA`` directory).
    Options the transport does not model raise ``NotImplementedError``; malformed input
    raises ``ValueError``.
    """

    fixture = read_fixture(path)

    if fixture.wavelengths != 1:
        raise NotImplementedError("one wavelength per prepared scene (worker.f90:1088)")

    if (fixture.source_location, fixture.detector_location) != (1, 1) or fixture.register_location not in (0, 1):
        # A ground detector needs view zeniths of 90 degrees or more, which cases do not take.
        raise NotImplementedError("only a TOA source and a TOA detector are supported (result grid TOA or surface)")

    if fixture.background != "Earth":
        raise NotImplementedError("only the Earth molecular background is restated")

    wavelength = fixture.wavelength_nm
    bmsca, bmabs, depol = molecular_optical_depths(fixture, wavelength, data_dir)

    if fixture.fix_depol:
        depol = fixture.depol_fixed

    inputs = Path(path).parents[1]

    if fixture.read_bmsca_from_file:
        bmsca = _depths_from_file(inputs / fixture.bmsca_file, wavelength, fixture.nz)

    if fixture.read_bmabs_from_file:
        bmabs = _depths_from_file(inputs / fixture.bmabs_file, wavelength, fixture.nz)

    species, baer = scatterer_tables(fixture, wavelength, data_dir)
    nx, ny, nz, kinds = fixture.nx, fixture.ny, fixture.nz, fixture.nscatt_types
    zunit = np.array([(fixture.alt[k + 1] - fixture.alt[k]) / 1000.0 for k in range(nz)])
    # worker.f90:946-1010: extinction per km by species (gas last), total, fractions.
    extinction = np.zeros((kinds + 1, nx, ny, nz))
    gas = (bmsca + bmabs) / zunit
    extinction[kinds] = gas[None, None, :]
    scattering = bmsca + bmabs
    ssalb_gas = np.where(scattering > EPS, bmsca / np.where(scattering > EPS, scattering, 1.0), 1.0)

    for isc in range(kinds):
        extinction[isc] = baer[isc] / zunit[None, None, :]

    beta = np.zeros((nx, ny, nz))

    for isc in range(kinds + 1):
        beta = beta + extinction[isc]

    fractions = np.zeros_like(extinction)
    positive = beta > EPS

    for isc in range(kinds + 1):
        fractions[isc][positive] = extinction[isc][positive] / beta[positive]

    scene = TransportScene(
        nx=nx,
        ny=ny,
        nz=nz,
        nmat=fixture.nmat,
        xunit=fixture.x_dim / float(nx),
        yunit=fixture.y_dim / float(ny),
        zunit=zunit,
        beta=beta,
        fractions=fractions,
        ssalb_gas=ssalb_gas,
        albedo=np.full((nx, ny), float(surface_albedo)),
        depol=float(depol),
        solar_zenith=float(solar_zenith),
        solar_azimuth=float(solar_azimuth),
        view_azimuth=float(view_azimuth),
        view_zeniths=tuple(view_zeniths),
        species=species,
        save_path=save_path,
        register=None if fixture.register_location == 1 else "surface",
        wavelength_nm=wavelength,
        altitude_m=fixture.alt[:nz].copy(),
        description=description or Path(path).name,
    )
    return PreparedFixture(scene, fixture, bmsca, bmabs, baer)


# Data files of the three shipped 340 nm fixtures, laid out as the reference's upstream tree
# (INPUT/testrun/*.in, DATA/Mie_data, DATA/abs_data); manifest.json lists their hashes.
UPSTREAM = Path(__file__).resolve().parent / "data" / "upstream"


def prepare_case(case, upstream: Path = UPSTREAM, *, save_path: bool = True, save_scats: bool = False):
    """Prepare a ``ReferenceCase`` as the reference's CPU preparer does for Metal.

    Checks of ``python/monki/execution/cpu.py:82-101``: the fixture's positional layout and the
    layer-AMF setting equal to ``save_path`` (each a ``ValueError``), then the preparation
    above with the case's geometry and albedo. The scattering-contribution setting must equal
    ``save_scats`` (the reference's CPU path refuses it; here it selects the decomposition,
    which replaces the layer paths as in mc.f90:826-865, so the AMF line is not checked then).
    Returns ``(SceneCase, PreparedFixture, provenance)``.
    """

    from .input import SceneCase
    from .metadata import file_hash

    upstream = Path(upstream)
    source = upstream / "INPUT" / case.input_file
    lines = check_layout(source)

    if len(case.view_zeniths) > 3:
        raise ValueError("case exceeds the transport limit of 3 viewing angles")

    if (lines[39].split()[0].lower() == ".true.") != save_scats:
        raise ValueError("save_scats must match the fixture's scattering-contribution setting")

    if not save_scats and (lines[38].split()[0].lower() == ".true.") != save_path:
        raise ValueError("save_path must match the fixture's layer AMF setting")

    prepared = prepare_fixture(
        source,
        upstream / "DATA",
        surface_albedo=case.surface_albedo,
        solar_zenith=case.solar_zenith,
        solar_azimuth=case.solar_azimuth,
        view_azimuth=case.view_azimuth,
        view_zeniths=case.view_zeniths,
        save_path=save_path and not save_scats,
        description=f"{case.input_file} ({case.name})",
    )
    prepared.scene.save_scats = save_scats
    scene_case = SceneCase(case.name, prepared.scene, case.photons_per_column, case.rounds, case.backward)
    provenance = {
        "preparer": "monki.preparation (rdinput, molecules_earth, get_O3_crabs, aerosols restated)",
        "case": {
            "name": case.name,
            "input_file": case.input_file,
            "surface_albedo": case.surface_albedo,
            "solar_zenith": case.solar_zenith,
            "solar_azimuth": case.solar_azimuth,
            "view_azimuth": case.view_azimuth,
            "view_zeniths": list(case.view_zeniths),
        },
        "source_input_sha256": file_hash(source),
        "wavelength_nm": prepared.scene.wavelength_nm,
        "depol": prepared.scene.depol,
    }
    return scene_case, prepared, provenance
