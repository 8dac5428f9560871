"""Product metadata shared by the driver and the backends: coordinates, dimension names,
units and placeholders of the NetCDF product, and file hashes for receipts."""

import hashlib
from pathlib import Path

import numpy as np

from .output import SimulationResult


def file_hash(path) -> str:
    digest = hashlib.sha256()

    with open(path, "rb") as stream:
        for block in iter(lambda: stream.read(1 << 20), b""):
            digest.update(block)

    return digest.hexdigest()


# Report_ncdf's add_attributes calls (kernels/cpu/worker.f90:1512-1596): long_name, units and
# plot_title of every product variable, for a TOA detector (signal "TOA radiance").
ATTRIBUTES = {
    "bmabs": ("Molecular absorption optical thickness", "-", "bmabs"),
    "bmsca": ("Molecular scattering optical thickness", "-", "bmsca"),
    "baer": ("Aerosol total optical thickness (baerabs+baersca)", "-", "baer"),
    "wav": ("Wavelength of the light", "nm", "wav"),
    "asurf": ("Surface albedo", "-", "asurf"),
    "n_photons_per_column": ("Number of photons used per column and per round", "-", "n_photons_per_column"),
    "n_rounds": ("Number of rounds used", "-", "n_rounds"),
    "x_unit": ("Spatial resolution in the x-direction", "km", "x_unit"),
    "y_unit": ("Spatial resolution in the y-direction", "km", "y_unit"),
    "z_unit": ("Spatial resolution in the z-direction (layer dependent)", "km", "z_unit"),
    "vmrgas": ("Volume mixing ratio of the absorbing gases", "ppm", "vmrgas"),
    "pres": ("Pressure", "Pa", "pres"),
    "temp": ("Temperature", "K", "temp"),
    "alt": ("Altitude", "m", "alt"),
    "theta0": ("Solar zenith angle", "deg", "solar_zenith_angle"),
    "phi0": ("Solar azimuth angle", "deg", "solar_azimuth_angle"),
    "phi": ("Viewing azimuth angle", "deg", "viewing_azimuth_angle"),
    "theta": ("Viewing zenith angle", "deg", "viewing_zenith_angle"),
    "radiance": ("TOA radiance", "-, assuming E0=pi", "radiance"),
    "radiance_convergence": ("radiance_convergence", "-", "radiance convergence history"),
    "std_radiance": ("Standard deviation of TOA radiance", "-", "std"),
    "DOP": ("Degree of polarization", "-", "DOP"),
    "AMF_layer": ("Layer AMF", "-", "AMF_layer"),
    # Scattering decomposition (Report_ncdf with save_scats, main.f90:1724-1850).
    "radiance_sing": ("TOA radiance single scattering", "-", "radiance_sing"),
    "radiance_mult": ("TOA radiance multiple scattering", "-", "radiance_mult"),
    "contrib_radiance_3D": ("TOA radiance 3D contribution", "[-]", "contrib_radiance"),
    "contrib_radiance_3D_sing": ("TOA radiance 3D singscat contribution", "-", "contrib_radiance_sing"),
    "contrib_radiance_3D_mult": ("TOA radiance 3D multscat contribution", "-", "contrib_radiance_mult"),
}


def product_attributes(name: str) -> dict:
    long_name, units, plot_title = ATTRIBUTES[name]
    return {"long_name": long_name, "units": units, "plot_title": plot_title}


def scene_metadata(scene, photons_per_column: int, path: Path = Path("scene.nc"), extra=None):
    """Coordinates and placeholders of the product, as ``Report_ncdf`` defines them.

    Variable and dimension names follow ``kernels/cpu/worker.f90:1270-1510`` reversed into
    C axis order (NetCDF's Python view of the Fortran declarations). ``extra`` maps further
    variable names to ``(value, dimensions, attributes)``, such as the optical inputs of a
    prepared fixture.
    """

    nx, ny, nz, nmat, angles = scene.nx, scene.ny, scene.nz, scene.nmat, scene.angles
    physical = nz - 1
    values = {
        "asurf": np.array([float(np.mean(scene.albedo))]),
        "n_photons_per_column": np.array([photons_per_column], dtype=np.int32),
        "n_rounds": np.array([0], dtype=np.int32),
        "theta0": np.array([scene.solar_zenith]),
        "theta": np.asarray(scene.view_zeniths, dtype=np.float64),
        "phi0": np.array([scene.solar_azimuth]),
        "phi": np.array([scene.view_azimuth]),
        "x_unit": np.array([scene.xunit]),
        "y_unit": np.array([scene.yunit]),
        "z_unit": scene.zunit[:physical].copy(),
        "alt": scene.altitude_m.copy(),
        "wav": np.array([scene.wavelength_nm]),
        "radiance": np.zeros((angles, ny, nx, nmat, 1)),
        "std_radiance": np.full((angles, ny, nx, nmat, 1), 999.0),
        "DOP": np.zeros((angles, ny, nx, 1)),
        "radiance_convergence": np.zeros((0, nmat, 1)),
    }
    dimensions = {
        "asurf": ("single",),
        "n_photons_per_column": ("single",),
        "n_rounds": ("single",),
        "theta0": ("single",),
        "theta": ("nvza",),
        "phi0": ("single",),
        "phi": ("single",),
        "x_unit": ("single",),
        "y_unit": ("single",),
        "z_unit": ("nz",),
        "alt": ("nlev",),
        "wav": ("nwavs",),
        "radiance": ("nvza", "ny", "nx", "nmat", "nwavs"),
        "std_radiance": ("nvza", "ny", "nx", "nmat", "nwavs"),
        "DOP": ("nvza", "ny", "nx", "nwavs"),
        "radiance_convergence": ("nrounds", "nmat", "nwavs"),
    }
    attributes = {name: product_attributes(name) for name in values}

    if scene.save_path:
        values["AMF_layer"] = np.zeros((angles, physical, ny, nx, 1))
        dimensions["AMF_layer"] = ("nvza", "nz", "ny", "nx", "nwavs")
        attributes["AMF_layer"] = product_attributes("AMF_layer")

    if getattr(scene, "save_scats", False):
        for name in ("radiance_sing", "radiance_mult"):
            values[name] = np.zeros((angles, ny, nx, nmat, 1))
            dimensions[name] = ("nvza", "ny", "nx", "nmat", "nwavs")
            attributes[name] = product_attributes(name)

        # Levels: 0 the surface, then the physical layers (nz_withsurf, main.f90:1519, 1542).
        for name in ("contrib_radiance_3D", "contrib_radiance_3D_sing", "contrib_radiance_3D_mult"):
            values[name] = np.zeros((angles, nz, ny, nx, nmat, 1))
            dimensions[name] = ("nvza", "nz_withsurf", "ny", "nx", "nmat", "nwavs")
            attributes[name] = product_attributes(name)

    if getattr(scene, "detector", "toa") == "surface":
        # Report_ncdf's loc_name (main.f90:723-729, 1704): "Surface " for a ground detector.
        for attrs in attributes.values():
            attrs["long_name"] = attrs["long_name"].replace("TOA radiance", "Surface radiance")

    for name, (value, dims, attrs) in (extra or {}).items():
        values[name] = value
        dimensions[name] = dims
        attributes[name] = attrs

    return SimulationResult(Path(path), values, dimensions, attributes)
