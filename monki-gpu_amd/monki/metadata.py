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


def scene_metadata(scene, photons_per_column: int, path: Path = Path("scene.nc")):
    """Coordinates and placeholders of the product, as ``Report_ncdf`` defines them.

    Variable and dimension names follow ``kernels/cpu/worker.f90:1270-1510`` reversed into
    C axis order (NetCDF's Python view of the Fortran declarations).
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
    attributes = {
        "theta0": {"units": "degrees", "long_name": "solar zenith angle"},
        "theta": {"units": "degrees", "long_name": "viewing zenith angle"},
        "phi0": {"units": "degrees", "long_name": "solar azimuth angle"},
        "phi": {"units": "degrees", "long_name": "viewing azimuth angle"},
        "x_unit": {"units": "km"},
        "y_unit": {"units": "km"},
        "z_unit": {"units": "km"},
        "alt": {"units": "m"},
        "wav": {"units": "nm"},
        "radiance": {"long_name": "TOA radiance", "units": "-"},
        "std_radiance": {"long_name": "TOA radiance standard error", "units": "-"},
        "DOP": {"long_name": "degree of polarization", "units": "-"},
    }

    if scene.save_path:
        values["AMF_layer"] = np.zeros((angles, physical, ny, nx, 1))
        dimensions["AMF_layer"] = ("nvza", "nz", "ny", "nx", "nwavs")
        attributes["AMF_layer"] = {"long_name": "Layer AMF", "units": "-"}

    return SimulationResult(Path(path), values, dimensions, attributes)
