"""Owned NetCDF result arrays, independent of the backend lifetime.

Same fields as the reference ``SimulationResult`` (``python/monki/output/__init__.py:10-74``):
values, dimension names and attributes per variable, exclusive-create saving and rejection of
missing values on load. netCDF4/HDF5 is not available here, so files are written as NetCDF-3
64-bit-offset through ``scipy.io.netcdf_file`` with the same variable names, dimension names
and attributes as the Fortran ``Report_ncdf`` product (``kernels/cpu/worker.f90:1270-1510``).
``load`` reads both that format and NetCDF-4 (HDF5) products such as the reference's archived
outputs (``monki.output.hdf5``).
"""

import os
from dataclasses import dataclass
from pathlib import Path
from typing import Dict, Tuple

import numpy as np
from scipy.io import netcdf_file


@dataclass(frozen=True)
class SimulationResult:
    """Own data, dimension names and units of one MONKI product."""

    path: Path
    values: Dict[str, np.ndarray]
    dimensions: Dict[str, Tuple[str, ...]]
    attributes: Dict[str, dict]

    @classmethod
    def load(cls, path) -> "SimulationResult":
        """Copy arrays before closing the file; reject masked/missing values."""

        path = Path(path).resolve()
        values, dimensions, attributes = {}, {}, {}

        with open(path, "rb") as stream:
            signature = stream.read(8)

        if signature == b"\x89HDF\r\n\x1a\n":
            from .hdf5 import read_netcdf4

            for name, (data, dims, attrs) in read_netcdf4(path).items():
                fill = attrs.get("_FillValue")

                if fill is not None and np.any(data == fill):
                    raise ValueError(f"{name}: masked/missing data in {path}")

                values[name], dimensions[name], attributes[name] = data, dims, attrs

            return cls(path, values, dimensions, attributes)

        with netcdf_file(path, "r", mmap=False) as dataset:
            for name, variable in dataset.variables.items():
                data = np.array(variable.data, copy=True)
                fill = getattr(variable, "_FillValue", None)

                if fill is not None and np.any(data == fill):
                    raise ValueError(f"{name}: masked/missing data in {path}")

                values[name] = data
                dimensions[name] = tuple(variable.dimensions)
                attributes[name] = {
                    key: (value.decode() if isinstance(value, bytes) else value)
                    for key, value in variable._attributes.items()
                }

        return cls(path, values, dimensions, attributes)

    def save(self) -> None:
        """Write owned arrays without replacing an existing scientific result."""

        self.path.parent.mkdir(parents=True, exist_ok=True)
        sizes = {}

        for name, value in self.values.items():
            for dimension, size in zip(self.dimensions[name], np.shape(value)):
                if dimension in sizes and sizes[dimension] != size:
                    raise ValueError(f"Inconsistent dimension {dimension}")

                sizes[dimension] = size

        # Exclusive creation protects existing outputs even against a racing writer.
        descriptor = os.open(self.path, os.O_CREAT | os.O_EXCL | os.O_WRONLY, 0o644)

        with os.fdopen(descriptor, "wb") as stream, netcdf_file(stream, "w", version=2) as out:
            for name, size in sizes.items():
                out.createDimension(name, size)

            for name, value in self.values.items():
                value = np.asarray(value)

                if value.dtype == np.int64:
                    value = value.astype(np.int32)

                variable = out.createVariable(name, value.dtype, self.dimensions[name])

                for key, attribute in self.attributes.get(name, {}).items():
                    setattr(variable, key, attribute)

                variable[...] = value

    @property
    def radiance(self):
        return self.values["radiance"]

    @property
    def standard_error(self):
        return self.values["std_radiance"]
