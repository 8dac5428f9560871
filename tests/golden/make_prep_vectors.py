"""Write prep_vectors.npz: prepared tables from the reference's Fortran preparation.

Builds tests/golden/prep_probe.f90 against the reference's unchanged rdsize, rdinput,
molecules_earth, get_O3_crabs and aerosols (make -C oracle prep; needs /root/reference) and
runs it on the three shipped 340 nm fixtures, reading them and their optical data where they
lie in the reference. Each case stores the probe's float64 stream: the transport.bin v1
tables (kernels/cpu/worker.f90:1102-1120) followed by bmsca, bmabs, depol, baer, alt, pres,
temp and vmrgas. One case varies the geometry arguments.
Run: python tests/golden/make_prep_vectors.py
"""

import json
import subprocess
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
UPSTREAM = Path("/root/reference/fortran/monki/upstream")
CASES = {
    "cloudfree": ("MONKI_MLS_cloudfree_340_340_test.in", 0.05, 45.0, 0.0, 0.0, (0.0, 30.0, 60.0)),
    "1layercloud": ("MONKI_MLS_1layercloud_340_340_test.in", 0.05, 45.0, 0.0, 0.0, (0.0, 30.0, 60.0)),
    "2boxclouds": ("MONKI_MLS_2boxclouds_340_340_test.in", 0.05, 45.0, 0.0, 0.0, (0.0, 30.0, 60.0)),
    "1layercloud_geometry": ("MONKI_MLS_1layercloud_340_340_test.in", 0.3, 60.0, 30.0, 120.0, (10.0, 45.0)),
}


def main():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "prep"], check=True)
    arrays, settings = {}, {}

    with tempfile.TemporaryDirectory() as scratch:
        for name, (fixture, albedo, sza, phi0, phi, views) in CASES.items():
            out = Path(scratch) / f"{name}.bin"
            command = [
                str(ROOT / "oracle/_ref/prep_probe"), str(UPSTREAM), str(UPSTREAM / "INPUT/testrun" / fixture),
                str(out), repr(albedo), repr(sza), repr(phi0), repr(phi), str(len(views)), *map(repr, views),
            ]  # fmt: skip
            subprocess.run(command, check=True, capture_output=True)
            arrays[name] = np.fromfile(out, dtype=np.float64)
            settings[name] = {
                "input_file": f"testrun/{fixture}",
                "surface_albedo": albedo,
                "solar_zenith": sza,
                "solar_azimuth": phi0,
                "view_azimuth": phi,
                "view_zeniths": list(views),
            }

    arrays["settings"] = np.frombuffer(json.dumps(settings).encode(), dtype=np.uint8)
    np.savez_compressed(Path(__file__).with_name("prep_vectors.npz"), **arrays)


if __name__ == "__main__":
    main()
