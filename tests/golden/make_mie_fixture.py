"""Write mie_reff2_340.json from the reference's 340 nm Mie data files.

Source: fortran/monki/upstream/DATA/Mie_data/Mie_nrvar_ni0/S.lu.00.340 (scattering matrix at
3601 angles; columns angle, F11, F22, F33, F44, F12, F34, -F12/F11) and S.sc.00.340 (header:
wavelength 0.34 um, m = 1.3442, gamma distribution r_eff 2 um, v_eff 0.1, integration bounds,
single-scattering albedo, asymmetry parameter). Every 20th angle is kept. Used to pin the
project's Mie routine (monki/optics.py). Run: python tests/golden/make_mie_fixture.py
"""

import json
from pathlib import Path

import numpy as np

DATA = Path("/root/reference/fortran/monki/upstream/DATA/Mie_data/Mie_nrvar_ni0")


def main():
    table = np.loadtxt(DATA / "S.lu.00.340")[::20]
    header = (DATA / "S.sc.00.340").read_text().splitlines()
    rmin = float(header[5].split()[2])
    rmax = float(header[5].split()[5])
    albedo = float(header[7].split()[-1])
    asymmetry = float(header[8].split()[-1])
    fixture = {
        "source": "reference DATA/Mie_data/Mie_nrvar_ni0/S.lu.00.340 (every 20th angle), S.sc.00.340",
        "wavelength_um": 0.34,
        "refractive_index": 1.3442,
        "reff_um": 2.0,
        "veff": 0.1,
        "rmin_um": rmin,
        "rmax_um": rmax,
        "ssalb": albedo,
        "asymmetry": asymmetry,
        "columns": ["angle_deg", "F11", "F22", "F33", "F44", "F12", "F34"],
        "rows": table[:, :7].tolist(),
    }
    path = Path(__file__).with_name("mie_reff2_340.json")
    path.write_text(json.dumps(fixture) + "\n")
    print(path, table.shape, rmin, rmax, albedo, asymmetry)


if __name__ == "__main__":
    main()
