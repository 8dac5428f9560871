"""Write helper_vectors.json: expected values from the reference's Fortran helpers.

Builds tests/golden/helper_probe.f90 against the reference's unchanged
mc_routines_general_mod.f90 (make -C oracle probe; needs /root/reference) and records
inputs and outputs. Geometry cases include the reference's own probe inputs
(tests/native/metal_probe.f90:11-15: depol 0.03, S = (1, 0.2, -0.1)) plus vertical,
coplanar and random cases. The table cases use the project's 181-angle water-cloud table.
Run: python tests/golden/make_helper_vectors.py
"""

import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "monki-gpu_amd"))

from monki import configs  # noqa: E402


def main():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "probe"], check=True)
    rng = np.random.default_rng(7)
    geometry = [
        (-0.7, 0.4, 0.3, 1.8, 0.03, 1.0, 0.2, -0.1),
        (-1.0, 0.6, 0.2, 0.8, 0.03, 1.0, 0.2, -0.1),
        (-0.4, 1.0, 0.9, 0.4, 0.03, 1.0, 0.2, -0.1),
        (-0.766, 1.0, 0.0, 0.0, 0.0295, 1.0, 0.0, 0.0),
        (0.3, 0.8660254037844386, 2.0, 0.0, 0.0295, 1.0, -0.3, 0.4),
        (-0.5, 0.5, 1.0, 1.0 + np.pi, 0.0295, 1.0, 0.1, 0.1),
    ]

    for _ in range(40):
        uo, u1 = rng.uniform(-0.999, 0.999, 2)
        phio, phi1 = rng.uniform(-np.pi, 3 * np.pi, 2)
        q, u = rng.uniform(-0.6, 0.6, 2)
        geometry.append((uo, u1, phio, phi1, 0.0295, 1.0, q, u))

    species = configs.water_cloud()
    # Float32-exact values: the product uploads float32 tables, so both sides see the same nodes.
    table = np.stack([species.cosines, species.f11, species.f22, species.f33, species.f12, species.f34,
                      species.f44], axis=1)  # fmt: skip
    table = table.astype(np.float32).astype(np.float64)
    cosines = list(rng.uniform(-1, 1, 30)) + [-1.0, 1.0, 0.999, float(species.cosines[100])]
    # Full Stokes vectors (nmat 4): the reference probe's geometry and random ones, with V.
    full = [(-0.7, 0.4, 0.3, 1.8, 0.03, 1.0, 0.2, -0.1, 0.05), (-0.4, 1.0, 0.9, 0.4, 0.03, 1.0, 0.2, -0.1, -0.3),
            (0.3, 0.8660254037844386, 2.0, 0.0, 0.0295, 1.0, -0.3, 0.4, 0.2)]  # fmt: skip

    for _ in range(30):
        uo, u1 = rng.uniform(-0.999, 0.999, 2)
        phio, phi1 = rng.uniform(-np.pi, 3 * np.pi, 2)
        q, u, v = rng.uniform(-0.5, 0.5, 3)
        full.append((uo, u1, phio, phi1, 0.0295, 1.0, q, u, v))
    text = [str(len(geometry))] + [" ".join(f"{v:.17g}" for v in g) for g in geometry]
    text += [str(len(table))] + [" ".join(f"{v:.17g}" for v in row) for row in table]
    text += [str(len(cosines))] + [f"{c:.17g}" for c in cosines]
    text += [str(len(full))] + [" ".join(f"{v:.17g}" for v in g) for g in full]
    output = subprocess.run(
        [str(ROOT / "oracle/_ref/helper_probe")], input="\n".join(text) + "\n",
        capture_output=True, text=True, check=True,
    ).stdout.split()  # fmt: skip
    values = np.array([float(v) for v in output])
    geometry_out = values[: 12 * len(geometry)].reshape(len(geometry), 12)
    table_end = 12 * len(geometry) + 4 * len(cosines)
    table_out = values[12 * len(geometry) : table_end].reshape(len(cosines), 4)
    full_out = values[table_end:].reshape(len(full), 11)
    vectors = {
        "source": "reference mc_routines_general_mod.f90 via tests/golden/helper_probe.f90 (amdflang)",
        "geometry_inputs": [list(map(float, g)) for g in geometry],
        "geometry_outputs": geometry_out.tolist(),
        "geometry_columns": ["cos_sca", "C1", "S1", "C2", "S2", "F11", "F22", "F33", "F12", "I", "Q", "U"],
        "table": table.tolist(),
        "table_cosines": [float(c) for c in cosines],
        "table_outputs": table_out.tolist(),
        "stokes4_inputs": [list(map(float, g)) for g in full],
        "stokes4_outputs": full_out.tolist(),
        "stokes4_columns": ["F44 Rayleigh", "I", "Q", "U", "V", "F34 table", "F44 table", "I", "Q", "U", "V"],
    }
    path = Path(__file__).with_name("helper_vectors.json")
    path.write_text(json.dumps(vectors, indent=1) + "\n")
    print(path, len(geometry), len(cosines))


if __name__ == "__main__":
    main()
