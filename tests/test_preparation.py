"""Optical preparation of the shipped fixtures without Fortran (CPU only).

Pins ``monki.preparation`` three ways:

- bit for bit against the reference's own preparation routines (rdsize, rdinput,
  molecules_earth, get_O3_crabs, aerosols; compiled unchanged with amdflang), whose outputs
  ``tests/golden/make_prep_vectors.py`` stored in ``prep_vectors.npz``;
- against the optical inputs that the original MONKI program wrote into the archived NetCDF-4
  products (``tests/golden/archived``), read with the package's HDF5 reader;
- against closed forms (hydrostatic column, linear interpolation) and the reader's
  list-directed record semantics.
"""

import hashlib
import json
import math
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
GOLDEN = ROOT / "tests" / "golden"
ARCHIVED = {
    "cloudfree": "MONKI_MLS_cloudfree_340_340_test_paper_as_005_th0_045_ph0_000_ph_000.nc",
    "1layercloud": "MONKI_MLS_1layercloud_340_340_test_paper_as_005_th0_045_ph0_000_ph_000.nc",
    "2boxclouds": "MONKI_MLS_2boxclouds_340_340_test_paper_as_005_th0_045_ph0_000_ph_000.nc",
}


def _vectors():
    with np.load(GOLDEN / "prep_vectors.npz") as data:
        arrays = {name: data[name] for name in data.files}

    settings = json.loads(arrays.pop("settings").tobytes().decode())
    return arrays, settings


def _prepare(settings, save_path=True):
    from monki.preparation import UPSTREAM, prepare_fixture

    return prepare_fixture(
        UPSTREAM / "INPUT" / settings["input_file"],
        UPSTREAM / "DATA",
        surface_albedo=settings["surface_albedo"],
        solar_zenith=settings["solar_zenith"],
        solar_azimuth=settings["solar_azimuth"],
        view_azimuth=settings["view_azimuth"],
        view_zeniths=settings["view_zeniths"],
        save_path=save_path,
    )


@pytest.mark.parametrize("name", ["cloudfree", "1layercloud", "2boxclouds", "1layercloud_geometry"])
def test_preparation_equals_reference_fortran_bit_for_bit(name):
    arrays, settings = _vectors()
    expected = arrays[name]
    save_path = bool(expected[10])
    prepared = _prepare(settings[name], save_path)
    tables = prepared.scene.transport_bin()
    np.testing.assert_array_equal(tables, expected[: tables.size])

    f = prepared.fixture
    nz, levels, kinds = f.nz, f.nz + 1, f.nscatt_types
    extra = expected[tables.size :]
    parts = {}

    for label, count in (("bmsca", nz), ("bmabs", nz), ("depol", 1), ("baer", kinds * f.nx * f.ny * nz),
                         ("alt", levels), ("pres", levels), ("temp", levels), ("vmrgas", f.ngases * levels)):  # fmt: skip
        parts[label], extra = extra[:count], extra[count:]

    assert extra.size == 0
    np.testing.assert_array_equal(prepared.bmsca, parts["bmsca"])
    np.testing.assert_array_equal(prepared.bmabs, parts["bmabs"])
    assert prepared.scene.depol == parts["depol"][0]
    np.testing.assert_array_equal(prepared.baer.ravel(order="F"), parts["baer"])

    for label in ("alt", "pres", "temp"):
        np.testing.assert_array_equal(getattr(f, label), parts[label])

    np.testing.assert_array_equal(f.vmrgas.ravel(order="F"), parts["vmrgas"])


@pytest.mark.parametrize("name", list(ARCHIVED))
def test_preparation_reproduces_archived_product_inputs(name):
    """The original program's products carry the optical inputs it prepared; ours equal them
    within the coordinate tolerance of the reference's product comparison (validation.py:12, 1e-12)."""

    from monki import SimulationResult
    from monki.metadata import scene_metadata
    from agreement import invariant_failures

    archived = SimulationResult.load(GOLDEN / "archived" / ARCHIVED[name])
    assert invariant_failures(archived) == []
    theta = tuple(archived.values["theta"])
    settings = {"input_file": f"testrun/MONKI_MLS_{name}_340_340_test.in", "surface_albedo": 0.05,
                "solar_zenith": 45.0, "solar_azimuth": 0.0, "view_azimuth": 0.0, "view_zeniths": theta[:3]}  # fmt: skip
    prepared = _prepare(settings, save_path="AMF_layer" in archived.values)
    ours = scene_metadata(prepared.scene, 1, extra=prepared.product_variables())

    for variable in ("pres", "temp", "alt", "bmabs", "bmsca", "baer", "vmrgas", "z_unit", "wav",
                     "x_unit", "y_unit", "theta0", "phi0", "phi", "asurf"):  # fmt: skip
        assert ours.dimensions[variable] == archived.dimensions[variable], variable
        assert ours.attributes[variable] == archived.attributes[variable], variable
        np.testing.assert_allclose(ours.values[variable], archived.values[variable], rtol=0, atol=1e-12, err_msg=variable)

    for variable in ("radiance", "std_radiance", "DOP", "radiance_convergence", "n_rounds"):
        assert ours.attributes[variable] == archived.attributes[variable], variable


def test_archived_products_read_as_netcdf4():
    from monki import SimulationResult

    box = SimulationResult.load(GOLDEN / "archived" / ARCHIVED["2boxclouds"])
    assert box.dimensions["radiance"] == ("nvza", "ny", "nx", "nmat", "nwavs")
    assert box.values["radiance"].shape == (5, 20, 30, 3, 1)
    np.testing.assert_array_equal(box.values["theta"], [0.0, 30.0, 45.0, 60.0, 75.0])
    assert int(box.values["n_photons_per_column"][0]) == 5000 and int(box.values["n_rounds"][0]) == 100
    assert "AMF_layer" not in box.values
    clear = SimulationResult.load(GOLDEN / "archived" / ARCHIVED["cloudfree"])
    assert clear.values["AMF_layer"].shape == (90, 33, 1, 1, 1)
    history = clear.values["radiance_convergence"][:, 0, 0]
    assert history.shape == (100,) and np.all(np.isfinite(history))
    assert clear.attributes["radiance"]["units"] == "-, assuming E0=pi"


def test_vendored_data_matches_manifest():
    from monki.preparation import UPSTREAM

    manifest = json.loads((UPSTREAM / "manifest.json").read_text())["files"]
    present = {str(p.relative_to(UPSTREAM)) for p in UPSTREAM.rglob("*") if p.is_file() and p.name != "manifest.json"}
    assert present == set(manifest)

    for name, digest in manifest.items():
        assert hashlib.sha256((UPSTREAM / name).read_bytes()).hexdigest() == digest, name

    reference = Path("/root/reference/fortran/monki/manifest.json")

    if reference.is_file():
        original = json.loads(reference.read_text())["files"]
        assert all(original[name] == digest for name, digest in manifest.items())


def test_list_directed_records(tmp_path):
    from monki.preparation import _Records

    path = tmp_path / "records.txt"
    path.write_text("  7  :seven\n\n 1.5, 2.5\n 3.5 / ignored 9\n.true. :flag\nname.dat      :text\nT\n")
    records = _Records(path)
    assert records.integer() == 7
    assert records.values(3) == ["1.5", "2.5", "3.5"]  # continues past a blank record
    assert records.logical() is True
    assert records.text(9) == "name.dat"
    assert records.logical() is True

    with pytest.raises(ValueError, match="end of file"):
        records.integer()


def _shifted_fixture(tmp_path, edit):
    from monki.preparation import UPSTREAM

    source = UPSTREAM / "INPUT/testrun/MONKI_MLS_2boxclouds_340_340_test.in"
    lines = source.read_text().splitlines()
    lines = edit(lines)
    target = tmp_path / "testrun" / "case.in"
    target.parent.mkdir(exist_ok=True)
    target.write_text("\n".join(lines) + "\n")
    return target


def test_layout_and_input_errors(tmp_path):
    from monki.preparation import check_layout, read_fixture

    with pytest.raises(ValueError, match="Unsupported positional input"):
        check_layout(_shifted_fixture(tmp_path, lambda lines: [""] + lines))

    def cloud_outside(lines):
        lines[69] = lines[69].replace("     4     11     8     12", "     4     31     8     12")
        return lines

    with pytest.raises(ValueError, match="x range"):
        read_fixture(_shifted_fixture(tmp_path, cloud_outside))

    with pytest.raises(ValueError, match="more lines than levels"):
        read_fixture(_shifted_fixture(tmp_path, lambda lines: lines + [lines[-1]]))

    fixture = read_fixture(_shifted_fixture(tmp_path, lambda lines: lines + ["", "# note"]))
    assert fixture.nz == 34 and fixture.cloud_abundance.shape == (1, 30, 20, 34)
    # Two overlapping boxes in layer 13 (level line 13): COT 10 over x 4..11 and 20 over 14..17.
    layer = fixture.cloud_abundance[0, :, :, 12]
    assert layer[3, 7] == 10.0 and layer[13, 8] == 20.0 and layer.sum() == 10.0 * 8 * 5 + 20.0 * 4 * 3


def test_hydrostatic_column_closed_form(tmp_path):
    """gas_col_type 1: N_k = (p_k - p_k+1) / (m g), so bmsca / dp is one constant (the
    scattering cross section over m g), and the column agrees with the ideal-gas column
    of the same profile to a few percent."""

    from dataclasses import replace

    from monki.preparation import UPSTREAM, molecular_optical_depths, read_fixture

    fixture = read_fixture(UPSTREAM / "INPUT/testrun/MONKI_MLS_cloudfree_340_340_test.in")
    ideal, _, depol = molecular_optical_depths(fixture, 340.0, UPSTREAM / "DATA")
    hydrostatic, _, depol_h = molecular_optical_depths(replace(fixture, gas_col_type=1), 340.0, UPSTREAM / "DATA")
    dp = -np.diff(fixture.pres)[:-1]
    ratio = hydrostatic[:-1] / dp
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-12)
    assert hydrostatic[-1] == 0.0 and ideal[-1] == 0.0 and depol == depol_h
    assert abs(hydrostatic.sum() / ideal.sum() - 1) < 0.03


def test_linear_interpolation_matches_closed_form():
    from monki.preparation import interpolate_linear

    rng = np.random.default_rng(3)
    x = np.cumsum(rng.uniform(0.1, 1.0, 50))
    y = rng.normal(size=50)

    for xp in rng.uniform(x[0], x[-1], 40):
        assert math.isclose(interpolate_linear(x, y, xp), float(np.interp(xp, x, y)), rel_tol=1e-12, abs_tol=1e-12)

    assert interpolate_linear(x, y, x[7]) == y[7] and interpolate_linear(x, y, x[-1]) == y[-1]
    below, above = x[0] - 0.5, x[-1] + 0.5
    assert math.isclose(interpolate_linear(x, y, below), y[0] - 0.5 * (y[1] - y[0]) / (x[1] - x[0]), rel_tol=1e-12)
    assert math.isclose(interpolate_linear(x, y, above), y[-1] + 0.5 * (y[-1] - y[-2]) / (x[-1] - x[-2]), rel_tol=1e-12)


def test_reference_case_prepares_a_scene_case():
    from monki import ReferenceCase, SceneCase
    from monki.preparation import prepare_case

    case = ReferenceCase("layered", "testrun/MONKI_MLS_1layercloud_340_340_test.in", 100, 8,
                         surface_albedo=0.1, solar_zenith=30.0, view_zeniths=(0.0, 45.0))  # fmt: skip
    scene_case, prepared, provenance = prepare_case(case)
    assert isinstance(scene_case, SceneCase) and scene_case.rounds == 8 and scene_case.photons_per_column == 100
    scene = scene_case.scene
    assert scene.nx == scene.ny == 1 and scene.nz == 34 and scene.nmat == 3 and len(scene.species) == 1
    assert np.all(scene.albedo == 0.1) and scene.view_zeniths == (0.0, 45.0) and scene.solar_zenith == 30.0
    # The cloud sits in layer 9 (7-8 km), COT 5: beta = 5 / 1 km plus the gas.
    assert scene.varying_layers().size == 0
    assert math.isclose(prepared.baer[0, 0, 0, 8], 5.0) and scene.beta[0, 0, 8] > 5.0
    assert provenance["case"]["input_file"] == case.input_file and len(provenance["source_input_sha256"]) == 64


def test_fixture_with_a_ground_result_grid(tmp_path):
    """A fixture whose result-grid line (main.f90:747-757) is 0 prepares a scene that registers
    at the ground; a ground detector is refused (cases keep view zeniths below 90 degrees)."""

    import shutil

    from monki import ReferenceCase
    from monki.preparation import UPSTREAM, prepare_case

    upstream = tmp_path / "upstream"
    shutil.copytree(UPSTREAM / "INPUT", upstream / "INPUT")
    (upstream / "DATA").symlink_to((UPSTREAM / "DATA").resolve())
    name = "testrun/MONKI_MLS_cloudfree_340_340_test.in"
    source = upstream / "INPUT" / name
    lines = source.read_text().splitlines(keepends=True)
    assert "definition of output nx*ny grid" in lines[11] and lines[11].startswith("1")
    lines[11] = "0" + lines[11][1:]
    source.write_text("".join(lines))
    scene_case, _, _ = prepare_case(ReferenceCase("grid", name), upstream)
    assert scene_case.scene.register == "surface" and scene_case.scene.detector == "toa"

    lines[10] = "0" + lines[10][1:]  # detector at the ground
    source.write_text("".join(lines))
    with pytest.raises(NotImplementedError, match="TOA detector"):
        prepare_case(ReferenceCase("grid", name), upstream)
