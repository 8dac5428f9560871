"""The reference's three 340 nm fixtures through ``simulate(ReferenceCase, HIP())``.

The product is compared with the archived MONKI outputs (``tests/golden/archived``) by the
reference's own statistical rule (``validation.py:148-260``: combined standard errors with
a Bonferroni normal bound, family error rate 1e-6), at the case settings of the reference's
validation suite (``validation/cases/*.json``; ``run_suite.py:39-60``). Optical preparation
runs on the host from the package's fixture data; transport runs on the GPU.
"""

import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ARCHIVED = Path(__file__).resolve().parent / "golden" / "archived"
CASES = {
    "clear": ("testrun/MONKI_MLS_cloudfree_340_340_test.in", 2000, "cloudfree"),
    "layered-cloud": ("testrun/MONKI_MLS_1layercloud_340_340_test.in", 2000, "1layercloud"),
    "box-clouds": ("testrun/MONKI_MLS_2boxclouds_340_340_test.in", 200, "2boxclouds"),
}


def archived(stem: str):
    from monki import SimulationResult

    return SimulationResult.load(ARCHIVED / f"MONKI_MLS_{stem}_340_340_test_paper_as_005_th0_045_ph0_000_ph_000.nc")


@pytest.mark.parametrize("name", list(CASES))
def test_fixture_matches_archived_product(tmp_path, name):
    from monki import ReferenceCase, simulate
    from monki.execution import HIP
    from agreement import agree, invariant_failures, views

    input_file, photons, stem = CASES[name]
    case = ReferenceCase(name, input_file, photons, 32, view_zeniths=(0.0, 30.0, 60.0))
    result = simulate(case, HIP(save_path=name != "box-clouds"), tmp_path / f"{name}.nc")
    assert invariant_failures(result) == []
    golden = views(archived(stem), case.view_zeniths)
    comparison = agree(golden, result)
    print(json.dumps(comparison.__dict__, default=str))
    assert comparison.passed, comparison.problems
    receipt = json.loads((tmp_path / f"{name}.run.json").read_text())
    assert receipt["preparation"]["case"]["input_file"] == input_file
    assert receipt["status"] == "complete" and receipt["completed_rounds"] == 32


def test_batched_fixtures_equal_single_sessions():
    from monki import ReferenceCase
    from monki.execution import HIP, HIPBatch, BatchControls

    cases = [
        ReferenceCase("clear", CASES["clear"][0], 300, 2),
        ReferenceCase("layered", CASES["layered-cloud"][0], 400, 2, solar_zenith=60.0),
    ]

    with HIPBatch(tally_shards=8) as batch:
        batch.prepare(cases)
        batched = batch.run_round(1)
        assert batch.case_provenance()[1]["preparation"]["case"]["solar_zenith"] == 60.0

    for case, result in zip(cases, batched):
        with HIP(tally_shards=8) as single:
            single.prepare(case)
            alone = single.run_batch(BatchControls(1, case.photons_per_column))

        np.testing.assert_allclose(result.stokes, alone.stokes, rtol=2e-5, atol=2e-5)
        np.testing.assert_allclose(result.weighted_paths, alone.weighted_paths, rtol=2e-5, atol=2e-5)
