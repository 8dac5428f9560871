"""Further GPU parity cases of the reference's test_metal.py / test_cpu.py.

- off-plane geometry (solar and view azimuths apart, so U carries signal) and scalar (I only)
  transport against the reference Fortran transport (test_metal.py:143-153, :204-226);
- the uniform-top detector shortcut against explicit cell-by-cell detector traversal of the
  same histories (test_metal.py:357-393);
- an adaptive AMF-convergence run equals a fixed run of its completed round count
  (test_cpu.py:141-161).
"""

import json
from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def compare_with_fortran(scene, ppc, rounds, seed, tmp_path, name, backward=False):
    """Radiance of HIP simulate against the reference Fortran rounds, within combined SE."""

    from monki import SceneCase, simulate
    from monki.execution import HIP
    from agreement import invariant_failures, z_statistics
    from oracle.oracle import run_reference

    gpu = simulate(SceneCase(name, scene, ppc, rounds, backward=backward), HIP(seed=seed), tmp_path / f"{name}.nc")
    assert invariant_failures(gpu) == []
    # One thread: the reference's compiler RNG sequence is then deterministic.
    tallies, _, _ = run_reference(scene, ppc, rounds, threads=1, backward=backward)
    nmat, nx, ny, views = scene.nmat, scene.nx, scene.ny, scene.angles
    stokes = tallies[:, : nmat * nx * ny * views].reshape(rounds, nmat, nx, ny, views, order="F") / ppc
    if backward and nmat >= 3:
        stokes[:, 2] *= -1  # the product reverses U of backward transport (simulation.py:128-131)
    mean = stokes.mean(axis=0)
    se = stokes.std(axis=0, ddof=1) / np.sqrt(rounds)
    a = gpu.radiance[..., 0].transpose(3, 2, 1, 0)
    b = gpu.standard_error[..., 0].transpose(3, 2, 1, 0)
    stats = z_statistics(a, b, mean, se, floor=1e-12)
    print(name, json.dumps(stats))
    assert stats["max_abs_z"] < 5.0
    assert 0.6 < stats["chi_square_per_dof"] < 1.5
    assert stats["fraction_above_3"] < 0.01
    return gpu, mean, se


def test_off_plane_polarization_matches_reference_fortran(tmp_path):
    from monki import configs

    scene = configs.box_cloud(nx=8, ny=6, nmat=3, solar_zenith=50.0, box=(3, 2), offset=(2, 1))
    scene = replace(scene, solar_azimuth=30.0, view_azimuth=120.0)
    _, mean, se = compare_with_fortran(scene, 100, 40, 977, tmp_path, "off-plane")
    # Off the principal plane U is a signal, not noise: the comparison above resolves it.
    assert np.median(np.abs(mean[2]) / np.maximum(se[2], 1e-30)) > 3


def test_full_stokes_transport_matches_reference_fortran(tmp_path):
    """nmat 4 (I, Q, U, V) against the reference Fortran at nmat 4, off the principal plane."""

    from monki import configs

    scene = configs.box_cloud(nx=8, ny=6, nmat=4, solar_zenith=50.0, box=(3, 2), offset=(2, 1))
    scene = replace(scene, solar_azimuth=30.0, view_azimuth=120.0)
    gpu, mean, se = compare_with_fortran(scene, 100, 40, 5150, tmp_path, "full-stokes")
    assert gpu.radiance.shape[-2] == 4
    # V is a small signal; its domain mean agrees in sign and size with the Fortran's.
    v_gpu = gpu.radiance[..., 3, 0].mean()
    v_ref = mean[3].mean()
    assert abs(v_gpu - v_ref) < 5 * np.sqrt((se[3] ** 2).mean() / se[3].size) + 0.2 * abs(v_ref)


@pytest.mark.parametrize("nmat", [3, 4])
def test_backward_transport_matches_reference_fortran(tmp_path, nmat):
    """Backward transport (weight matrix, injection-column register, no mu0/mu) against the
    reference Fortran's backward_mode, off the principal plane."""

    from monki import configs

    scene = configs.box_cloud(nx=8, ny=6, nmat=nmat, solar_zenith=50.0, box=(3, 2), offset=(2, 1))
    scene = replace(scene, solar_azimuth=30.0, view_azimuth=120.0)
    gpu, mean, se = compare_with_fortran(scene, 100, 40, 6262, tmp_path, f"backward{nmat}", backward=True)
    receipt = json.loads((tmp_path / f"backward{nmat}.run.json").read_text())
    assert receipt["transport"] == "backward"
    assert np.median(np.abs(mean[2]) / np.maximum(se[2], 1e-30)) > 3  # U resolved


@pytest.mark.parametrize("backward", [False, True])
def test_decomposition_matches_reference_fortran(tmp_path, backward):
    """The single/multiple-scattering decomposition against the reference Fortran run with
    save_scats (calc_radiances_and_scats), forward and backward: per-round means within
    combined SE, and the NetCDF product of simulate carries those means."""

    from monki import SceneCase, configs, simulate
    from monki.execution import HIP, BatchControls
    from monki.execution.hip_tables import pack_tables
    from agreement import z_statistics
    from oracle.oracle import fortran_scats_row, run_reference

    scene = replace(configs.box_cloud(nx=8, ny=6, nmat=3, box=(3, 2), offset=(2, 1)), solar_azimuth=30.0,
                    view_azimuth=120.0, save_path=False, save_scats=True)  # fmt: skip
    rounds, ppc = 40, 100
    case = SceneCase("scats", scene, ppc, rounds, backward=backward)

    with HIP(save_path=False, save_scats=True, seed=77) as backend:
        backend.prepare(case)
        batches = backend.run_batches([BatchControls(index, ppc) for index in range(1, rounds + 1)])

    _, shape = pack_tables(scene, shards=1)
    flat = lambda b: np.concatenate([b.stokes.ravel(order="F")] + [b.scattering[k].ravel(order="F") for k in
                                    ("stokes_sing", "stokes_mult", "contrib_sing", "contrib_mult")])  # fmt: skip
    ours = np.stack([flat(b) for b in batches]) / ppc
    tallies, _, _ = run_reference(scene, ppc, rounds, threads=1, backward=backward, save_scats=True)
    theirs = np.stack([fortran_scats_row(row, shape) for row in tallies]) / ppc
    stats = z_statistics(
        ours.mean(axis=0), ours.std(axis=0, ddof=1) / np.sqrt(rounds),
        theirs.mean(axis=0), theirs.std(axis=0, ddof=1) / np.sqrt(rounds), floor=1e-7,
    )  # fmt: skip
    print("decomposition", json.dumps(stats))
    assert stats["max_abs_z"] < 5.5
    assert 0.7 < stats["chi_square_per_dof"] < 1.4

    product = simulate(case, HIP(save_path=False, save_scats=True, seed=77), tmp_path / "scats.nc")
    sing = sum(b.scattering["stokes_sing"] for b in batches) / rounds / ppc
    if backward:
        sing[2] *= -1  # the product reverses U of backward runs (main.f90:1190-1198)
    sing = sing.transpose(3, 2, 1, 0)
    # Same streams; float32 bank sums differ only in atomic order.
    np.testing.assert_allclose(product.values["radiance_sing"][..., 0], sing, rtol=2e-5, atol=2e-5 * np.abs(sing).max())
    total = product.values["contrib_radiance_3D"][..., 0]
    np.testing.assert_allclose(total, product.values["contrib_radiance_3D_sing"][..., 0]
                               + product.values["contrib_radiance_3D_mult"][..., 0], rtol=1e-12)  # fmt: skip
    assert total.shape == (3, scene.nz, scene.ny, scene.nx, 3)


def test_surface_detector_matches_reference_fortran(tmp_path):
    """Detector and result grid at the ground (k_detector = k_register = 0): views pointing
    down, against the reference Fortran with the same settings."""

    from monki import configs

    scene = configs.box_cloud(nx=8, ny=6, nmat=3, solar_zenith=50.0, box=(3, 2), offset=(2, 1))
    scene = replace(scene, solar_azimuth=30.0, view_azimuth=120.0, detector="surface",
                    view_zeniths=(180.0, 150.0, 120.0))  # fmt: skip
    gpu, mean, se = compare_with_fortran(scene, 100, 40, 8484, tmp_path, "surface")
    assert gpu.attributes["radiance"]["long_name"] == "Surface radiance"
    assert np.all(mean[0] > 0)  # every ground pixel sees diffuse light


@pytest.mark.parametrize("grid", ["surface", "toa"])
def test_result_grid_on_the_other_level_matches_reference_fortran(tmp_path, grid):
    """Result grid away from the detector (k_register != k_detector): TOA views registered at
    the ground, and ground views registered at TOA, against the reference Fortran."""

    from monki import configs

    scene = configs.box_cloud(nx=8, ny=6, nmat=3, solar_zenith=50.0, box=(3, 2), offset=(2, 1))
    scene = replace(scene, solar_azimuth=30.0, view_azimuth=120.0, register=grid)
    if grid == "toa":
        scene = replace(scene, detector="surface", view_zeniths=(180.0, 150.0, 120.0))
    compare_with_fortran(scene, 100, 40, 9191, tmp_path, f"grid-{grid}")


def test_scalar_transport_matches_reference_fortran(tmp_path):
    from monki import configs

    scene = configs.box_cloud(nx=8, ny=6, nmat=1, box=(3, 2), offset=(2, 1))
    compare_with_fortran(scene, 100, 40, 4242, tmp_path, "scalar")


@pytest.mark.parametrize("name", ["box-cloud", "layered-cloud"])
def test_uniform_tail_matches_explicit_detector_traversal(monkeypatch, name):
    """Forcing the explicit detector walk through every layer (uniform_top = last physical
    layer) leaves the tallies of the same histories unchanged up to float32 rounding."""

    import monki.execution.hip as hip_module
    from monki import SceneCase, configs
    from monki.execution import HIP, BatchControls

    scene = (
        configs.box_cloud(nx=8, ny=6, nmat=3, box=(3, 2), offset=(2, 1))
        if name == "box-cloud"
        else configs.layered_cloud(nx=5, ny=4, nmat=3)
    )
    case = SceneCase(name, scene, 300, 2)

    def run():
        with HIP(tally_shards=8) as backend:
            backend.prepare(case)
            return backend.run_batch(BatchControls(1, case.photons_per_column))

    shortcut = run()
    original = hip_module.pack_tables

    def explicit(tables, shards, **options):
        packed, shape = original(tables, shards=shards, **options)
        assert packed[8] < tables.nz - 1
        packed = packed.copy()
        packed[8] = tables.nz - 1
        return packed, shape

    monkeypatch.setattr(hip_module, "pack_tables", explicit)
    walked = run()
    # Per view, the domain sums agree to rounding; per pixel, all but boundary cases do.
    for a, b in ((shortcut.stokes, walked.stokes), (shortcut.weighted_paths, walked.weighted_paths)):
        axes = tuple(range(a.ndim - 1))
        np.testing.assert_allclose(a.sum(axis=axes), b.sum(axis=axes), rtol=1e-4)
        close = np.abs(a - b) <= 1e-4 * np.abs(a).max() + 1e-4 * np.abs(a)
        assert close.mean() > 0.999, close.mean()


def test_adaptive_run_matches_fixed_run_at_completed_count(tmp_path):
    from monki import AMFConvergence, SceneCase, configs, simulate
    from monki.execution import HIP

    scene = configs.clear_sky(nx=3, ny=2, nmat=3)
    policy = AMFConvergence(0.02, 32, 16, 1)
    adaptive = simulate(SceneCase("adaptive", scene, 400, 512), HIP(), tmp_path / "adaptive.nc", convergence=policy)
    receipt = json.loads((tmp_path / "adaptive.run.json").read_text())
    assert receipt["status"] == "converged"
    completed = receipt["completed_rounds"]
    assert 32 <= completed < 512 and completed % 16 == 0
    fixed = simulate(SceneCase("fixed", scene, 400, completed), HIP(), tmp_path / "fixed.nc")

    # Same Philox streams and rounds; float32 bank sums differ only in atomic order.
    assert int(adaptive.values["n_rounds"][0]) == int(fixed.values["n_rounds"][0]) == completed
    for variable in ("radiance", "AMF_layer", "DOP"):
        expected = fixed.values[variable]
        np.testing.assert_allclose(adaptive.values[variable], expected, rtol=2e-5, atol=2e-5 * np.abs(expected).max(),
                                   err_msg=variable)  # fmt: skip
    np.testing.assert_allclose(adaptive.values["std_radiance"], fixed.values["std_radiance"], rtol=1e-3,
                               atol=1e-3 * np.abs(fixed.values["std_radiance"]).max())  # fmt: skip
