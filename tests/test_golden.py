"""Pin the float64 oracle and the table generators against golden vectors and closed forms.

Golden sources: the Random123 Philox KAT (the reference's tests/integration/test_metal.py:43-59),
outputs of the reference's own Fortran helpers (tests/golden/helper_vectors.json, written by
make_helper_vectors.py from mc_routines_general_mod.f90), and the reference's 340 nm Mie LUT
(tests/golden/mie_reff2_340.json from DATA/Mie_data).
"""

import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).with_name("golden")


def helper_vectors():
    return json.loads((GOLDEN / "helper_vectors.json").read_text())


def test_philox_published_known_answers(oracle):
    # Random123 kat_vectors, Philox4x32-10.
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]  # fmt: skip
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]  # fmt: skip


def test_rayleigh_matrix_closed_form(oracle):
    # Hansen & Travis with depolarization; at depol 0: F11 = 3/16 (1 + c^2), F33 = 3c/8.
    for c in (-1.0, -0.3, 0.0, 0.5, 1.0):
        f = oracle.rayleigh(0.0, c)
        np.testing.assert_allclose(f, [3 / 16 * (1 + c * c), 3 / 16 * (1 + c * c), 3 * c / 8, -3 / 16 * (1 - c * c)], atol=1e-15)
    # Normalization: mean of F11 over the sphere is 1/4 (the /4 radiance convention).
    mu = np.linspace(-1, 1, 20001)
    for depol in (0.0, 0.0295, 0.1):
        f11 = np.array([oracle.rayleigh(depol, c)[0] for c in mu[::100]])
        assert abs(np.trapezoid(f11, mu[::100]) / 2 - 0.25) < 1e-4
    # Rejection bound used by rayl_scatter_pol: F11(c) + |F12(c)| equals F11(1) for all c.
    for c in np.linspace(-1, 1, 41):
        f = oracle.rayleigh(0.0295, c)
        assert f[0] + abs(f[3]) == pytest.approx(oracle.rayleigh(0.0295, 1.0)[0], rel=1e-13)


def test_oracle_helpers_match_reference_fortran(oracle):
    vectors = helper_vectors()

    for inputs, expected in zip(vectors["geometry_inputs"], vectors["geometry_outputs"]):
        uo, u1, phio, phi1, depol, i, q, u = inputs
        cosine = expected[0]
        f = oracle.rayleigh(depol, cosine)
        np.testing.assert_allclose(f, expected[5:9], rtol=1e-12, atol=1e-15)
        r = oracle.rotation(cosine, uo, u1, phio, phi1)
        np.testing.assert_allclose(r, expected[1:5], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(oracle.scatter(f, r, [i, q, u]), expected[9:12], rtol=1e-9, atol=1e-12)


def test_rotation_is_orthonormal(oracle):
    rng = np.random.default_rng(3)

    for _ in range(200):
        uo, u1 = rng.uniform(-1, 1, 2)
        phio, phi1 = rng.uniform(-7, 7, 2)
        c = uo * u1 + np.sqrt(1 - uo**2) * np.sqrt(1 - u1**2) * np.cos(phio - phi1)
        c1, s1, c2, s2 = oracle.rotation(c, uo, u1, phio, phi1)
        assert c1 * c1 + s1 * s1 == pytest.approx(1, abs=1e-9)
        assert c2 * c2 + s2 * s2 == pytest.approx(1, abs=1e-9)


def test_tabulated_interpolation_matches_reference_fortran(oracle):
    from monki.scene import Species

    vectors = helper_vectors()
    table = np.array(vectors["table"])
    species = Species("probe", 1.0, *table.T)

    for cosine, expected in zip(vectors["table_cosines"], vectors["table_outputs"]):
        # The oracle interpolates the float32 table the product uploads.
        np.testing.assert_allclose(oracle.tabulated(species, cosine), expected, rtol=1e-12, atol=1e-14)


def test_full_stokes_helpers_match_reference_fortran(oracle):
    """nmat 4: Rayleigh F44, tabulated F34/F44 and the rotated 4x4 phase matrix times a full
    Stokes vector, against ray_scat_matrix / mie_scat_matrix / matmul_stokes_rot at nmat 4."""

    from monki.scene import Species

    vectors = helper_vectors()
    table = np.array(vectors["table"])
    species = Species("probe", 1.0, *table.T)
    assert np.any(species.f34 != 0)  # the water-cloud table has a circular-polarization term

    for inputs, expected in zip(vectors["stokes4_inputs"], vectors["stokes4_outputs"]):
        uo, u1, phio, phi1, depol, *stokes = inputs
        cosine = uo * u1 + np.sqrt(1 - uo**2) * np.sqrt(1 - u1**2) * np.cos(phio - phi1)
        r = oracle.rotation(cosine, uo, u1, phio, phi1)
        rayleigh = [*oracle.rayleigh(depol, cosine), 0.0, oracle.rayleigh_f44(depol, cosine)]
        assert rayleigh[5] == pytest.approx(expected[0], rel=1e-12, abs=1e-15)
        np.testing.assert_allclose(oracle.scatter4(rayleigh, r, stokes), expected[1:5], rtol=1e-9, atol=1e-12)
        v = oracle.tabulated_v(species, cosine)
        np.testing.assert_allclose(v, expected[5:7], rtol=1e-9, atol=1e-12)
        tabulated = [*oracle.tabulated(species, cosine), *v]
        np.testing.assert_allclose(oracle.scatter4(tabulated, r, stokes), expected[7:11], rtol=1e-9, atol=1e-12)


def test_full_stokes_matrix_reduces_to_iqu(oracle):
    # With V = 0 and F34 = 0 the first three rows equal the nmat 3 product; at depol 0 the
    # Rayleigh F44 equals F33.
    rng = np.random.default_rng(5)

    for _ in range(50):
        uo, u1 = rng.uniform(-0.99, 0.99, 2)
        phio, phi1 = rng.uniform(-3, 3, 2)
        c = uo * u1 + np.sqrt(1 - uo**2) * np.sqrt(1 - u1**2) * np.cos(phio - phi1)
        f = oracle.rayleigh(0.0295, c)
        r = oracle.rotation(c, uo, u1, phio, phi1)
        s = [1.0, *rng.uniform(-0.5, 0.5, 2)]
        full = oracle.scatter4([*f, 0.0, oracle.rayleigh_f44(0.0295, c)], r, [*s, 0.0])
        np.testing.assert_allclose(full[:3], oracle.scatter(f, r, s), rtol=1e-13, atol=1e-16)
        assert full[3] == 0.0
        assert oracle.rayleigh_f44(0.0, c) == pytest.approx(oracle.rayleigh(0.0, c)[2], rel=1e-14, abs=1e-16)


def test_scattering_direction_geometry(oracle):
    # The new direction must make the sampled scattering angle with the old one.
    rng = np.random.default_rng(11)

    for _ in range(500):
        cos_in = rng.uniform(-0.999, 0.999)
        cos_sca = rng.uniform(-1, 1)
        sigma = rng.uniform(-np.pi, np.pi)
        cos_out, dphi = oracle.new_direction(cos_in, cos_sca, sigma)
        s_in, s_out = np.sqrt(1 - cos_in**2), np.sqrt(max(0.0, 1 - cos_out**2))
        # old direction at azimuth 0, new at azimuth dphi
        dot = s_in * s_out * np.cos(dphi) + cos_in * cos_out
        assert dot == pytest.approx(cos_sca, abs=2e-7)


def test_mie_routine_matches_reference_lut():
    from monki.optics import mie_gamma

    fixture = json.loads((GOLDEN / "mie_reff2_340.json").read_text())
    rows = np.array(fixture["rows"])
    table = mie_gamma(
        fixture["wavelength_um"], fixture["refractive_index"] + 0j, fixture["reff_um"], fixture["veff"],
        angles=3601, radii=4000, rmin_um=fixture["rmin_um"], rmax_um=fixture["rmax_um"],
    )  # fmt: skip
    ours = {name: table[name][::-1][::20] for name in ("f11", "f22", "f33", "f44", "f12", "f34")}
    assert table["ssalb"] == pytest.approx(fixture["ssalb"], abs=1e-9)
    assert table["asymmetry"] == pytest.approx(fixture["asymmetry"], abs=5e-4)
    np.testing.assert_allclose(ours["f11"], rows[:, 1], rtol=0.015)
    np.testing.assert_allclose(ours["f33"], rows[:, 3], rtol=0.015, atol=2e-3 * rows[:, 1].max() / 816)
    # Degree of linear polarization -F12/F11: within 0.02 at every angle, 0.002 on average
    # (the size quadratures differ: 100 x 50 Gauss points in the reference's Mie program).
    for column, name in ((5, "f12"), (6, "f34")):
        difference = np.abs(ours[name] / ours["f11"] - rows[:, column] / rows[:, 1])
        assert difference.max() < 0.02 and difference.mean() < 0.002, (name, difference.max())


def test_bound_cdf_matches_trapezoid_definition():
    from monki.scene import Species

    mu = np.linspace(-1, 1, 181)
    flat = Species("flat", 1.0, mu, np.ones_like(mu), np.ones_like(mu), mu, np.zeros_like(mu))
    np.testing.assert_allclose(flat.bound_cdf(3), (mu + 1) / 2, atol=1e-15)
    # With |F12| the bound integrates F11 + |F12| (main_routines_general_mod.f90:834-866).
    f12 = -0.5 * (1 - mu * mu)
    pol = Species("pol", 1.0, mu, np.ones_like(mu), np.ones_like(mu), mu, f12)
    bound = 1 + np.abs(f12)
    steps = np.concatenate([[0], np.cumsum(np.diff(mu) * (bound[1:] + bound[:-1]) / 2)])
    np.testing.assert_allclose(pol.bound_cdf(3), steps / steps[-1], rtol=1e-14)
    np.testing.assert_allclose(pol.bound_cdf(1), (mu + 1) / 2, atol=1e-15)


# ---- Direction sampling: distributions fixed by the phase matrix ---------------------------


def _chi_square_p(counts, probabilities):
    from scipy.stats import chi2

    expected = probabilities * counts.sum()
    statistic = float(((counts - expected) ** 2 / expected).sum())
    return statistic, float(chi2.sf(statistic, len(counts) - 1))


def _packed(scene):
    from monki.execution.hip_tables import pack_tables

    return pack_tables(scene, shards=1)[0]


def test_rayleigh_sampling_follows_f11(oracle):
    """Unpolarized light: cos(theta) has density F11(c) / 2 (mc_routines_general_mod.f90:492-670)."""

    from monki import configs

    scene = configs.clear_sky(nx=2, ny=2, nmat=1)
    rows = oracle.sample(_packed(scene), -1, 0.0, 0.0, 0.3, 200_000)
    delta = (1 - scene.depol) / (1 + scene.depol / 2)
    edges = np.linspace(-1, 1, 41)
    primitive = delta * 0.75 * (edges + edges**3 / 3) + (1 - delta) * edges
    counts, _ = np.histogram(rows[:, 0], edges)
    statistic, p = _chi_square_p(counts, np.diff(primitive) / 2.0)
    assert p > 1e-3, statistic


@pytest.mark.parametrize("q, u", [(0.8, 0.0), (0.0, 0.8), (-0.5, -0.6)])
def test_rayleigh_sampling_of_sigma_follows_linear_polarization(oracle, q, u):
    """Polarized light: p(sigma) = (2 - Delta Q cos 2 sigma + Delta U sin 2 sigma) / 4 pi, the
    cos-integral of F11 + F12 (Q cos 2 sigma - U sin 2 sigma) with the forward sign -1
    (mc_routines_general_mod.f90:583-587). A flipped sign of either term fails."""

    from monki import configs

    scene = configs.clear_sky(nx=2, ny=2, nmat=3)
    rows = oracle.sample(_packed(scene), -1, q, u, -0.4, 200_000)
    delta = (1 - scene.depol) / (1 + scene.depol / 2)
    edges = np.linspace(-np.pi, np.pi, 25)
    primitive = 2 * edges - delta * q * np.sin(2 * edges) / 2 - delta * u * np.cos(2 * edges) / 2
    counts, _ = np.histogram(rows[:, 1], edges)
    statistic, p = _chi_square_p(counts, np.diff(primitive) / (4 * np.pi))
    assert p > 1e-3, statistic


@pytest.mark.parametrize("species", [0, 1])
def test_tabulated_sampling_follows_trapezoid_cdf(oracle, species):
    """nmat = 1: every draw is accepted, so the probability of a node interval is its trapezoid
    area of F11 (main_routines_general_mod.f90:822-866 with mie_scatter_pol_bound)."""

    from monki import configs

    scene = configs.layered_cloud(nx=2, ny=2, nmat=1)
    sp = scene.species[species]
    cosines = sp.cosines.astype(np.float32).astype(np.float64)
    f11 = sp.f11.astype(np.float32).astype(np.float64)
    rows = oracle.sample(_packed(scene), species, 0.0, 0.0, 0.7, 200_000)
    nodes = np.concatenate([np.arange(0, cosines.size - 1, 9), [cosines.size - 1]])
    areas = np.concatenate([[0.0], np.cumsum(np.diff(cosines) * (f11[1:] + f11[:-1]) / 2)])
    probabilities = np.diff(areas[nodes]) / areas[-1]
    counts, _ = np.histogram(rows[:, 0], cosines[nodes])
    statistic, p = _chi_square_p(counts, probabilities)
    assert p > 1e-3, statistic


def test_polarized_mie_sampling_follows_bounded_proposal(oracle):
    """nmat = 3: the cosine is drawn piecewise uniformly from the trapezoid CDF of F11 + |F12|
    and accepted with probability F11 / (F11 + |F12|) for unpolarized light, so a node interval
    has probability (its bound area / width) times the integral of F11 / (F11 + |F12|)."""

    from monki import configs

    scene = configs.layered_cloud(nx=2, ny=2, nmat=3)
    sp = scene.species[0]
    c = sp.cosines.astype(np.float32).astype(np.float64)
    f11 = sp.f11.astype(np.float32).astype(np.float64)
    f12 = sp.f12.astype(np.float32).astype(np.float64)
    rows = oracle.sample(_packed(scene), 0, 0.0, 0.0, 0.2, 200_000)
    bound = f11 + np.abs(f12)
    weights = np.empty(c.size - 1)

    for n in range(c.size - 1):
        t = np.linspace(0, 1, 65)
        a = f11[n] + t * (f11[n + 1] - f11[n])
        b = f12[n] + t * (f12[n + 1] - f12[n])
        area = (c[n + 1] - c[n]) * (bound[n] + bound[n + 1]) / 2
        weights[n] = area * np.trapezoid(a / (a + np.abs(b)), t)

    nodes = np.concatenate([np.arange(0, c.size - 1, 9), [c.size - 1]])
    cumulative = np.concatenate([[0.0], np.cumsum(weights)])
    probabilities = np.diff(cumulative[nodes]) / cumulative[-1]
    counts, _ = np.histogram(rows[:, 0], c[nodes])
    statistic, p = _chi_square_p(counts, probabilities)
    assert p > 1e-3, statistic


# ---- Whole histories: closed forms and the reference Fortran --------------------------------


def absorbing_column(optical_depth=0.6, albedo=0.3, solar_zenith=35.0, nmat=1):
    """One periodic column, three 1-km layers, pure absorption (single-scattering albedo 0)."""

    from monki.scene import VACUUM_THICKNESS_KM, TransportScene

    zunit = np.array([1.0, 1.0, 1.0, VACUUM_THICKNESS_KM])
    beta = np.zeros((1, 1, 4))
    beta[0, 0, :3] = optical_depth / 3.0
    return TransportScene(
        nx=1, ny=1, nz=4, nmat=nmat, xunit=1.0, yunit=1.0, zunit=zunit, beta=beta,
        fractions=np.ones((1, 1, 1, 4)), ssalb_gas=np.zeros(4), albedo=np.full((1, 1), albedo),
        depol=0.03, solar_zenith=solar_zenith, solar_azimuth=0.0, view_azimuth=0.0,
        view_zeniths=(0.0, 30.0, 60.0),
    )  # fmt: skip


def absorbing_column_expectation(scene, optical_depth):
    """Only the surface reflection contributes: I = A mu0 exp(-tau/mu0 - tau/mu) per photon
    (Lambertian estimate S1 mu w, times mu0/mu, mc.f90:738-822 and 1456-1474), and every
    contribution's layer path is dz/mu0 + dz/mu, so each layer AMF is 1/mu0 + 1/mu exactly."""

    mu0 = np.cos(np.deg2rad(scene.solar_zenith))
    mu = np.cos(np.deg2rad(scene.view_zeniths))
    radiance = scene.albedo[0, 0] * mu0 * np.exp(-optical_depth / mu0 - optical_depth / mu)
    return radiance, 1 / mu0 + 1 / mu


def test_oracle_absorbing_column_closed_form(oracle):
    tau = 0.6
    scene = absorbing_column(tau)
    photons = 100_000
    tallies, _, _, flag = oracle.run(_packed(scene), [photons, 0, 1, 7, 100000, 0, photons, 1])
    assert flag == 0
    radiance, amf = absorbing_column_expectation(scene, tau)
    stokes = tallies[0, :3] / photons
    paths = tallies[0, 3 : 3 + 12].reshape(4, 3, order="F")
    weights = tallies[0, 15:18]
    # The surface is reached with probability exp(-tau/mu0): binomial standard error.
    p = np.exp(-tau / np.cos(np.deg2rad(scene.solar_zenith)))
    np.testing.assert_allclose(stokes, radiance, rtol=5 * np.sqrt((1 - p) / (p * photons)))
    # Angles reach the transport as float32 radians: AMFs carry their rounding (~1e-7).
    np.testing.assert_allclose(paths[:3] / weights / 1.0, np.broadcast_to(amf, (3, 3)), rtol=1e-6)
    # The vacuum layer holds the launch path only (the detector walk stops below it).
    mu0 = np.cos(np.deg2rad(scene.solar_zenith))
    np.testing.assert_allclose(paths[3] / weights, scene.zunit[3] / mu0, rtol=1e-6)


@pytest.mark.skipif(not (Path(__file__).parents[1] / "oracle" / "_ref" / "monki_ref").is_file(),
                    reason="reference Fortran build missing (make -C oracle ref)")  # fmt: skip
@pytest.mark.parametrize("name", ["clear-IQU", "box-cloud", "box-cloud-IQUV", "box-cloud-backward",
                                  "box-cloud-IQUV-backward", "box-cloud-surface-detector", "box-cloud-ground-grid",
                                  "box-cloud-surface-detector-toa-grid"])  # fmt: skip
def test_oracle_matches_reference_fortran_statistically(oracle, name):
    """The float64 oracle against the reference's unchanged mc.f90 (its own compiler random
    numbers): radiance and layer paths per pixel/view within combined standard errors, for
    I/Q/U, the full Stokes vector (nmat 4) and backward transport (backward_mode)."""

    from dataclasses import replace

    from monki import configs
    from monki.execution.hip_tables import pack_tables
    from agreement import z_statistics
    from oracle.oracle import run_reference

    box = configs.box_cloud(nx=4, ny=3, nmat=3, box=(2, 1), offset=(1, 1))
    # Off the principal plane, so that U (and with nmat 4, V) carry signal.
    off_plane = replace(box, solar_azimuth=30.0, view_azimuth=120.0)
    scene = {
        "clear-IQU": configs.clear_sky(nx=3, ny=2, nmat=3),
        "box-cloud": box,
        "box-cloud-IQUV": replace(off_plane, nmat=4),
        "box-cloud-backward": off_plane,
        "box-cloud-IQUV-backward": replace(off_plane, nmat=4),
        # Detector and grid at the ground: views pointing down (zenith 180, 150, 120).
        "box-cloud-surface-detector": replace(off_plane, detector="surface", view_zeniths=(180.0, 150.0, 120.0)),
        # Result grid on the other level: TOA views registered at the ground, and back.
        "box-cloud-ground-grid": replace(off_plane, register="surface"),
        "box-cloud-surface-detector-toa-grid": replace(off_plane, detector="surface", register="toa",
                                                       view_zeniths=(180.0, 150.0, 120.0)),  # fmt: skip
    }[name]
    backward = name.endswith("backward")
    rounds, ppc = 40, 60
    per_round = ppc * scene.nx * scene.ny
    packed, _ = pack_tables(scene, shards=1, backward=backward)
    ours, _, _, flag = oracle.run(packed, [per_round * rounds, 0, 1, 5, 100000, 0, per_round, rounds])
    assert flag == 0
    # One thread: the reference RNG sequence (and so the test) is then deterministic.
    theirs, _, _ = run_reference(scene, ppc, rounds, threads=1, backward=backward)
    a, b = ours / ppc, theirs / ppc
    stats = z_statistics(
        a.mean(axis=0), a.std(axis=0, ddof=1) / np.sqrt(rounds),
        b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(rounds), floor=1e-12,
    )  # fmt: skip
    print(name, stats)
    assert stats["max_abs_z"] < 5.5
    assert 0.75 < stats["chi_square_per_dof"] < 1.35


@pytest.mark.skipif(not (Path(__file__).parents[1] / "oracle" / "_ref" / "monki_ref").is_file(),
                    reason="reference Fortran build missing (make -C oracle ref)")  # fmt: skip
@pytest.mark.parametrize("backward", [False, True])
def test_oracle_decomposition_matches_reference_fortran(oracle, backward):
    """Single/multiple-scattering decomposition (calc_radiances_and_scats): the oracle's
    Stokes_sing/mult and 3-D contributions against the reference's unchanged mc.f90 run with
    save_scats (forward, and backward with contributions at the first scattering cell), within
    combined SE; and the identities the decomposition fixes."""

    from dataclasses import replace

    from monki import configs
    from monki.execution.hip_tables import pack_tables
    from agreement import z_statistics
    from oracle.oracle import fortran_scats_row, run_reference

    scene = replace(configs.box_cloud(nx=4, ny=3, nmat=3, box=(2, 1), offset=(1, 1)), solar_azimuth=30.0,
                    view_azimuth=120.0, save_path=False, save_scats=True)  # fmt: skip
    rounds, ppc = 40, 60
    per_round = ppc * scene.nx * scene.ny
    packed, shape = pack_tables(scene, shards=1, backward=backward)
    ours, _, _, flag = oracle.run(packed, [per_round * rounds, 0, 1, 5, 100000, 0, per_round, rounds])
    assert flag == 0
    theirs, _, _ = run_reference(scene, ppc, rounds, threads=1, backward=backward, save_scats=True)
    theirs = np.stack([fortran_scats_row(row, shape) for row in theirs])
    assert ours.shape == theirs.shape
    nmat, nx, ny, nz, views = shape
    stokes = nmat * nx * ny * views

    for tallies in (ours, theirs):
        # Stokes = single + multiple, and the contributions by event cell sum to the Stokes
        # tallies of each view and component.
        np.testing.assert_allclose(tallies[:, :stokes], tallies[:, stokes : 2 * stokes] + tallies[:, 2 * stokes : 3 * stokes],
                                   rtol=1e-12, atol=1e-12)  # fmt: skip
        contrib = tallies[:, 3 * stokes :].reshape(rounds, nmat, nx * ny * nz, views, 2, order="F")
        by_view = tallies[:, :stokes].reshape(rounds, nmat, nx * ny, views, order="F").sum(axis=2)
        np.testing.assert_allclose(contrib.sum(axis=(2, 4)), by_view, rtol=1e-9, atol=1e-9)

    a, b = ours / ppc, theirs / ppc
    # Floor: single-scattered U at nadir is zero up to the rounding of the float32 geometry
    # (1e-9 here against I ~ 0.1), which its near-zero round-to-round spread cannot absorb.
    stats = z_statistics(
        a.mean(axis=0), a.std(axis=0, ddof=1) / np.sqrt(rounds),
        b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(rounds), floor=1e-7,
    )  # fmt: skip
    print(stats)
    assert stats["max_abs_z"] < 5.5
    assert 0.75 < stats["chi_square_per_dof"] < 1.35
    # Single scattering is a resolved share of the total.
    share = a[:, stokes : 2 * stokes : nmat].sum() / a[:, 0:stokes:nmat].sum()
    assert 0.05 < share < 0.95


# ---- Learned proposal: the float64 restatement of the network ------------------------------


def test_proposal_oracle_gradient_matches_finite_differences():
    """The oracle's analytic gradient (the checker of the device backward pass) against central
    differences of its own loss: a dropped term, a sign or a transposed factor fails."""

    from oracle import proposal as net

    rng = np.random.default_rng(3)
    params = net.initial_parameters(5) + rng.normal(0, 0.05, net.PARAMETERS)
    features = rng.uniform(-1, 1, (6, net.INPUTS))
    directions = rng.normal(size=(6, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    weights = rng.uniform(0.2, 2.0, 6)
    loss, grad = net.loss_and_gradient(params, features, directions, weights)
    picks = np.concatenate([rng.choice(net.PARAMETERS, 150, replace=False),
                            np.arange(net.PARAMETERS - 40, net.PARAMETERS)])  # fmt: skip

    for index in picks:
        step = np.zeros(net.PARAMETERS)
        step[index] = 1e-6
        up = net.loss_and_gradient(params + step, features, directions, weights)[0]
        down = net.loss_and_gradient(params - step, features, directions, weights)[0]
        assert grad[index] == pytest.approx((up - down) / 2e-6, rel=1e-4, abs=1e-9), index


def test_proposal_oracle_density_is_normalized():
    """q integrates to one over the sphere (vMF normalization kappa / (4 pi sinh kappa))."""

    from oracle import proposal as net

    rng = np.random.default_rng(4)
    outputs = rng.normal(0, 1.5, (1, net.OUTPUTS))
    outputs[0, 3::5] = [-3.0, 0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0]  # kappa from 0.05 to ~6
    cos_theta, phi = np.meshgrid(np.linspace(-1, 1, 801), np.linspace(0, 2 * np.pi, 801), indexing="ij")
    s = np.sqrt(1 - cos_theta**2)
    w = np.stack([s * np.cos(phi), s * np.sin(phi), cos_theta], axis=-1).reshape(-1, 3)
    q = np.exp(net.log_density(np.repeat(outputs, len(w), axis=0), w)).reshape(cos_theta.shape)
    total = np.trapezoid(np.trapezoid(q, phi[0], axis=1), cos_theta[:, 0])
    assert total == pytest.approx(1.0, rel=2e-3)
