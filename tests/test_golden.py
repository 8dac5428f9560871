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
