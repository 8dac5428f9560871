"""Synthetic scenes of the benchmark workloads (clear sky, layered cloud, 3-D box clouds).

Layer grid: the levels of the reference's 340 nm clear-sky fixture
(``fortran/monki/upstream/INPUT/testrun/MONKI_MLS_cloudfree_340_340_test.in:58-91``, 0-100 km)
with one extra level at 60 km, giving 34 physical layers plus the vacuum layer (35 in all).

* Rayleigh: exponential with an 8 km scale height, column optical depth 0.70, single
  scattering albedo 1, depolarization from the Bodhaine et al. King factors at 340 nm.
* Ozone: absorption optical depth 0.01 per column, Gaussian in altitude centred at 22 km with
  a 5 km standard deviation; it lowers the gas single-scattering albedo per layer.
* Water cloud: Mie, two-parameter gamma distribution, r_eff = 10 um, v_eff = 0.1,
  m = 1.3442 (the refractive index of the reference's 340 nm Mie data,
  ``DATA/Mie_data/Mie_nrvar_ni0/S.sc.00.340:3``), 181 equal-angle nodes.
* Aerosol: polarized Henyey-Greenstein, g = 0.7, single-scattering albedo 0.92, column
  optical depth 0.2, exponential with a 1.5 km scale height.

Tabulated phase matrices are rescaled so that the trapezoid integral of F11 over the cosine
grid is exactly two: linear interpolation in cosine is then normalized, and the local
estimate and the direction sampling describe the same phase function.
"""

from functools import lru_cache

import numpy as np

from . import optics
from .scene import VACUUM_THICKNESS_KM, Species, TransportScene

FIXTURE_LEVELS_KM = (
    0.0, 0.5, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0, 11.0, 12.0, 13.0, 14.0,
    15.0, 16.0, 17.0, 18.0, 19.0, 20.0, 21.0, 22.0, 23.0, 24.0, 25.0, 30.0, 35.0, 40.0,
    45.0, 50.0, 70.0, 100.0,
)  # fmt: skip
LEVELS_KM = tuple(sorted(FIXTURE_LEVELS_KM + (60.0,)))
WAVELENGTH_NM = 340.0
WATER_INDEX_340 = 1.3442


def _normalized(table: dict) -> dict:
    scale = np.trapezoid(table["f11"], table["cosines"]) / 2.0
    out = dict(table)

    for name in ("f11", "f22", "f33", "f44", "f12", "f34"):
        out[name] = table[name] / scale

    return out


@lru_cache(maxsize=4)
def water_cloud_table(angles: int = 181) -> dict:
    table = optics.mie_gamma(WAVELENGTH_NM / 1000.0, WATER_INDEX_340 + 0j, 10.0, 0.1, angles=angles)
    return _normalized(table)


@lru_cache(maxsize=4)
def aerosol_table(angles: int = 181) -> dict:
    return _normalized(optics.henyey_greenstein(0.7, angles=angles))


def water_cloud(angles: int = 181) -> Species:
    table = water_cloud_table(angles)
    return optics.species_from_table("water_cloud", table, ssalb=min(1.0, table["ssalb"]))


def aerosol(angles: int = 181) -> Species:
    return optics.species_from_table("aerosol_hg", aerosol_table(angles), ssalb=0.92)


def _gas_profile(levels, rayleigh_tau=0.70, ozone_tau=0.01):
    rayleigh = optics.exponential_layer_depths(levels, rayleigh_tau, 8.0)
    ozone = optics.gaussian_layer_depths(levels, ozone_tau, 22.0, 5.0)
    return rayleigh, ozone


def _assemble(
    name,
    nx,
    ny,
    levels,
    species_depths,
    species,
    nmat,
    solar_zenith,
    albedo,
    view_zeniths=(0.0, 30.0, 60.0),
    xunit=1.0,
    yunit=1.0,
    save_path=True,
    wavelength_nm=WAVELENGTH_NM,
    rayleigh_tau=0.70,
    ozone_tau=0.01,
):
    """Build tables from per-column optical depths (tau arrays shaped (nx, ny, layers))."""

    levels = np.asarray(levels, dtype=np.float64)
    thickness = np.diff(levels)
    physical = thickness.size
    rayleigh, ozone = _gas_profile(levels, rayleigh_tau, ozone_tau)
    gas = np.broadcast_to(rayleigh + ozone, (nx, ny, physical))
    depths = [np.broadcast_to(np.asarray(d, dtype=np.float64), (nx, ny, physical)) for d in species_depths]
    total = gas + sum(depths) if depths else gas.copy()
    nz = physical + 1
    zunit = np.concatenate([thickness, [VACUUM_THICKNESS_KM]])
    beta = np.zeros((nx, ny, nz))
    beta[:, :, :physical] = total / thickness
    fractions = np.zeros((len(depths) + 1, nx, ny, nz))

    with np.errstate(invalid="ignore", divide="ignore"):
        for index, depth in enumerate(depths):
            fractions[index, :, :, :physical] = np.where(total > 0, depth / total, 0.0)

        fractions[-1, :, :, :physical] = np.where(total > 0, gas / total, 1.0)

    ssalb_gas = np.ones(nz)
    ssalb_gas[:physical] = rayleigh / (rayleigh + ozone)
    return TransportScene(
        nx=nx,
        ny=ny,
        nz=nz,
        nmat=nmat,
        xunit=xunit,
        yunit=yunit,
        zunit=zunit,
        beta=beta,
        fractions=fractions,
        ssalb_gas=ssalb_gas,
        albedo=np.broadcast_to(np.asarray(albedo, dtype=np.float64), (nx, ny)).copy(),
        depol=optics.rayleigh_depolarization(wavelength_nm),
        solar_zenith=float(solar_zenith),
        solar_azimuth=0.0,
        view_azimuth=0.0,
        view_zeniths=tuple(view_zeniths),
        species=list(species),
        save_path=save_path,
        wavelength_nm=wavelength_nm,
        altitude_m=1000.0 * levels,
        description=name,
    )


def clear_sky(nx: int = 30, ny: int = 20, nmat: int = 1, solar_zenith: float = 40.0, albedo=0.05):
    """Plain clear sky: Rayleigh + ozone over a Lambertian surface (30 x 20 x 35)."""

    return _assemble("clear-sky", nx, ny, LEVELS_KM, [], [], nmat, solar_zenith, albedo)


def cloud_layer_depths(levels, base_km: float, top_km: float, depth: float):
    levels = np.asarray(levels)
    lower = np.clip(levels[:-1], base_km, top_km)
    upper = np.clip(levels[1:], base_km, top_km)
    overlap = upper - lower
    return depth * overlap / overlap.sum()


def layered_cloud(nx: int = 30, ny: int = 20, nmat: int = 3, solar_zenith: float = 40.0, albedo=0.05):
    """Homogeneous water-cloud slab (2-3 km, tau 10) plus HG aerosol, with polarization."""

    cloud = cloud_layer_depths(LEVELS_KM, 2.0, 3.0, 10.0)
    aer = optics.exponential_layer_depths(LEVELS_KM, 0.2, 1.5)
    return _assemble(
        "layered-cloud", nx, ny, LEVELS_KM, [cloud, aer], [water_cloud(), aerosol()], nmat,
        solar_zenith, albedo,
    )  # fmt: skip


BOX_LAYERS = (3, 4, 5, 6)  # zero-based physical layers 2-6 km


def box_cloud(
    nx: int = 30,
    ny: int = 20,
    nmat: int = 3,
    solar_zenith: float = 40.0,
    albedo=0.05,
    optical_depth: float = 20.0,
    box=(10, 6),
    offset=(0, 0),
):
    """A cloud block of box[0] x box[1] columns over layers 3-6 in the clear-sky grid."""

    physical = len(LEVELS_KM) - 1
    depth = np.zeros((nx, ny, physical))
    thickness = np.diff(np.asarray(LEVELS_KM))[list(BOX_LAYERS)]
    per_layer = optical_depth * thickness / thickness.sum()
    x0, y0 = offset
    columns_x = [(x0 + i) % nx for i in range(box[0])]
    columns_y = [(y0 + j) % ny for j in range(box[1])]

    for i in columns_x:
        for j in columns_y:
            depth[i, j, list(BOX_LAYERS)] = per_layer

    return _assemble(
        "box-cloud", nx, ny, LEVELS_KM, [depth], [water_cloud()], nmat, solar_zenith, albedo
    )


def box_cloud_batch(count: int = 64, seed: int = 2026):
    """Scenes drawn from one seed: SZA U(0, 70), albedo U(0.02, 0.3), cloud optical depth
    lognormal (median 15, sigma 0.5), grid 16-30 x 12-20 columns, box 10 x 6 x layers 3-6."""

    rng = np.random.default_rng(seed)
    scenes = []

    for _ in range(count):
        sza = float(rng.uniform(0.0, 70.0))
        albedo = float(rng.uniform(0.02, 0.3))
        cot = float(np.exp(np.log(15.0) + 0.5 * rng.standard_normal()))
        nx = int(rng.integers(16, 31))
        ny = int(rng.integers(12, 21))
        offset = (int(rng.integers(0, nx)), int(rng.integers(0, ny)))
        scenes.append(box_cloud(nx, ny, 3, sza, albedo, cot, offset=offset))

    return scenes


# ---- Spectral 3-D stochastic cloud field ---------------------------------------------------

# 59 physical layers + vacuum (60): 0.25 km to 5 km, 1 km to 20 km, 2 km to 50 km, then 9 to 100 km.
SPECTRAL_LEVELS_KM = tuple(
    np.round(np.concatenate([np.arange(0.0, 5.0, 0.25), np.arange(5.0, 20.0, 1.0), np.arange(20.0, 50.0, 2.0),
                             np.linspace(50.0, 100.0, 10)]), 6)  # fmt: skip
)


def kolmogorov_field(n: int, seed: int) -> np.ndarray:
    """Zero-mean, unit-variance periodic Gaussian field with 2-D power |k|^-8/3 (1-D k^-5/3)."""

    rng = np.random.default_rng(seed)
    k = np.fft.fftfreq(n) * n
    kx, ky = np.meshgrid(k, k, indexing="ij")
    magnitude = np.hypot(kx, ky)
    amplitude = np.zeros_like(magnitude)
    np.power(magnitude, -8.0 / 6.0, out=amplitude, where=magnitude > 0)
    noise = np.fft.fft2(rng.standard_normal((n, n)))
    field = np.real(np.fft.ifft2(noise * amplitude))
    return (field - field.mean()) / field.std()


def stochastic_cloud_field(n: int = 128, mean_tau: float = 12.0, cover: float = 0.6, sigma: float = 0.8,
                           seed: int = 7) -> np.ndarray:
    """Column cloud optical depth on an n x n periodic grid: a Gaussian field with a
    Kolmogorov horizontal spectrum (1-D spectrum k^-5/3, 2-D power |k|^-8/3), thresholded to the
    cloud cover, lognormal (exp(sigma g)) in the cloudy columns, scaled to the mean optical depth
    of the cloudy columns."""

    field = kolmogorov_field(n, seed)
    cloudy = field > np.quantile(field, 1.0 - cover)
    tau = np.where(cloudy, np.exp(sigma * field), 0.0)
    return tau * (mean_tau / tau[cloudy].mean())


def spectral_wavelengths(count: int = 64) -> np.ndarray:
    return np.linspace(300.0, 500.0, count)


def spectral_cloud_field(wavelength_nm: float, field: np.ndarray, nmat: int = 3, solar_zenith: float = 40.0,
                         albedo=0.05) -> TransportScene:
    """One wavelength of the spectral case: the field's cloud between 1 and 4 km (extinction
    uniform in height), Rayleigh column 0.70 (lambda / 340 nm)^-4.08, a synthetic ozone column
    0.01 exp(-(lambda - 340 nm) / 12 nm) (Huggins-band-like decrease), depolarization at the
    wavelength. The water-cloud Mie table is the 340 nm one at every wavelength (droplet
    scattering varies little over 300-500 nm for r_eff 10 um)."""

    levels = np.asarray(SPECTRAL_LEVELS_KM)
    n = field.shape[0]
    cloud = cloud_layer_depths(levels, 1.0, 4.0, 1.0)[None, None, :] * field[:, :, None]
    rayleigh_tau = 0.70 * (wavelength_nm / 340.0) ** -4.08
    ozone_tau = 0.01 * np.exp(-(wavelength_nm - 340.0) / 12.0)
    scene = _assemble(
        f"spectral-{wavelength_nm:.1f}nm", n, field.shape[1], levels, [cloud], [water_cloud()], nmat, solar_zenith,
        albedo, wavelength_nm=float(wavelength_nm), rayleigh_tau=rayleigh_tau, ozone_tau=ozone_tau,
    )  # fmt: skip
    thickness = np.diff(levels)
    rayleigh, ozone = _gas_profile(levels, rayleigh_tau, ozone_tau)
    return _factor_extinction(scene, cloud / thickness, (rayleigh + ozone) / thickness)


def _factor_extinction(scene: TransportScene, species_extinction, gas_extinction) -> TransportScene:
    """Rebuild the extinction and species-fraction tables of a one-species scene from float32
    parts (species per cell, gas per layer), so that beta = species + gas and the fraction =
    species / beta hold exactly in float32, and keep the parts: every wavelength of one cloud
    field then carries the same species array (differences from the float64 tables: one float32
    rounding)."""

    nx, ny, nz = scene.nx, scene.ny, scene.nz
    species = np.zeros((nx, ny, nz), dtype=np.float32)
    species[:, :, : nz - 1] = np.asarray(species_extinction, dtype=np.float64)
    gas = np.zeros(nz, dtype=np.float32)
    gas[: nz - 1] = np.asarray(gas_extinction, dtype=np.float64)
    beta = species + gas[None, None, :]
    positive = beta > 0

    with np.errstate(invalid="ignore", divide="ignore"):
        fraction = np.where(positive, species / beta, np.float32(0))

    fractions = np.zeros_like(scene.fractions)
    fractions[0] = fraction
    fractions[-1] = np.where(positive, 1.0 - fractions[0], scene.fractions[-1])
    scene.beta = beta.astype(np.float64)
    scene.fractions = fractions
    scene.extinction_parts = (species, gas)
    return scene


def spectral_scenes(count: int = 64, n: int = 128, seed: int = 7, nmat: int = 3) -> list:
    """The spectral batch: one scene per wavelength, all on the same cloud field."""

    field = stochastic_cloud_field(n, seed=seed)
    return [spectral_cloud_field(w, field, nmat) for w in spectral_wavelengths(count)]
