"""Optical inputs of the synthetic scenes: Rayleigh, ozone, Mie and Henyey-Greenstein tables.

These produce the per-layer optical depths and tabulated phase matrices that the reference
obtains from its Fortran preparation (``main_routines_general_mod.f90``). The Mie routine
is the Bohren & Huffman (1983, appendix A) series with logarithmic-derivative downward
recurrence, averaged over the two-parameter gamma size distribution of Hansen & Travis
(1974, eq. 2.56). Phase matrices use the reference's normalization: the mean of F11 over the
sphere is one, and F12 < 0 means polarization perpendicular to the scattering plane (the
sign of ``ray_scat_matrix``, ``mc_routines_general_mod.f90:110-125``).
"""

import numpy as np

from .scene import Species

# main_routines_general_mod.f90:68-71 — dry-air composition in volume percent.
AIR_FRACTIONS = {"N2": 78.084, "O2": 20.9476, "Ar": 0.934, "CO2": 0.0375}


def rayleigh_depolarization(wavelength_nm: float) -> float:
    """Depolarization factor from the Bodhaine et al. (1999) King factors.

    Restates ``main_routines_general_mod.f90:199-218``.
    """

    sig2 = (1.0e3 / wavelength_nm) ** 2
    king = {
        "N2": 1.034 + 3.17e-4 * sig2,
        "O2": 1.096 + 1.385e-3 * sig2 + 1.448e-4 * sig2**2,
        "Ar": 1.0,
        "CO2": 1.15,
    }
    total = sum(AIR_FRACTIONS.values())
    dep = sum(AIR_FRACTIONS[name] * king[name] for name in king) / total
    return (6.0 * dep - 6.0) / (3.0 + 7.0 * dep)


def exponential_layer_depths(levels_km, total: float, scale_height_km: float) -> np.ndarray:
    """Optical depth per layer of an exponential profile normalized to ``total``."""

    levels = np.asarray(levels_km, dtype=np.float64)
    e = np.exp(-levels / scale_height_km)
    return total * (e[:-1] - e[1:]) / (e[0] - e[-1])


def gaussian_layer_depths(levels_km, total: float, centre_km: float, width_km: float):
    """Optical depth per layer of a Gaussian profile normalized to ``total``."""

    from math import erf, sqrt

    levels = np.asarray(levels_km, dtype=np.float64)
    cdf = np.array([0.5 * (1 + erf((z - centre_km) / (sqrt(2) * width_km))) for z in levels])
    depth = np.diff(cdf)
    return total * depth / depth.sum()


def henyey_greenstein(g: float, angles: int = 181, polarization: float = 0.3) -> dict:
    """Tabulated polarized Henyey-Greenstein matrix on an equal-angle grid.

    F11 is the Henyey-Greenstein function (mean one over the sphere). The polarized
    elements take the Rayleigh angular shapes scaled to F11: F22 = F11,
    F33 = F44 = F11 * 2 mu / (1 + mu^2), F12 = -p F11 (1 - mu^2) / (1 + mu^2) with
    p = ``polarization``, F34 = 0. For p <= 1 this keeps F12^2 + F33^2 <= F11^2.
    """

    theta = np.linspace(np.pi, 0.0, angles)
    mu = np.cos(theta)
    mu[0], mu[-1] = -1.0, 1.0
    f11 = (1 - g * g) / (1 + g * g - 2 * g * mu) ** 1.5
    shape33 = 2 * mu / (1 + mu * mu)
    shape12 = -(1 - mu * mu) / (1 + mu * mu)
    return {
        "cosines": mu,
        "f11": f11,
        "f22": f11.copy(),
        "f33": f11 * shape33,
        "f44": f11 * shape33,
        "f12": polarization * f11 * shape12,
        "f34": np.zeros_like(f11),
    }


def _mie_coefficients(x: np.ndarray, m: complex, nmax: int):
    """Return a_n, b_n (radii x nmax) with terms beyond each radius' nstop set to zero."""

    x = np.asarray(x, dtype=np.float64)
    nstop = np.floor(x + 4.0 * np.cbrt(x) + 2.0).astype(int)
    mx = m * x
    nmx = int(max(nmax, np.max(np.abs(mx)))) + 16
    # Downward recurrence of D_n(mx) (Bohren & Huffman, eq. 4.89).
    d = np.zeros((nmx + 1, x.size), dtype=np.complex128)

    for n in range(nmx, 0, -1):
        d[n - 1] = n / mx - 1.0 / (d[n] + n / mx)

    a = np.zeros((x.size, nmax), dtype=np.complex128)
    b = np.zeros((x.size, nmax), dtype=np.complex128)
    psi0, psi1 = np.cos(x), np.sin(x)
    chi0, chi1 = -np.sin(x), np.cos(x)
    xi1 = psi1 - 1j * chi1

    for n in range(1, nmax + 1):
        active = n <= nstop
        psi = (2 * n - 1) / x * psi1 - psi0
        chi = (2 * n - 1) / x * chi1 - chi0
        xi = psi - 1j * chi
        da = d[n] / m + n / x
        db = m * d[n] + n / x
        an = (da * psi - psi1) / (da * xi - xi1)
        bn = (db * psi - psi1) / (db * xi - xi1)
        a[:, n - 1] = np.where(active, an, 0)
        b[:, n - 1] = np.where(active, bn, 0)
        # Freeze the recurrences past nstop: upward psi is unstable there.
        psi0, psi1 = np.where(active, psi1, psi0), np.where(active, psi, psi1)
        chi0, chi1 = np.where(active, chi1, chi0), np.where(active, chi, chi1)
        xi1 = psi1 - 1j * chi1

    return a, b, nstop


def _angular_functions(mu: np.ndarray, nmax: int):
    """pi_n(mu), tau_n(mu) for n = 1..nmax (Bohren & Huffman, eqs. 4.46-4.47)."""

    pi = np.zeros((nmax, mu.size))
    tau = np.zeros((nmax, mu.size))
    p0 = np.zeros_like(mu)
    p1 = np.ones_like(mu)

    for n in range(1, nmax + 1):
        if n == 1:
            p = p1
        else:
            p = ((2 * n - 1) * mu * p1 - n * p0) / (n - 1)
            p0, p1 = p1, p
        pi[n - 1] = p
        tau[n - 1] = n * mu * p - (n + 1) * p0

    return pi, tau


def mie_gamma(
    wavelength_um: float,
    refractive_index: complex,
    reff_um: float,
    veff: float,
    angles: int = 181,
    radii: int = 1200,
    rmin_um: float = None,
    rmax_um: float = None,
) -> dict:
    """Size-averaged Mie scattering matrix on an equal-angle grid from 180 to 0 degrees.

    Returns cosines (ascending), F11..F44 normalized like the reference LUTs
    (``S.lu.*``: mean F11 over the sphere is one), the single-scattering albedo and the
    asymmetry parameter.
    """

    alpha = (1.0 - 3.0 * veff) / veff  # n(r) ~ r^alpha exp(-r / (reff * veff))
    scale = reff_um * veff
    rmax = scale * (alpha + 1 + 14.0 * np.sqrt(alpha + 1)) if rmax_um is None else rmax_um
    rmin = max(1e-4, scale * max(0.0, alpha + 1 - 12.0 * np.sqrt(alpha + 1)))
    rmin = rmin if rmin_um is None else rmin_um
    nodes, weights = np.polynomial.legendre.leggauss(40)
    edges = np.linspace(rmin, rmax, radii // 40 + 1)
    r = np.concatenate([0.5 * (hi - lo) * nodes + 0.5 * (hi + lo) for lo, hi in zip(edges, edges[1:])])
    w = np.concatenate([0.5 * (hi - lo) * weights for lo, hi in zip(edges, edges[1:])])
    logn = alpha * np.log(r) - r / scale
    number = w * np.exp(logn - logn.max())
    k = 2.0 * np.pi / wavelength_um
    x = k * r
    nmax = int(np.max(np.floor(x + 4.0 * np.cbrt(x) + 2.0)))
    a, b, _ = _mie_coefficients(x, refractive_index, nmax)
    n = np.arange(1, nmax + 1)
    theta = np.linspace(np.pi, 0.0, angles)
    mu = np.cos(theta)
    mu[0], mu[-1] = -1.0, 1.0
    pi, tau = _angular_functions(mu, nmax)
    coefficient = (2 * n + 1) / (n * (n + 1))
    s1 = (a * coefficient) @ pi + (b * coefficient) @ tau
    s2 = (a * coefficient) @ tau + (b * coefficient) @ pi
    # Per-particle cross sections times k^2 / pi (so they share the x^2 scaling).
    csca = 2.0 * np.sum((2 * n + 1) * (np.abs(a) ** 2 + np.abs(b) ** 2), axis=1)
    cext = 2.0 * np.sum((2 * n + 1) * np.real(a + b), axis=1)
    i1, i2 = np.abs(s1) ** 2, np.abs(s2) ** 2
    cross = s2 * np.conj(s1)
    total = number @ csca
    # 4 pi <|S|^2 / 2> / (k^2 <C_sca>) with k^2 C_sca = pi * csca.
    norm = 4.0 / total
    f11 = norm * (number @ ((i1 + i2) / 2))
    f12 = norm * (number @ ((i2 - i1) / 2))
    f33 = norm * (number @ np.real(cross))
    f34 = -norm * (number @ np.imag(cross))
    mean_mu = 0.5 * np.trapezoid(f11 * mu, mu)
    return {
        "cosines": mu,
        "f11": f11,
        "f22": f11.copy(),
        "f33": f33,
        "f44": f33.copy(),
        "f12": f12,
        "f34": f34,
        "ssalb": float(total / (number @ cext)),
        "asymmetry": float(mean_mu / (0.5 * np.trapezoid(f11, mu))),
    }


def species_from_table(name: str, table: dict, ssalb: float) -> Species:
    return Species(
        name=name,
        ssalb=float(ssalb),
        cosines=table["cosines"],
        f11=table["f11"],
        f22=table["f22"],
        f33=table["f33"],
        f12=table["f12"],
        f34=table.get("f34"),
        f44=table.get("f44"),
    )
