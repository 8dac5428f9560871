"""Scientific output checks and deterministic/statistical comparisons.

Restates ``python/monki/validation.py:20-253``: physical invariants of a product
(``check_result``) and the comparison of two products either deterministically or
statistically with combined standard errors and a Bonferroni normal bound
``|d| <= atol + z * hypot(SE_a, SE_b)`` at a family error rate (``validation.py:229-239``).
``chi_square`` adds the per-value chi-square of the normalized differences.
"""

from dataclasses import asdict, dataclass
from statistics import NormalDist
from typing import Tuple

import numpy as np

from .output import SimulationResult

COORDINATES = ("wav", "theta", "theta0", "phi", "phi0", "asurf", "x_unit", "y_unit", "z_unit")
SCENE_VARIABLES = ("pres", "temp", "alt", "bmabs", "bmsca", "baer", "vmrgas")
STOKES_PHYSICAL_TOLERANCE = 1e-2  # mc.f90 stokes_criterium_error
COORDINATE_TOLERANCE = 1e-12
UNCERTAINTY_SENTINEL = 999
MINIMUM_UNCERTAINTY_ROUNDS = 31


def select_views(result: SimulationResult, angles) -> SimulationResult:
    """Select exact forward viewing coordinates from a retained result."""

    if "theta" not in result.values or result.dimensions["theta"] != ("nvza",):
        raise ValueError("view selection requires forward theta(nvza)")

    indices = []

    for angle in angles:
        matches = np.flatnonzero(np.isclose(result.values["theta"], angle, rtol=0, atol=1e-12))

        if len(matches) != 1:
            raise ValueError(f"view {angle} is absent or ambiguous")

        indices.append(int(matches[0]))

    values = {
        name: np.take(data, indices, axis=result.dimensions[name].index("nvza"))
        if "nvza" in result.dimensions[name]
        else data.copy()
        for name, data in result.values.items()
    }
    return SimulationResult(result.path, values, dict(result.dimensions), dict(result.attributes))


@dataclass(frozen=True)
class Comparison:
    passed: bool
    mode: str
    compared_values: int
    failed_values: int
    max_absolute_error: float
    max_normalized_error: float
    details: Tuple[str, ...]

    def to_dict(self):
        return asdict(self)


def check_result(result: SimulationResult, *, require_uncertainty: bool = True) -> list:
    """Check Stokes geometry, positivity, polarization bounds and available uncertainty."""

    failures = []

    for name in ("radiance", "std_radiance", "DOP", "wav", "n_rounds", "n_photons_per_column"):
        if name not in result.values:
            failures.append(f"missing variable: {name}")

    if failures:
        return failures

    for name, data in result.values.items():
        if np.issubdtype(np.asarray(data).dtype, np.number) and not np.all(np.isfinite(data)):
            failures.append(f"non-finite values: {name}")

    if "nmat" not in result.dimensions["radiance"]:
        return failures + ["radiance has no nmat dimension"]

    axis = result.dimensions["radiance"].index("nmat")
    stokes = np.moveaxis(result.radiance, axis, -1)

    if stokes.shape[-1] not in (1, 3, 4):
        failures.append("Stokes vector must contain 1, 3, or 4 elements")

    intensity = stokes[..., 0]

    if np.any(intensity < -1e-12):
        failures.append("negative intensity")

    polarized = np.linalg.norm(stokes[..., 1:], axis=-1)

    if np.any(polarized > intensity + STOKES_PHYSICAL_TOLERANCE):
        failures.append("polarized magnitude exceeds intensity")

    expected_dop = np.divide(polarized, intensity, out=np.zeros_like(intensity), where=intensity > 0)

    if result.values["DOP"].shape != intensity.shape or not np.allclose(
        result.values["DOP"], expected_dop, rtol=1e-9, atol=1e-10
    ):
        failures.append("DOP is inconsistent with Stokes components")

    error = result.standard_error

    if (
        error.shape != result.radiance.shape
        or result.dimensions["std_radiance"] != result.dimensions["radiance"]
    ):
        failures.append("standard-error shape/dimensions differ from radiance")

    if np.any(error < 0):
        failures.append("negative standard error")

    if require_uncertainty and (
        int(np.asarray(result.values["n_rounds"]).ravel()[0]) < MINIMUM_UNCERTAINTY_ROUNDS
        or np.any(error == UNCERTAINTY_SENTINEL)
    ):
        failures.append("uncertainty unavailable: need more than 30 completed rounds")

    if np.any(result.values["wav"] <= 0):
        failures.append("non-positive wavelength")

    for name in ("n_rounds", "n_photons_per_column"):
        if np.any(result.values[name] <= 0):
            failures.append(f"non-positive {name}")

    return failures


def compare(
    reference: SimulationResult,
    candidate: SimulationResult,
    *,
    mode: str = "statistical",
    absolute_tolerance: float = 1e-10,
    relative_tolerance: float = 1e-10,
    family_error_rate: float = 1e-6,
) -> Comparison:
    """Compare aligned outputs; statistical mode checks radiance against combined SE."""

    if mode not in {"statistical", "deterministic"}:
        raise ValueError("mode must be statistical or deterministic")

    if not all(np.isfinite(x) and x >= 0 for x in (absolute_tolerance, relative_tolerance)):
        raise ValueError("tolerances must be finite and non-negative")

    if not 0 < family_error_rate < 1:
        raise ValueError("family_error_rate must be in (0, 1)")

    statistical = mode == "statistical"
    problems = [f"reference: {x}" for x in check_result(reference, require_uncertainty=statistical)]
    problems += [f"candidate: {x}" for x in check_result(candidate, require_uncertainty=statistical)]

    if set(reference.values) != set(candidate.values):
        problems.append("variable inventories differ")

    for name in reference.values.keys() & candidate.values.keys():
        different = (
            reference.dimensions[name] != candidate.dimensions[name]
            or np.shape(reference.values[name]) != np.shape(candidate.values[name])
        )

        if different and not (statistical and name == "radiance_convergence"):
            problems.append(f"shape/dimension mismatch: {name}")

        if reference.attributes.get(name, {}).get("units") != candidate.attributes.get(
            name, {}
        ).get("units"):
            problems.append(f"units mismatch: {name}")

    for name in (*COORDINATES, *SCENE_VARIABLES):
        if name in reference.values and name in candidate.values:
            a, b = reference.values[name], candidate.values[name]

            if np.shape(a) != np.shape(b) or not np.allclose(a, b, rtol=0, atol=COORDINATE_TOLERANCE):
                problems.append(f"coordinate mismatch: {name}")

    if problems:
        return Comparison(False, mode, 0, 0, 0.0, 0.0, tuple(problems))

    names = ["radiance"] if statistical else sorted(reference.values)
    count = failed = 0
    max_absolute = max_normalized = 0.0

    for name in names:
        a, b = np.asarray(reference.values[name]), np.asarray(candidate.values[name])

        if not np.issubdtype(a.dtype, np.number):
            if not np.array_equal(a, b):
                problems.append(f"non-numeric values differ: {name}")

            continue

        if a.shape != b.shape:
            continue

        delta = np.abs(a.astype(float) - b.astype(float))

        if statistical:
            tail = max(family_error_rate / (2 * a.size), np.finfo(float).eps)
            z = NormalDist().inv_cdf(1 - tail)
            bound = absolute_tolerance + z * np.hypot(
                reference.standard_error, candidate.standard_error
            )
        else:
            bound = absolute_tolerance + relative_tolerance * np.abs(a)

        bad = delta > bound
        count += a.size
        failed += int(np.count_nonzero(bad))
        max_absolute = max(max_absolute, float(np.max(delta, initial=0)))
        normalized = delta / np.maximum(bound, np.finfo(float).tiny)
        max_normalized = max(max_normalized, float(np.max(normalized, initial=0)))

        if np.any(bad):
            problems.append(f"{name}: {np.count_nonzero(bad)}/{a.size} values exceed tolerance")

    return Comparison(not problems, mode, count, failed, max_absolute, max_normalized, tuple(problems))


def combined_se_statistics(a, se_a, b, se_b, floor: float = 0.0) -> dict:
    """Per-value |a - b| / hypot(se_a, se_b): maximum, chi-square and degrees of freedom.

    Values whose combined SE is not above ``floor`` (for example Q/U at a nadir view, where
    both estimates are exactly zero) are excluded.
    """

    a, b = np.asarray(a, float), np.asarray(b, float)
    combined = np.hypot(np.asarray(se_a, float), np.asarray(se_b, float))
    used = combined > floor
    z = (a - b)[used] / combined[used]
    return {
        "count": int(z.size),
        "max_abs_z": float(np.max(np.abs(z), initial=0.0)),
        "chi_square": float(np.sum(z * z)),
        "chi_square_per_dof": float(np.sum(z * z) / max(z.size, 1)),
        "fraction_above_3": float(np.mean(np.abs(z) > 3)) if z.size else 0.0,
    }
