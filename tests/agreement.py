"""Agreement checks between MONKI products, for the tests.

The acceptance criteria are the reference's product checks (python/monki/validation.py:
20-253), restated:

* physical invariants of one product: the variables a product needs, finite numbers, a
  Stokes axis of 1, 3 or 4 components, I >= 0, |(Q, U, V)| <= I + 1e-2 (mc.f90's
  stokes_criterium_error), a DOP consistent with the Stokes vector, standard errors shaped
  like the radiance and non-negative, and (for statistical use) more than 30 rounds so that
  no standard error carries the 999 sentinel;
* two products agree statistically when every radiance value satisfies
  |a - b| <= atol + z hypot(SE_a, SE_b), z the two-sided normal quantile at the family error
  rate split over all values (Bonferroni); deterministically when every numeric variable
  satisfies |a - b| <= atol + rtol |a|. Both first require equal variable sets, dimensions,
  units, coordinates and scene variables (coordinates within 1e-12).

``z_statistics`` adds the chi-square of the normalized differences used by the GPU tests.
"""

from dataclasses import dataclass, field
from statistics import NormalDist

import numpy as np

from monki.output import SimulationResult

REQUIRED = ("radiance", "std_radiance", "DOP", "wav", "n_rounds", "n_photons_per_column")
ALIGNED = ("wav", "theta", "theta0", "phi", "phi0", "asurf", "x_unit", "y_unit", "z_unit",
           "pres", "temp", "alt", "bmabs", "bmsca", "baer", "vmrgas")  # fmt: skip
POLARIZATION_SLACK = 1e-2
ALIGNMENT_TOLERANCE = 1e-12
SENTINEL = 999
ROUNDS_FOR_UNCERTAINTY = 31


def views(result: SimulationResult, angles) -> SimulationResult:
    """The product restricted to the given forward view zeniths (each must occur exactly once)."""

    if result.dimensions.get("theta") != ("nvza",):
        raise ValueError("selecting views needs a forward theta(nvza) coordinate")

    theta = np.asarray(result.values["theta"], dtype=float)
    picks = []

    for angle in angles:
        hits = np.nonzero(np.abs(theta - angle) <= ALIGNMENT_TOLERANCE)[0]

        if hits.size != 1:
            raise ValueError(f"view zenith {angle} occurs {hits.size} times")

        picks.append(int(hits[0]))

    values = {}

    for name, data in result.values.items():
        axes = result.dimensions[name]
        values[name] = np.take(data, picks, axis=axes.index("nvza")) if "nvza" in axes else np.array(data, copy=True)

    return SimulationResult(result.path, values, dict(result.dimensions), dict(result.attributes))


def invariant_failures(result: SimulationResult, *, need_uncertainty: bool = True) -> list:
    """The physical invariants a product breaks (empty when it is sound)."""

    absent = [name for name in REQUIRED if name not in result.values]

    if absent:
        return [f"variable missing: {name}" for name in absent]

    broken = [f"not finite: {name}" for name, data in result.values.items()
              if np.issubdtype(np.asarray(data).dtype, np.number) and not np.isfinite(data).all()]  # fmt: skip
    axes = result.dimensions["radiance"]

    if "nmat" not in axes:
        return broken + ["radiance lacks the Stokes (nmat) axis"]

    stokes = np.moveaxis(np.asarray(result.radiance, dtype=float), axes.index("nmat"), -1)
    intensity, polarized = stokes[..., 0], np.sqrt((stokes[..., 1:] ** 2).sum(axis=-1))

    if stokes.shape[-1] not in (1, 3, 4):
        broken.append(f"Stokes axis of {stokes.shape[-1]} components")

    if (intensity < -1e-12).any():
        broken.append("intensity below zero")

    if (polarized > intensity + POLARIZATION_SLACK).any():
        broken.append("polarized part larger than the intensity")

    dop = np.zeros_like(intensity)
    np.divide(polarized, intensity, out=dop, where=intensity > 0)
    stored = np.asarray(result.values["DOP"])

    if stored.shape != dop.shape or not np.allclose(stored, dop, rtol=1e-9, atol=1e-10):
        broken.append("DOP does not follow from the Stokes vector")

    error = np.asarray(result.standard_error)

    if error.shape != np.shape(result.radiance) or result.dimensions.get("std_radiance") != axes:
        broken.append("standard error shaped unlike the radiance")

    if (error < 0).any():
        broken.append("standard error below zero")

    rounds = int(np.asarray(result.values["n_rounds"]).ravel()[0])

    if need_uncertainty and (rounds < ROUNDS_FOR_UNCERTAINTY or (error == SENTINEL).any()):
        broken.append(f"no standard error: {rounds} rounds, more than 30 needed")

    if (np.asarray(result.values["wav"]) <= 0).any():
        broken.append("wavelength not positive")

    broken += [f"{name} not positive" for name in ("n_rounds", "n_photons_per_column")
               if (np.asarray(result.values[name]) <= 0).any()]  # fmt: skip
    return broken


@dataclass(frozen=True)
class Agreement:
    passed: bool
    statistical: bool
    values: int = 0
    outside: int = 0
    largest_difference: float = 0.0
    largest_ratio: float = 0.0  # |a - b| / bound
    problems: tuple = field(default_factory=tuple)


def _structure_problems(a: SimulationResult, b: SimulationResult, statistical: bool) -> list:
    problems = [f"reference: {p}" for p in invariant_failures(a, need_uncertainty=statistical)]
    problems += [f"candidate: {p}" for p in invariant_failures(b, need_uncertainty=statistical)]

    if a.values.keys() != b.values.keys():
        problems.append("different variables")

    for name in sorted(a.values.keys() & b.values.keys()):
        shaped = a.dimensions[name] == b.dimensions[name] and np.shape(a.values[name]) == np.shape(b.values[name])

        # Statistical agreement allows different round counts, hence convergence histories.
        if not shaped and not (statistical and name == "radiance_convergence"):
            problems.append(f"{name}: dimensions or shape differ")

        if a.attributes.get(name, {}).get("units") != b.attributes.get(name, {}).get("units"):
            problems.append(f"{name}: units differ")

    for name in ALIGNED:
        if name in a.values and name in b.values:
            x, y = np.asarray(a.values[name]), np.asarray(b.values[name])

            if x.shape != y.shape or (np.abs(x - y) > ALIGNMENT_TOLERANCE).any():
                problems.append(f"{name}: not aligned")

    return problems


def agree(reference: SimulationResult, candidate: SimulationResult, *, statistical: bool = True,
          atol: float = 1e-10, rtol: float = 1e-10, family_error_rate: float = 1e-6) -> Agreement:  # fmt: skip
    """Statistical agreement of the radiances, or deterministic agreement of every variable."""

    if not (np.isfinite(atol) and np.isfinite(rtol) and atol >= 0 and rtol >= 0):
        raise ValueError("tolerances must be finite and not negative")

    if not 0.0 < family_error_rate < 1.0:
        raise ValueError("the family error rate is a probability in (0, 1)")

    problems = _structure_problems(reference, candidate, statistical)

    if problems:
        return Agreement(False, statistical, problems=tuple(problems))

    count = outside = 0
    largest = ratio = 0.0

    for name in ["radiance"] if statistical else sorted(reference.values):
        a, b = np.asarray(reference.values[name]), np.asarray(candidate.values[name])

        if not np.issubdtype(a.dtype, np.number):
            if not np.array_equal(a, b):
                problems.append(f"{name}: values differ")
            continue

        if a.shape != b.shape:
            continue

        difference = np.abs(a.astype(float) - b.astype(float))

        if statistical:
            # Two-sided tail of family_error_rate / values, floored at machine epsilon.
            tail = max(family_error_rate / (2.0 * a.size), np.finfo(float).eps)
            bound = atol + NormalDist().inv_cdf(1.0 - tail) * np.hypot(reference.standard_error,
                                                                      candidate.standard_error)  # fmt: skip
        else:
            bound = atol + rtol * np.abs(a)

        beyond = difference > bound
        count += a.size
        outside += int(beyond.sum())
        largest = max(largest, float(difference.max(initial=0.0)))
        ratio = max(ratio, float((difference / np.maximum(bound, np.finfo(float).tiny)).max(initial=0.0)))

        if beyond.any():
            problems.append(f"{name}: {int(beyond.sum())} of {a.size} values outside the bound")

    return Agreement(not problems, statistical, count, outside, largest, ratio, tuple(problems))


def z_statistics(a, se_a, b, se_b, floor: float = 0.0) -> dict:
    """Per-value z = (a - b) / hypot(se_a, se_b) over values whose combined SE exceeds floor
    (Q and U at a nadir view are exactly zero in both): largest |z|, chi-square, per degree of
    freedom, and the fraction beyond 3."""

    a, b = np.asarray(a, float), np.asarray(b, float)
    combined = np.hypot(np.asarray(se_a, float), np.asarray(se_b, float))
    kept = combined > floor
    z = (a - b)[kept] / combined[kept]
    square = float(np.sum(z * z))
    return {
        "count": int(z.size),
        "max_abs_z": float(np.abs(z).max(initial=0.0)),
        "chi_square": square,
        "chi_square_per_dof": square / max(z.size, 1),
        "fraction_above_3": float(np.mean(np.abs(z) > 3)) if z.size else 0.0,
    }
