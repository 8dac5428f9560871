"""Stopping on the estimated precision of the layer air-mass factors.

Estimator. For one (pixel, layer, view) cell, round r gives a weighted layer path P_r and a
detector weight W_r; the layer AMF is the ratio of means A = R / dz with R = mean(P) / mean(W).
Its standard error is the delta-method (linearised) error of a ratio of means,

    SE(A) = sqrt( sum_r (P_r - R W_r)^2 / (n (n - 1)) ) / mean(W) / dz,

the estimator of the reference's stopping rule (python/monki/convergence.py:76-99). The sum
is expanded in centred second moments, S_PP - 2 R S_PW + R^2 S_WW, so only five arrays per
cell are kept whatever the number of rounds.

Accumulation. Rounds are buffered and folded in as groups with the pairwise update of Chan,
Golub and LeVeque (1979): a group's own means and centred co-moments are computed in one
vectorised pass, then merged into the running ones with the correction term
delta_P delta_W n_a n_b / (n_a + n_b). A device that keeps its own online moments (the batch
backend) hands them over through ``RatioMoments.from_moments``.

Policy. ``AMFConvergence`` keeps the reference's fields and meaning (convergence.py:10-34):
rounds run in groups of ``check_every``; from ``minimum_rounds`` on, every group end is a
check, which passes when every physical cell's relative SE is finite and below
``relative_error``; the run stops after ``consecutive_checks`` passing checks in a row, and
``case.rounds`` stays a hard cap. The vacuum layer (last) never enters.
"""

import math
from dataclasses import dataclass

import numpy as np

from .execution.tallies import BatchControls, RawTallies

# Largest group folded in at once: bounds the buffer of a large grid (a 128 x 128 x 60 scene
# holds 23 MB of paths per round).
_GROUP_LIMIT = 8


def _positive_int(value) -> bool:
    return isinstance(value, int) and not isinstance(value, bool) and value > 0


@dataclass(frozen=True)
class AMFConvergence:
    """Target relative standard error of every layer AMF and how often it is checked."""

    relative_error: float = 0.01
    minimum_rounds: int = 128
    check_every: int = 32
    consecutive_checks: int = 3

    def __post_init__(self):
        target = self.relative_error

        if isinstance(target, bool) or not isinstance(target, (int, float)) or not math.isfinite(target):
            raise ValueError(f"relative_error={target!r}: expected a finite number")

        if not 0.0 < target < 1.0:
            raise ValueError(f"relative_error={target!r}: expected a fraction strictly between 0 and 1")

        for name in ("minimum_rounds", "check_every", "consecutive_checks"):
            if not _positive_int(getattr(self, name)):
                raise ValueError(f"{name}={getattr(self, name)!r}: expected an integer of at least 1")

        if self.minimum_rounds < 32:
            # The error estimate of the mean over rounds is unstable below about 30 rounds, the
            # same bound that makes the radiance SE unavailable through round 30.
            raise ValueError(f"minimum_rounds={self.minimum_rounds}: checks start at 32 rounds or later")


class RatioMoments:
    """Running means and centred co-moments of per-round layer paths and detector weights.

    Shapes follow the tallies: paths (nx, ny, nz - 1, views) over the physical layers, weights
    (nx, ny, 1, views) so that they broadcast over layers.
    """

    def __init__(self):
        self.count = 0
        self.mean_path = self.mean_weight = None
        self.s_pp = self.s_ww = self.s_pw = None
        self._pending_paths = []
        self._pending_weights = []

    @classmethod
    def from_moments(cls, count, mean_path, mean_weight, s_pp, s_ww, s_pw):
        """Moments accumulated elsewhere (weights given as (nx, ny, views))."""

        moments = cls()
        moments.count = int(count)
        moments.mean_path = np.asarray(mean_path, dtype=np.float64)
        moments.s_pp = np.asarray(s_pp, dtype=np.float64)
        moments.s_pw = np.asarray(s_pw, dtype=np.float64)
        moments.mean_weight = np.asarray(mean_weight, dtype=np.float64)[:, :, None, :]
        moments.s_ww = np.asarray(s_ww, dtype=np.float64)[:, :, None, :]
        return moments

    @property
    def rounds(self) -> int:
        return self.count + len(self._pending_paths)

    def add(self, batch: RawTallies) -> None:
        """Take one complete round (its physical layers)."""

        if batch.weighted_paths is None or batch.weights is None:
            raise ValueError("the AMF precision needs layer-path tallies: run with save_path=True")

        paths = np.asarray(batch.weighted_paths, dtype=np.float64)[:, :, :-1, :]
        weights = np.asarray(batch.weights, dtype=np.float64)[:, :, None, :]

        if not (np.isfinite(paths).all() and np.isfinite(weights).all()):
            raise ValueError(f"round {batch.controls.round_index}: layer-path or weight tally is not finite")

        self._pending_paths.append(paths)
        self._pending_weights.append(weights)

        if len(self._pending_paths) >= _GROUP_LIMIT:
            self._fold()

    def _fold(self) -> None:
        if not self._pending_paths:
            return

        paths = np.stack(self._pending_paths)
        weights = np.stack(self._pending_weights)
        self._pending_paths, self._pending_weights = [], []
        size = paths.shape[0]
        group_path, group_weight = paths.mean(axis=0), weights.mean(axis=0)
        centred_path, centred_weight = paths - group_path, weights - group_weight
        group_pp = np.einsum("r...,r...->...", centred_path, centred_path)
        group_ww = np.einsum("r...,r...->...", centred_weight, centred_weight)
        group_pw = np.einsum("r...,r...->...", centred_path, np.broadcast_to(centred_weight, centred_path.shape))

        if self.count == 0:
            self.count = size
            self.mean_path, self.mean_weight = group_path, group_weight
            self.s_pp, self.s_ww, self.s_pw = group_pp, group_ww, group_pw
            return

        total = self.count + size
        step_path, step_weight = group_path - self.mean_path, group_weight - self.mean_weight
        mix = self.count * size / total
        self.s_pp = self.s_pp + group_pp + step_path * step_path * mix
        self.s_ww = self.s_ww + group_ww + step_weight * step_weight * mix
        self.s_pw = self.s_pw + group_pw + step_path * step_weight * mix
        self.mean_path = self.mean_path + step_path * (size / total)
        self.mean_weight = self.mean_weight + step_weight * (size / total)
        self.count = total

    def estimate(self, thickness):
        """Layer AMF and its standard error, both (nx, ny, nz - 1, views); thickness in km."""

        self._fold()

        if self.count < 2:
            raise ValueError(f"the AMF error needs two or more rounds, have {self.count}")

        dz = np.asarray(thickness, dtype=np.float64).reshape(1, 1, -1, 1)

        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = self.mean_path / self.mean_weight
            spread = self.s_pp - 2.0 * ratio * self.s_pw + ratio * ratio * self.s_ww
            # In exact arithmetic the spread is a sum of squares; only roundoff (relative to the
            # magnitude of the terms that cancel) may take it below zero.
            magnitude = self.s_pp + np.abs(2.0 * ratio * self.s_pw) + ratio * ratio * self.s_ww
            floor = -128.0 * np.finfo(np.float64).eps * np.maximum(magnitude, np.finfo(np.float64).tiny)

            if np.any(spread < floor):
                raise ArithmeticError("AMF residual sum of squares is negative beyond roundoff")

            error = np.sqrt(np.clip(spread, 0.0, None) / (self.count * (self.count - 1.0))) / self.mean_weight

        return ratio / dz, error / dz

    def relative_error(self, thickness) -> np.ndarray:
        """SE / |AMF| per cell; infinite where the AMF is zero."""

        amf, error = self.estimate(thickness)
        relative = np.full(amf.shape, np.inf)
        np.divide(error, np.abs(amf), out=relative, where=amf != 0)
        return relative


class PrecisionStopping:
    """Schedules rounds in check groups and records each precision check."""

    def __init__(self, policy: AMFConvergence, thickness):
        self.policy = policy
        self.thickness = np.asarray(thickness, dtype=np.float64)
        self.moments = RatioMoments()
        self.history = []
        self.converged = False
        self._passes = 0

    @property
    def streak(self) -> int:
        return self._passes

    def schedule(self, backend, case):
        """Yield the rounds of ``case`` as the backend returns them, group by group, until the
        policy is met or the round cap is reached."""

        every = self.policy.check_every

        for first in range(1, case.rounds + 1, every):
            last = min(case.rounds, first + every - 1)
            requests = [BatchControls(index, case.photons_per_column) for index in range(first, last + 1)]
            returned = backend.run_batches(requests)

            if len(returned) != len(requests):
                raise ValueError(f"backend returned {len(returned)} rounds for {len(requests)} requested")

            for request, batch in zip(requests, returned):
                if not isinstance(batch, RawTallies) or batch.controls != request:
                    raise ValueError(f"backend answered round {request.round_index} with another round's tallies")

                self.moments.add(batch)
                yield batch

            if self.check(last):
                return

    def check(self, rounds: int) -> bool:
        """The precision check at a group end after ``rounds`` rounds; True once converged."""

        if rounds < self.policy.minimum_rounds:
            return False

        relative = self.moments.relative_error(self.thickness)
        worst = float(relative.max())
        finite = bool(np.isfinite(relative).all())
        below = finite and worst < self.policy.relative_error
        self._passes = self._passes + 1 if below else 0
        self.converged = self._passes >= self.policy.consecutive_checks
        self.history.append(
            {"rounds": rounds, "max_relative_se": worst if finite else None, "below_target": below,
             "streak": self._passes}  # fmt: skip
        )
        return self.converged
