"""Round statistics, the NetCDF product and the run receipt, for any transport backend.

What is computed (the product definition of python/monki/simulation.py:73-185, after the
accumulate routine of Victor's main.f90:1353-1478):

* radiance: the mean over rounds of each round's Stokes tally per photon per column;
* its standard error: the standard error of that mean over rounds, sqrt(s^2 / n) with the
  unbiased sample variance s^2, reported as 999 (unavailable) through round 30;
* the radiance convergence history: after each round, the running mean averaged over pixels
  and views, per Stokes component;
* DOP = |(Q, U, V)| / I (0 where I <= 1e-16);
* layer AMFs = summed weighted layer paths / summed detector weights / layer thickness (km),
  vacuum layer dropped; with AMF stopping, their delta-method standard errors.

Backward transport reverses U in the product (simulation.py:128-131). Here the per-round
moments are updated online (Welford): mean and centred sum of squares, which avoids the
cancellation of sum-of-squares forms over thousands of rounds.
"""

import json
from dataclasses import asdict
from pathlib import Path
from typing import Optional

import numpy as np

from .convergence import AMFConvergence, PrecisionStopping
from .execution.backend import Backend
from .execution.tallies import BatchControls, RawTallies
from .metadata import file_hash
from .output import SimulationResult

# Standard errors are reported only after this many rounds; earlier ones carry the sentinel.
UNCERTAIN_THROUGH = 30
UNAVAILABLE = 999.0
DOP_FLOOR = 1e-16


def _product_order(array: np.ndarray) -> np.ndarray:
    """Tallies in Fortran axis order (component, x, y, ..., view) to the product's reversed
    axes with a trailing wavelength axis of one."""

    return np.ascontiguousarray(np.transpose(array))[..., np.newaxis]


class RoundStatistics:
    """Accumulates one case's rounds in float64 and writes its product and receipt.

    ``add`` takes whole rounds in order (1, 2, ...); ``add_stokes`` + ``attach`` are the same
    for backends that keep the layer paths on the device. ``simulate`` drives one instance,
    ``simulate_batch`` one per scene in lockstep.
    """

    def __init__(self, case, metadata, output, convergence: Optional[AMFConvergence] = None):
        if convergence is not None and case.rounds < convergence.minimum_rounds:
            raise ValueError(f"round cap {case.rounds} is below the policy's minimum of {convergence.minimum_rounds}")

        self.output = Path(output).resolve()
        self.receipt = self.output.with_suffix(".run.json")

        for path in (self.output, self.receipt):
            if path.exists():
                raise FileExistsError(f"{path} exists; results are never overwritten")

        if metadata is None:
            raise RuntimeError("the backend was prepared without product metadata")

        self.case = case
        self.metadata = metadata
        self.rounds = 0
        self.mean = self.squares = None  # per-photon Stokes per round: running mean, centred squares
        self.history = []
        self.paths = self.weights = None
        self.scattering = None
        self.stopping = None

        if convergence is not None:
            if "AMF_layer" not in metadata.values:
                raise ValueError("the AMF precision needs layer-path tallies: run with save_path=True")

            self.stopping = PrecisionStopping(convergence, np.diff(metadata.values["alt"]) / 1000.0)

    @property
    def completed(self) -> int:
        return self.rounds

    def add(self, batch) -> None:
        """The next round's tallies (``RawTallies``)."""

        if not isinstance(batch, RawTallies):
            raise TypeError(f"backends return RawTallies, got {type(batch).__name__}")

        first = self.rounds == 0
        self.add_stokes(batch.controls, batch.stokes)

        if first:
            if batch.weighted_paths is not None:
                self.paths = np.zeros_like(batch.weighted_paths, dtype=np.float64)
                self.weights = np.zeros_like(batch.weights, dtype=np.float64)

            if batch.scattering is not None:
                self.scattering = {name: np.zeros_like(part, dtype=np.float64) for name, part in batch.scattering.items()}

        if self.scattering is not None:
            if batch.scattering is None:
                raise ValueError(f"round {self.rounds} lacks the scattering decomposition of earlier rounds")

            for name, part in batch.scattering.items():
                self.scattering[name] += part

        if self.paths is not None:
            if batch.weighted_paths is None or batch.weights is None:
                raise ValueError(f"round {self.rounds} lacks the layer-path tallies of earlier rounds")

            self.paths += batch.weighted_paths
            self.weights += batch.weights

    def add_stokes(self, controls, stokes) -> None:
        """The next round's Stokes tally (nmat, nx, ny, views)."""

        expected = BatchControls(self.rounds + 1, self.case.photons_per_column)

        if controls != expected:
            raise ValueError(f"expected tallies of {expected}, got {controls}")

        sample = np.asarray(stokes, dtype=np.float64) / self.case.photons_per_column

        if self.mean is None:
            self.mean = np.zeros_like(sample)
            self.squares = np.zeros_like(sample)

        self.rounds += 1
        step = sample - self.mean
        self.mean = self.mean + step / self.rounds
        self.squares = self.squares + step * (sample - self.mean)
        # Domain average of the running mean per Stokes component.
        self.history.append(self.mean.reshape(self.mean.shape[0], -1).mean(axis=1))

    def attach(self, moments) -> None:
        """Layer-path sums and AMF moments kept on the device (``HIPBatch.moments``) for the
        rounds added so far."""

        if moments.rounds != self.rounds:
            raise ValueError(f"device moments cover {moments.rounds} rounds, the host {self.rounds}")

        self.paths, self.weights = moments.paths, moments.weights

        if self.stopping is not None:
            self.stopping.moments = moments.amf_moments()

    # -- product ----------------------------------------------------------------------------

    def _standard_error(self) -> np.ndarray:
        if self.rounds <= UNCERTAIN_THROUGH:
            return np.full_like(self.mean, UNAVAILABLE)

        if np.any(self.squares < 0.0):  # a centred sum of squares only falls below zero by a bug
            raise ArithmeticError("negative centred sum of squares in the round statistics")

        return np.sqrt(self.squares / ((self.rounds - 1) * self.rounds))

    def _radiance_fields(self) -> dict:
        radiance, error = self.mean.copy(), self._standard_error()
        history = np.array(self.history)

        if self.case.backward and radiance.shape[0] >= 3:
            radiance[2] = -radiance[2]
            history[:, 2] = -history[:, 2]

        polarized = np.linalg.norm(radiance[1:], axis=0)
        dop = np.zeros_like(radiance[0])
        np.divide(polarized, radiance[0], out=dop, where=radiance[0] > DOP_FLOOR)
        return {
            "radiance": _product_order(radiance),
            "std_radiance": _product_order(error),
            "DOP": _product_order(dop),
            "radiance_convergence": history[..., np.newaxis],
        }

    def _amf_fields(self, altitudes) -> dict:
        if self.paths is None:
            return {}

        if not np.all(self.weights > 0):
            raise ValueError("a pixel and view without detector weight has no layer AMF")

        # Path lengths are in km and altitudes in m; the last (vacuum) layer is not reported.
        dz_km = np.diff(altitudes) / 1000.0
        amf = self.paths[:, :, :-1, :] / self.weights[:, :, np.newaxis, :] / dz_km[:, np.newaxis]
        return {"AMF_layer": _product_order(amf)}

    def _decomposition_fields(self) -> dict:
        if self.scattering is None:
            return {}

        per_photon = 1.0 / (self.rounds * self.case.photons_per_column)  # main.f90:1168-1171
        parts = {name: value * per_photon for name, value in self.scattering.items()}

        if self.case.backward and parts["stokes_sing"].shape[0] >= 3:
            for part in parts.values():
                part[2] = -part[2]  # main.f90:1190-1198

        return {
            "radiance_sing": _product_order(parts["stokes_sing"]),
            "radiance_mult": _product_order(parts["stokes_mult"]),
            "contrib_radiance_3D": _product_order(parts["contrib_sing"] + parts["contrib_mult"]),
            "contrib_radiance_3D_sing": _product_order(parts["contrib_sing"]),
            "contrib_radiance_3D_mult": _product_order(parts["contrib_mult"]),
        }

    def finish(self, backend_provenance: dict) -> SimulationResult:
        """Write the NetCDF product and the ``.run.json`` receipt; return the product."""

        if self.rounds == 0:
            raise ValueError("no round was run: nothing to report")

        values = {name: np.array(value, copy=True) for name, value in self.metadata.values.items()}
        dimensions = dict(self.metadata.dimensions)
        attributes = dict(self.metadata.attributes)
        values.update(self._radiance_fields())
        values.update(self._amf_fields(values["alt"]))
        values.update(self._decomposition_fields())
        values["n_rounds"] = np.full_like(values["n_rounds"], self.rounds)
        values["n_photons_per_column"] = np.full_like(values["n_photons_per_column"], self.case.photons_per_column)
        stopping = self.stopping

        if stopping is not None:
            values["std_AMF_layer"] = _product_order(stopping.moments.estimate(stopping.thickness)[1])
            dimensions["std_AMF_layer"] = dimensions["AMF_layer"]
            attributes["std_AMF_layer"] = {"long_name": "Layer AMF standard error", "units": "-"}

        result = SimulationResult(self.output, values, dimensions, attributes)
        result.save()
        _, nx, ny, _ = self.mean.shape
        receipt = dict(backend_provenance)
        receipt.update(
            status="complete", completed_rounds=self.rounds,
            total_photons=self.rounds * self.case.photons_per_column * nx * ny, output_sha256=file_hash(self.output),
        )  # fmt: skip

        if stopping is not None:
            receipt["status"] = "converged" if stopping.converged else "max_rounds_reached"
            receipt["amf_convergence"] = {
                "policy": asdict(stopping.policy), "converged": stopping.converged, "checks": stopping.history,
                "uncertainty": "pointwise relative standard error",
            }  # fmt: skip

        with self.receipt.open("x") as stream:
            json.dump(receipt, stream, indent=2, default=str)
            stream.write("\n")

        return result


def _check_request(case, output: Path, convergence: Optional[AMFConvergence]) -> None:
    """What can be refused before a backend touches the device."""

    if convergence is not None and case.rounds < convergence.minimum_rounds:
        raise ValueError(f"round cap {case.rounds} is below the policy's minimum of {convergence.minimum_rounds}")

    for path in (output, output.with_suffix(".run.json")):
        if path.exists():
            raise FileExistsError(f"{path} exists; results are never overwritten")


def simulate(
    case,
    backend: Backend,
    output,
    *,
    convergence: Optional[AMFConvergence] = None,
) -> SimulationResult:
    """Trace ``case.rounds`` equal rounds on ``backend`` and write the NetCDF product and its
    ``.run.json`` receipt (the reference's ``simulate``, simulation.py:17-189).

    With ``convergence``, ``case.rounds`` is the round cap and the run stops earlier once the
    policy is met; the receipt's status says ``converged`` or ``max_rounds_reached``. The
    backend is prepared here and closed on return or failure.
    """

    output = Path(output).resolve()
    _check_request(case, output, convergence)

    with backend:
        backend.prepare(case)
        statistics = RoundStatistics(case, backend.metadata, output, convergence)

        if statistics.stopping is None:
            rounds = backend.run_batches([BatchControls(index, case.photons_per_column)
                                          for index in range(1, case.rounds + 1)])  # fmt: skip

            if len(rounds) != case.rounds:
                raise ValueError(f"backend returned {len(rounds)} rounds for {case.rounds} requested")
        else:
            rounds = statistics.stopping.schedule(backend, case)

        for batch in rounds:
            statistics.add(batch)

        return statistics.finish(backend.provenance)


def simulate_batch(cases, backend, outputs, *, convergence: Optional[AMFConvergence] = None) -> list:
    """Run the cases of a batch backend (``HIPBatch``) round by round in lockstep and save one
    product and receipt per case. No reference counterpart: the reference runs one case per
    session. With ``AMFConvergence`` each case stops on its own policy; rounds continue while any
    case still runs (finished cases' later rounds are traced and dropped)."""

    cases = list(cases)
    outputs = [Path(output).resolve() for output in outputs]

    if len(outputs) != len(cases):
        raise ValueError("one output path per case is required")

    with backend:
        backend.prepare(cases)
        statistics = [
            RoundStatistics(case, metadata, output, convergence)
            for case, metadata, output in zip(cases, backend.metadata, outputs)
        ]
        running = [True] * len(cases)
        # Backends that keep the layer paths on the device return only Stokes per round; the
        # sums and AMF moments are read at each precision check and at the end.
        on_device = hasattr(backend, "accumulate_round")
        paths = ["AMF_layer" in metadata.values for metadata in backend.metadata]

        for index in range(1, max(case.rounds for case in cases) + 1):
            if not any(running):
                break

            if on_device:
                rounds = backend.accumulate_round(index, active=running)
            else:
                rounds = backend.run_round(index)

            for number, (stats, batch) in enumerate(zip(statistics, rounds)):
                if not running[number]:
                    continue

                stopping = stats.stopping

                if on_device:
                    stats.add_stokes(BatchControls(index, stats.case.photons_per_column), batch)
                else:
                    stats.add(batch)

                    if stopping is not None:
                        stopping.moments.add(batch)

                if stopping is not None:
                    group_end = index % stopping.policy.check_every == 0 or index == stats.case.rounds

                    if group_end and on_device and index >= stopping.policy.minimum_rounds:
                        stats.attach(backend.moments(number))

                    if group_end and stopping.check(index):
                        running[number] = False

                if index >= stats.case.rounds:
                    running[number] = False

        if on_device:
            for number, stats in enumerate(statistics):
                if paths[number]:
                    stats.attach(backend.moments(number))

        return [stats.finish(provenance) for stats, provenance in zip(statistics, backend.case_provenance())]
