"""Shared round statistics and the NetCDF product, independent of the transport backend.

``simulate`` restates ``python/monki/simulation.py:17-189`` (itself following Victor's
``main.f90`` accumulate routine): float64 accumulation of per-round tallies, the standard
error of the mean over rounds with the 999 sentinel through round 30, the radiance
convergence history, DOP, layer AMFs (vacuum layer dropped) and the ``.run.json`` receipt.
"""

import json
from dataclasses import asdict
from pathlib import Path
from typing import Optional

import numpy as np

from .convergence import AMFConvergence, AMFStopping
from .execution.backend import Backend
from .execution.tallies import BatchControls, RawTallies
from .metadata import file_hash
from .output import SimulationResult


class RoundStatistics:
    """Round accumulation and product of one case (the body of simulation.py:45-189).

    ``add`` takes the rounds in order; ``finish`` computes the statistics, writes the NetCDF
    product and the ``.run.json`` receipt. ``simulate`` drives one instance from one backend;
    ``simulate_batch`` drives one per scene of a batch in lockstep.
    """

    def __init__(self, case, metadata, output, convergence: Optional[AMFConvergence] = None):
        if convergence is not None and case.rounds < convergence.minimum_rounds:
            raise ValueError("round cap must be at least minimum_rounds")

        self.output = Path(output).resolve()
        self.receipt = self.output.with_suffix(".run.json")

        if self.output.exists() or self.receipt.exists():
            raise FileExistsError(f"Refusing to overwrite an existing result: {self.output}")

        if metadata is None:
            raise RuntimeError("Backend did not supply prepared scene metadata")

        self.case = case
        self.metadata = metadata
        self.total = self.rounds_sum = self.rounds_squared = self.paths = self.weights = None
        self.scattering = None  # sums of the decomposition arrays, when the backend returns them
        self.history = []
        self.stopping = None

        if convergence is not None:
            if "AMF_layer" not in metadata.values:
                raise ValueError("AMF stopping requires layer path diagnostics")

            thickness = np.diff(metadata.values["alt"]) / 1000
            self.stopping = AMFStopping(convergence, thickness)

    @property
    def completed(self) -> int:
        return len(self.history)

    def add(self, batch) -> None:
        """Accumulate the next round (round indices 1, 2, ... in order)."""

        if not isinstance(batch, RawTallies):
            raise TypeError("Backend must return RawTallies")

        if self.total is None and batch.weighted_paths is not None:
            self.paths = np.zeros_like(batch.weighted_paths)
            self.weights = np.zeros_like(batch.weights)

        if self.total is None and batch.scattering is not None:
            self.scattering = {name: np.zeros_like(value) for name, value in batch.scattering.items()}

        self.add_stokes(batch.controls, batch.stokes)

        if self.scattering is not None:
            if batch.scattering is None:
                raise ValueError("Backend dropped the scattering decomposition during a session")

            for name, value in batch.scattering.items():
                self.scattering[name] += value

        if self.paths is not None:
            if batch.weighted_paths is None or batch.weights is None:
                raise ValueError("Backend dropped path diagnostics during a session")

            self.paths += batch.weighted_paths
            self.weights += batch.weights

    def add_stokes(self, controls, stokes) -> None:
        """Accumulate the Stokes tallies of the next round; the layer paths of device-accumulated
        rounds arrive through ``attach``."""

        case = self.case
        index = self.completed + 1

        if controls != BatchControls(index, case.photons_per_column):
            raise ValueError("Backend returned tallies for a different round")

        if self.total is None:
            self.total = np.zeros_like(stokes)
            self.rounds_sum = np.zeros_like(stokes)
            self.rounds_squared = np.zeros_like(stokes)

        self.total += stokes
        normalized = stokes / case.photons_per_column
        self.rounds_sum += normalized
        self.rounds_squared += normalized**2
        _, nx, ny, angles = self.total.shape
        # Successive spatial sums and divisions as in the original accumulate.
        self.history.append(self.rounds_sum.sum(axis=1).sum(axis=1).sum(axis=1) / nx / ny / angles / index)

    def attach(self, moments) -> None:
        """Take the layer-path sums and AMF moments a batch backend accumulated on the device
        (``HIPBatch.moments``) for the rounds added so far."""

        if moments.rounds != self.completed:
            raise ValueError("Device moments cover a different number of rounds")

        self.paths, self.weights = moments.paths, moments.weights

        if self.stopping is not None:
            self.stopping.moments = moments.amf_moments()

    def finish(self, backend_provenance: dict) -> SimulationResult:
        case, metadata, output = self.case, self.metadata, self.output
        total, rounds_sum, rounds_squared = self.total, self.rounds_sum, self.rounds_squared
        paths, weights, stopping = self.paths, self.weights, self.stopping
        completed = self.completed

        if completed == 0:
            raise ValueError("No statistical round was run")

        _, nx, ny, _ = total.shape
        radiance = total / completed / case.photons_per_column
        uncertainty = np.full_like(radiance, 999.0)

        # 999 marks an unavailable uncertainty through round 30 (simulation.py:113-124).
        if completed > 30:
            variance_sum = rounds_squared - (1.0 / completed) * rounds_sum**2
            tolerance = 64 * np.finfo(float).eps * np.maximum(rounds_squared, 1e-300)

            if np.any(variance_sum < -tolerance):
                raise ArithmeticError("Negative round variance exceeds floating-point roundoff")

            uncertainty = np.sqrt(np.maximum(variance_sum, 0) / (completed - 1)) / np.sqrt(
                completed
            )

        radiance_history = np.asarray(self.history)

        if case.backward and radiance.shape[0] >= 3:
            radiance[2] *= -1
            radiance_history[:, 2] *= -1

        intensity = radiance[0]
        polarization = np.sqrt(np.sum(radiance[1:] ** 2, axis=0))
        dop = np.divide(
            polarization, intensity, out=np.zeros_like(intensity), where=intensity > 1e-16
        )
        values = {name: value.copy() for name, value in metadata.values.items()}
        values["radiance"] = radiance.transpose(3, 2, 1, 0)[..., None]
        values["std_radiance"] = uncertainty.transpose(3, 2, 1, 0)[..., None]
        values["DOP"] = dop.transpose(2, 1, 0)[..., None]
        values["radiance_convergence"] = radiance_history[..., None]
        values["n_rounds"] = np.full_like(values["n_rounds"], completed)
        values["n_photons_per_column"] = np.full_like(
            values["n_photons_per_column"], case.photons_per_column
        )

        if paths is not None:
            if np.any(weights <= 0):
                raise ValueError("Layer AMF is undefined for a pixel with no detected weight")

            # Lengths are km, altitude metadata is m; the vacuum layer is dropped.
            thickness = np.diff(values["alt"]) / 1000.0
            amf = paths[:, :, :-1, :] / weights[:, :, None, :] / thickness[None, None, :, None]
            values["AMF_layer"] = amf.transpose(3, 2, 1, 0)[..., None]

        if self.scattering is not None:
            # Sums over rounds / rounds / photons per column (main.f90:1168-1171), U reversed
            # for backward transport like the radiance; the 3-D total is single + multiple.
            scale = 1.0 / completed / case.photons_per_column
            sing, mult = self.scattering["stokes_sing"] * scale, self.scattering["stokes_mult"] * scale
            contrib_sing = self.scattering["contrib_sing"] * scale
            contrib_mult = self.scattering["contrib_mult"] * scale
            if case.backward and sing.shape[0] >= 3:
                for part in (sing, mult, contrib_sing, contrib_mult):
                    part[2] *= -1  # main.f90:1190-1198

            values["radiance_sing"] = sing.transpose(3, 2, 1, 0)[..., None]
            values["radiance_mult"] = mult.transpose(3, 2, 1, 0)[..., None]
            values["contrib_radiance_3D"] = (contrib_sing + contrib_mult).transpose(4, 3, 2, 1, 0)[..., None]
            values["contrib_radiance_3D_sing"] = contrib_sing.transpose(4, 3, 2, 1, 0)[..., None]
            values["contrib_radiance_3D_mult"] = contrib_mult.transpose(4, 3, 2, 1, 0)[..., None]

        dimensions = dict(metadata.dimensions)
        attributes = dict(metadata.attributes)

        if stopping is not None:
            _, amf_error = stopping.moments.estimate(stopping.thickness)
            values["std_AMF_layer"] = amf_error.transpose(3, 2, 1, 0)[..., None]
            dimensions["std_AMF_layer"] = dimensions["AMF_layer"]
            attributes["std_AMF_layer"] = {"long_name": "Layer AMF standard error", "units": "-"}

        result = SimulationResult(output, values, dimensions, attributes)
        result.save()
        provenance = {
            **backend_provenance,
            "status": "complete",
            "completed_rounds": completed,
            "total_photons": completed * case.photons_per_column * nx * ny,
            "output_sha256": file_hash(output),
        }

        if stopping is not None:
            provenance["status"] = "converged" if stopping.converged else "max_rounds_reached"
            provenance["amf_convergence"] = {
                "policy": asdict(stopping.policy),
                "converged": stopping.converged,
                "checks": stopping.history,
                "uncertainty": "pointwise relative standard error",
            }

        with self.receipt.open("x") as stream:
            stream.write(json.dumps(provenance, indent=2, default=str) + "\n")

        return result


def simulate(
    case,
    backend: Backend,
    output,
    *,
    convergence: Optional[AMFConvergence] = None,
) -> SimulationResult:
    """Run equally sized rounds and save a complete NetCDF product plus receipt.

    With ``AMFConvergence``, ``case.rounds`` is the maximum round budget; the receipt
    records ``converged`` or ``max_rounds_reached``.
    """

    if convergence is not None and case.rounds < convergence.minimum_rounds:
        raise ValueError("round cap must be at least minimum_rounds")

    output = Path(output).resolve()

    if output.exists() or output.with_suffix(".run.json").exists():
        raise FileExistsError(f"Refusing to overwrite an existing result: {output}")

    with backend:
        backend.prepare(case)
        statistics = RoundStatistics(case, backend.metadata, output, convergence)

        if statistics.stopping is not None:
            batches = statistics.stopping.batches(backend, case)
        else:
            requests = [BatchControls(index, case.photons_per_column) for index in range(1, case.rounds + 1)]
            batches = backend.run_batches(requests)

            if len(batches) != case.rounds:
                raise ValueError("Backend returned the wrong number of statistical rounds")

        for batch in batches:
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
