"""Many scenes in one persistent transport launch per round (batched HIP backend).

No reference counterpart: the reference prepares and runs one scene per session
(``python/monki/execution/metal.py:81-213``). ``HIPBatch`` packs every scene with the same
table layout as ``HIP`` (``hip_tables.pack_tables``), uploads them into one session, and traces a
round of all scenes with one launch: waves start spread over the scenes and move to the next
scene when theirs runs out of photons. The round (clears, transport, per-scene float64 shard
reductions) is captured once as a HIP graph and replayed; only the per-scene control words
change between rounds. ``simulate_batch`` consumes it.
"""

import ctypes
import hashlib
import platform
from dataclasses import dataclass, replace
from pathlib import Path
from typing import List, Optional

import numpy as np

from ..input import ReferenceCase, SceneCase
from ..metadata import file_hash, scene_metadata
from ..preparation import UPSTREAM, prepare_case
from .hip import DOUBLE_POINTER, FLOAT_POINTER, HIP, LIBRARY, UINT_POINTER, all_finite, load_library
from .hip_tables import pack_tables, tally_count
from .proposal import DirectedSettings, Proposal, ProposalSettings, apply_steering, sharpest_species
from .tallies import BatchControls, RawTallies


def declare(native) -> None:
    native.monki_hip_batch_create.argtypes = [
        ctypes.c_uint32, FLOAT_POINTER, ctypes.POINTER(ctypes.c_size_t), ctypes.c_uint32, ctypes.c_uint32,
        ctypes.c_char_p, ctypes.c_size_t,
    ]  # fmt: skip
    native.monki_hip_batch_create.restype = ctypes.c_void_p
    native.monki_hip_batch_run.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, UINT_POINTER, DOUBLE_POINTER,
        ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
    ]  # fmt: skip
    native.monki_hip_batch_run.restype = ctypes.c_int
    native.monki_hip_batch_accumulate.argtypes = native.monki_hip_batch_run.argtypes
    native.monki_hip_batch_accumulate.restype = ctypes.c_int
    native.monki_hip_batch_moments.argtypes = [
        ctypes.c_void_p, ctypes.c_uint32, DOUBLE_POINTER, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t,
    ]  # fmt: skip
    native.monki_hip_batch_moments.restype = ctypes.c_int


@dataclass
class DeviceMoments:
    """Sums and online AMF moments of one scene's rounds, accumulated on the device.

    ``rounds`` rounds; ``paths`` (nx, ny, nz, views) and ``weights`` (nx, ny, views) are the
    sums of the rounds' outputs; the moments are the online means and centred (co)moments of RatioMoments
    (physical layers only; weight moments per pixel and view).
    """

    rounds: int
    paths: np.ndarray
    weights: np.ndarray
    path_mean: np.ndarray
    path_m2: np.ndarray
    covariance: np.ndarray
    weight_mean: np.ndarray
    weight_m2: np.ndarray

    def amf_moments(self):
        from ..convergence import RatioMoments  # convergence imports this package

        return RatioMoments.from_moments(self.rounds, self.path_mean, self.weight_mean, self.path_m2, self.weight_m2,
                                         self.covariance)  # fmt: skip


class HIPBatch:
    """A batch of ``SceneCase`` traced together, one round at a time, on one GPU.

    Settings: ``tally_shards`` (per scene), ``block`` (lanes per block: 64, 128, 256) and
    ``use_graph`` (replay the captured round) -- the parameters ``autotune`` chooses.
    """

    def __init__(
        self,
        device: int = 0,
        *,
        save_path: bool = True,
        seed: int = 23274,
        max_steps: int = 100_000,
        tally_shards: int = 32,
        block: int = 256,
        use_graph: bool = True,
        library: Path = LIBRARY,
        upstream: Path = UPSTREAM,
        proposal: Optional[ProposalSettings] = None,
        directed: Optional[DirectedSettings] = None,
    ):
        for name, value, minimum, maximum in (
            ("device", device, 0, 2**32 - 1),
            ("seed", seed, 0, 2**32 - 1),
            ("max_steps", max_steps, 1, 100_000),
            ("tally_shards", tally_shards, 1, 64),
        ):
            if type(value) is not int or not minimum <= value <= maximum:
                raise ValueError(f"{name} must be an integer in [{minimum}, {maximum}]")

        if block not in (64, 128, 256):
            raise ValueError("block must be 64, 128 or 256")

        if proposal is not None and not isinstance(proposal, ProposalSettings):
            raise ValueError("proposal must be ProposalSettings or None")

        if directed is not None and (not isinstance(directed, DirectedSettings) or proposal is not None):
            raise ValueError("directed must be DirectedSettings or None, and excludes a proposal")

        self.device = device
        self.save_path = save_path
        self.seed = seed
        self.max_steps = max_steps
        self.tally_shards = tally_shards
        self.block = block
        self.use_graph = bool(use_graph)
        self.library = Path(library)
        self.upstream = Path(upstream)
        self.metadata: List = []
        self.provenance: dict = {}
        self.proposal = proposal
        self.directed = directed
        self._network: Optional[Proposal] = None
        self._native = None
        self._handle = None
        self._cases: List[SceneCase] = []
        self._preparations: List = []
        self._shapes = []
        self._counts = []
        self._round = 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def prepare(self, cases) -> None:
        self.close()
        cases = list(cases)

        if not cases or not all(isinstance(case, (ReferenceCase, SceneCase)) for case in cases):
            raise TypeError("HIPBatch.prepare expects a non-empty list of ReferenceCase or SceneCase")

        if any(case.backward for case in cases):
            raise NotImplementedError("HIPBatch runs forward transport; run backward cases with HIP")

        extras, preparations = [None] * len(cases), [None] * len(cases)

        for index, case in enumerate(cases):
            if isinstance(case, ReferenceCase):
                cases[index], prepared, preparations[index] = prepare_case(case, self.upstream, save_path=self.save_path)
                extras[index] = prepared.product_variables()

        packed, shapes, counts, offsets = [], [], [], [0]

        for case in cases:
            tables = replace(case.scene, save_path=self.save_path)
            buffer, shape = pack_tables(tables, shards=self.tally_shards)
            packed.append(buffer)
            shapes.append(shape)
            nmat, nx, ny, nz, views = shape

            if nmat == 4 or case.scene.save_scats or case.scene.detector != "toa" or case.scene.register not in (None, "toa"):
                raise NotImplementedError("HIPBatch takes nmat 1 or 3, TOA detectors, no scattering decomposition; use HIP")

            counts.append(tally_count(nmat, nx, ny, nz, views, self.save_path))
            offsets.append(offsets[-1] + buffer.size)

            if case.photons_per_column * nx * ny >= 2**31:
                raise ValueError("HIP photon identity exceeds its 32-bit round range")

        tables = np.concatenate(packed).astype(np.float32)
        offset_array = (ctypes.c_size_t * len(offsets))(*offsets)
        native = load_library(self.library)
        declare(native)
        error = ctypes.create_string_buffer(8192)
        handle = native.monki_hip_batch_create(
            self.device, tables.ctypes.data_as(FLOAT_POINTER), offset_array, len(cases), self.block, error,
            len(error),
        )  # fmt: skip

        if not handle:
            raise RuntimeError(error.value.decode() or "HIP batch session creation failed")

        self._native, self._handle = native, handle
        self._cases, self._shapes, self._counts = cases, shapes, counts
        self._round = 0
        self.metadata = [
            scene_metadata(replace(case.scene, save_path=self.save_path), case.photons_per_column, extra=extra)
            for case, extra in zip(cases, extras)
        ]
        self._preparations = preparations
        self.provenance = {
            "backend": "HIPBatch",
            "device": native.monki_hip_device_name(handle).decode(),
            "scenes": len(cases),
            "rng": "Philox4x32-10; seed/round/photon/draw-block",
            "seed": self.seed,
            "transport_precision": "float32",
            "round_precision": "float64",
            "tally_shards": self.tally_shards,
            "block": self.block,
            "graph": self.use_graph,
            "max_steps": self.max_steps,
            "library_sha256": file_hash(self.library),
            "tables_sha256": hashlib.sha256(tables.tobytes()).hexdigest(),
            "python_version": platform.python_version(),
            "gpu_rounds": 0,
        }

        try:
            self._steer(cases)
        except BaseException:
            self.close()
            raise

    def _steer(self, cases) -> None:
        """Guided or detector-directed rounds for the whole batch: one network, trained on the
        first case by the single-scene backend, or the view lobes; the lobes take the sharpest
        tabulated phase function of the first case (it must exist in every scene)."""

        native, handle = self._native, self._handle
        settings = self.proposal or self.directed

        if settings is None:
            return

        lobes = sharpest_species(cases[0].scene)
        error = ctypes.create_string_buffer(4096)

        if self.proposal is not None:
            trainer = HIP(self.device, proposal=self.proposal, seed=self.seed, max_steps=self.max_steps,
                          library=self.library)  # fmt: skip

            with trainer:
                trainer.prepare(cases[0])
                self.provenance["proposal"] = dict(trainer.provenance["proposal"], trained_on=cases[0].name)
                self._network = trainer.release_network()

            status = native.monki_hip_attach_proposal(
                handle, self._network.handle, settings.alpha, settings.species_mask(), min(settings.max_order, 2**32 - 1),
                error, len(error),
            )  # fmt: skip
            fraction = settings.lobe_fraction if lobes >= 0 else 0.0
            # Without a tabulated species the guided tracer has no view lobes, at the surface too.
            settings = replace(settings, surface=settings.surface and lobes >= 0)
            self.provenance["method"] = "guided (view lobes + learned vMF mixture, defensive physical sampler)"
        else:
            status = native.monki_hip_set_directed(
                handle, settings.alpha, settings.kappa, settings.species_mask() & 0xFFFFFFFF,
                min(settings.max_order, 2**32 - 1), error, len(error),
            )  # fmt: skip
            lobes = lobes if settings.lobes == "phase" else -1
            fraction = 1.0
            self.provenance["directed"] = {key: getattr(settings, key) for key in settings.__dataclass_fields__}
            self.provenance["method"] = "detector-directed (view lobes, defensive physical sampler)"

        if status:
            raise RuntimeError(error.value.decode())

        apply_steering(native, handle, settings, lobes, fraction)
        self.provenance["lobe_species"] = lobes

    def case_provenance(self) -> list:
        records = []

        for case, preparation in zip(self._cases, self._preparations):
            record = {**self.provenance, "case": case.name, "scene": case.scene.description}

            if preparation is not None:
                record["preparation"] = preparation

            records.append(record)

        return records

    def run_round(self, round_index: int) -> List[RawTallies]:
        """One statistical round of every scene (round indices sequential from one)."""

        if self._handle is None:
            raise RuntimeError("HIPBatch must be prepared before running a round")

        if round_index != self._round + 1:
            raise ValueError("round_index must be sequential, starting at one")

        photons = [case.photons_per_column * shape[1] * shape[2] for case, shape in zip(self._cases, self._shapes)]
        words = (ctypes.c_uint32 * len(photons))(*photons)
        out = np.zeros(sum(self._counts), dtype=np.float64)
        error = ctypes.create_string_buffer(8192)

        try:
            status = self._native.monki_hip_batch_run(
                self._handle, round_index, self.seed, self.max_steps, words, out.ctypes.data_as(DOUBLE_POINTER),
                int(self.use_graph), error, len(error),
            )  # fmt: skip

            if status or not all_finite(out):
                raise RuntimeError(error.value.decode() or "Nonfinite HIP batch tallies")
        except BaseException:
            self.close()
            raise

        self._round = round_index
        self.provenance["gpu_rounds"] += 1
        # Every array returned is a view of this round's own buffer (no copies).
        results, start = [], 0

        for case, shape, count in zip(self._cases, self._shapes, self._counts):
            values = out[start : start + count]
            start += count
            nmat, nx, ny, nz, views = shape
            stokes_count = nmat * nx * ny * views
            stokes = values[:stokes_count].reshape((nmat, nx, ny, views), order="F")
            paths = weights = None

            if self.save_path:
                path_count = nx * ny * nz * views
                paths = values[stokes_count : stokes_count + path_count].reshape((nx, ny, nz, views), order="F")
                weights = values[stokes_count + path_count :].reshape((nx, ny, views), order="F")

            results.append(RawTallies(BatchControls(round_index, case.photons_per_column), stokes, paths, weights))

        return results

    def accumulate_round(self, round_index: int, active=None) -> List[Optional[np.ndarray]]:
        """One round of every active scene with the path and weight tallies kept on the device.

        Returns each scene's Stokes tallies (nmat, nx, ny, views) for this round, ``None`` for
        a scene that is not active (``active[i]`` false: nothing traced or accumulated). The
        sums and AMF moments are read with ``moments``. Round indices are sequential from one,
        shared with ``run_round``.
        """

        if self._handle is None:
            raise RuntimeError("HIPBatch must be prepared before running a round")

        if round_index != self._round + 1:
            raise ValueError("round_index must be sequential, starting at one")

        active = [True] * len(self._cases) if active is None else [bool(a) for a in active]

        if len(active) != len(self._cases):
            raise ValueError("one active flag per case is required")

        photons = [
            case.photons_per_column * shape[1] * shape[2] if on else 0
            for case, shape, on in zip(self._cases, self._shapes, active)
        ]  # fmt: skip
        words = (ctypes.c_uint32 * len(photons))(*photons)
        sizes = [shape[0] * shape[1] * shape[2] * shape[4] for shape in self._shapes]
        stokes = np.zeros(sum(sizes), dtype=np.float64)
        error = ctypes.create_string_buffer(8192)

        try:
            status = self._native.monki_hip_batch_accumulate(
                self._handle, round_index, self.seed, self.max_steps, words, stokes.ctypes.data_as(DOUBLE_POINTER),
                int(self.use_graph), error, len(error),
            )  # fmt: skip

            if status or not all_finite(stokes):
                raise RuntimeError(error.value.decode() or "Nonfinite HIP batch tallies")
        except BaseException:
            self.close()
            raise

        self._round = round_index
        self.provenance["gpu_rounds"] += 1
        # Views of this round's own buffer (no copies).
        results, start = [], 0

        for shape, size, on in zip(self._shapes, sizes, active):
            values = stokes[start : start + size]
            start += size
            nmat, nx, ny, _, views = shape
            results.append(values.reshape((nmat, nx, ny, views), order="F") if on else None)

        return results

    def moments(self, number: int) -> DeviceMoments:
        """The device sums and AMF moments of case ``number`` (layer paths required)."""

        if self._handle is None:
            raise RuntimeError("HIPBatch must be prepared before reading moments")

        if not self.save_path:
            raise ValueError("device moments need layer path diagnostics (save_path)")

        _, nx, ny, nz, views = self._shapes[number]
        p, w, layers = nx * ny * nz * views, nx * ny * views, nx * ny * (nz - 1) * views
        out = np.zeros(1 + p + 3 * w + 3 * layers, dtype=np.float64)
        error = ctypes.create_string_buffer(8192)

        if self._native.monki_hip_batch_moments(
            self._handle, number, out.ctypes.data_as(DOUBLE_POINTER), out.size, error, len(error)
        ):  # fmt: skip
            raise RuntimeError(error.value.decode() or "HIP batch moments failed")

        cuts = np.cumsum([1, p, w, layers, layers, layers, w])
        parts = np.split(out, cuts)
        full, weight_shape, physical = (nx, ny, nz, views), (nx, ny, views), (nx, ny, nz - 1, views)
        return DeviceMoments(
            int(parts[0][0]),
            parts[1].reshape(full, order="F"),
            parts[2].reshape(weight_shape, order="F"),
            parts[3].reshape(physical, order="F"),
            parts[4].reshape(physical, order="F"),
            parts[5].reshape(physical, order="F"),
            parts[6].reshape(weight_shape, order="F"),
            parts[7].reshape(weight_shape, order="F"),
        )

    def kernel_seconds(self) -> float:
        """Device seconds of the last round (HIP events around the launch or graph replay)."""

        return float(self._native.monki_hip_kernel_seconds(self._handle))

    def counters(self) -> dict:
        from .hip import COUNTER_NAMES

        values = (ctypes.c_uint64 * len(COUNTER_NAMES))()

        if self._native.monki_hip_counters(self._handle, values, len(values)):
            raise RuntimeError("HIP counter readback failed")

        return dict(zip(COUNTER_NAMES, (int(v) for v in values)))

    def close(self) -> None:
        if self._handle is not None and self._native is not None:
            self._native.monki_hip_close(self._handle)

        if self._network is not None:
            self._network.close()

        self._network = None

        self._handle = None
        self._native = None
