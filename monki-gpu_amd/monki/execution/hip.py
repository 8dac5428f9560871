"""MI355X transport backend: prepared tables resident in HBM, histories traced on the GPU.

Counterpart of the reference ``Metal`` backend (``python/monki/execution/metal.py:23-349``):
same settings validation before any device access, same session lifecycle (prepare replaces,
close is idempotent, any failure closes the session), same grouping of equal rounds into one
dispatch with independent random streams and tally banks, and the same ``RawTallies``
output. The native side is ``libmonki_hip.so`` (``include/monki_hip.h``), compiled ahead of
time for gfx950. There is no CPU transport fallback: without the library or a GPU,
``prepare`` raises.
"""

import contextlib
import ctypes
import hashlib
import itertools
import platform
import time
from dataclasses import replace
from pathlib import Path
from typing import List, Optional

import numpy as np

from ..input import ReferenceCase, SceneCase
from ..metadata import file_hash, scene_metadata
from ..preparation import UPSTREAM, prepare_case
from .backend import Backend
from .hip_tables import pack_tables, split_scattering, tally_count
from .proposal import (DirectedSettings, Proposal, ProposalSettings, apply_steering, declare, log_rounds,
                       sharpest_species)
from .tallies import BatchControls, RawTallies

PACKAGE_ROOT = Path(__file__).resolve().parents[2]
LIBRARY = PACKAGE_ROOT / "lib" / "libmonki_hip.so"
ABI_VERSION = 2
# Seed offset of the proposal's training rounds: their streams never meet the estimation's.
TRAINING_SEED_OFFSET = 0x6A09E667
TRACE_WORDS = 8
COUNTER_NAMES = (
    "photons",
    "atmospheric_events",
    "surface_events",
    "walk_cells",
    "detector_cells",
    "table_reads",
    "sampling_attempts",
    "tally_atomics",
)

FLOAT_POINTER = ctypes.POINTER(ctypes.c_float)
DOUBLE_POINTER = ctypes.POINTER(ctypes.c_double)
UINT_POINTER = ctypes.POINTER(ctypes.c_uint32)


def all_finite(values: np.ndarray) -> bool:
    """Whether a float64 tally array holds no NaN or infinity, by one dot product (a BLAS pass,
    about a tenth of the time of ``np.isfinite(values).all()`` on a round's millions of values).
    Magnitudes of 1e154 and above also read as not finite; tallies stay many orders below."""

    return bool(np.isfinite(np.dot(values, values)))
U64_POINTER = ctypes.POINTER(ctypes.c_uint64)


def load_library(path: Path = LIBRARY):
    """Load libmonki_hip.so and declare its C-ABI; raise if it is missing or stale."""

    if not Path(path).is_file():
        raise FileNotFoundError(
            f"{path} is missing: build it with python -c 'import __graft_entry__ as g; g.build()'"
        )

    native = ctypes.CDLL(str(path))
    abi = getattr(native, "monki_hip_abi_version", None)

    if abi is None or abi() != ABI_VERSION:
        raise RuntimeError("Rebuild libmonki_hip.so: its ABI version does not match")

    native.monki_hip_create.argtypes = [
        ctypes.c_uint32,
        FLOAT_POINTER,
        ctypes.c_size_t,
        ctypes.c_size_t,
        ctypes.c_uint32,
        ctypes.c_char_p,
        ctypes.c_size_t,
    ]
    native.monki_hip_create.restype = ctypes.c_void_p

    native.monki_hip_run_static.argtypes = [
        ctypes.c_uint32, FLOAT_POINTER, ctypes.c_size_t, UINT_POINTER, DOUBLE_POINTER, ctypes.c_char_p, ctypes.c_size_t,
    ]  # fmt: skip
    native.monki_hip_run_static.restype = ctypes.c_int

    for name in ("monki_hip_run", "monki_hip_run_fp64"):
        function = getattr(native, name)
        function.argtypes = [ctypes.c_void_p, UINT_POINTER, DOUBLE_POINTER, ctypes.c_char_p, ctypes.c_size_t]
        function.restype = ctypes.c_int

    native.monki_hip_trace.argtypes = [
        ctypes.c_void_p,
        UINT_POINTER,
        ctypes.c_uint32,
        DOUBLE_POINTER,
        FLOAT_POINTER,
        ctypes.c_size_t,
        ctypes.c_char_p,
        ctypes.c_size_t,
    ]
    native.monki_hip_trace.restype = ctypes.c_int
    native.monki_hip_counters.argtypes = [ctypes.c_void_p, U64_POINTER, ctypes.c_size_t]
    native.monki_hip_counters.restype = ctypes.c_int
    native.monki_hip_gpu_seconds.argtypes = [ctypes.c_void_p]
    native.monki_hip_gpu_seconds.restype = ctypes.c_double
    native.monki_hip_kernel_seconds.argtypes = [ctypes.c_void_p]
    native.monki_hip_kernel_seconds.restype = ctypes.c_double
    native.monki_hip_layer_rows.argtypes = [ctypes.c_void_p]
    native.monki_hip_layer_rows.restype = ctypes.c_int
    native.monki_hip_device_name.argtypes = [ctypes.c_void_p]
    native.monki_hip_device_name.restype = ctypes.c_char_p
    native.monki_hip_close.argtypes = [ctypes.c_void_p]
    native.monki_hip_close.restype = None
    declare(native)
    return native


class HIP(Backend):
    """Retain device tables across rounds; all photon histories execute on the GPU.

    Transport is float32 (``precision=64`` runs the same kernels in float64 on the same
    random streams); per-round shard reduction is float64 on the device; cross-round
    statistics are float64 on the host. One caller per instance.

    With ``proposal=ProposalSettings(...)`` the backend is the guided tracer: ``prepare`` logs
    scattering events of plain rounds on their own random streams, trains the proposal network
    on the device, and every estimation round then samples directions from the learned
    mixture (with the phase function as defensive component) and reweights.
    """

    def __init__(
        self,
        device: int = 0,
        *,
        save_path: bool = True,
        save_scats: bool = False,
        profile: bool = False,
        seed: int = 23274,
        photons_per_dispatch: int = 1 << 24,
        tally_shards: int = 32,
        max_steps: int = 100_000,
        precision: int = 32,
        library: Path = LIBRARY,
        proposal: Optional[ProposalSettings] = None,
        upstream: Path = UPSTREAM,
        directed: Optional[DirectedSettings] = None,
    ):
        for name, value, minimum, maximum in (
            ("device", device, 0, 2**32 - 1),
            ("seed", seed, 0, 2**32 - 1),
            ("tally_shards", tally_shards, 1, 64),
            ("photons_per_dispatch", photons_per_dispatch, 1, 1 << 30),
            ("max_steps", max_steps, 1, 100_000),
        ):
            if type(value) is not int or not minimum <= value <= maximum:
                raise ValueError(f"{name} must be an integer in [{minimum}, {maximum}]")

        for name, value in (("save_path", save_path), ("save_scats", save_scats), ("profile", profile)):
            if type(value) is not bool:
                raise ValueError(f"{name} must be a bool")

        if save_scats and save_path:
            # calc_radiances_and_scats replaces the path tallies (mc.f90:826-865).
            raise ValueError("save_scats and save_path are exclusive: set save_path=False")

        if precision not in (32, 64) or type(precision) is not int:
            raise ValueError("precision must be 32 or 64")

        if proposal is not None and not isinstance(proposal, ProposalSettings):
            raise ValueError("proposal must be ProposalSettings or None")

        if proposal is not None and precision != 32:
            raise ValueError("the guided tracer runs in float32")

        if directed is not None and not isinstance(directed, DirectedSettings):
            raise ValueError("directed must be DirectedSettings or None")

        if directed is not None and (precision != 32 or proposal is not None):
            raise ValueError("detector-directed sampling runs in float32 without a proposal")

        self.device = device
        self.save_path = save_path
        self.save_scats = save_scats
        self.profile = profile
        self.seed = seed
        self.photons_per_dispatch = photons_per_dispatch
        self.tally_shards = tally_shards
        self.max_steps = max_steps
        self.precision = precision
        self.library = Path(library)
        self.proposal = proposal
        self.directed = directed
        self.upstream = Path(upstream)
        self.metadata = None
        self.provenance: dict = {}
        self._native = None
        self._handle = None
        self._shape = (0, 0, 0, 0, 0)
        self._configuration = None
        self._round = 0
        self._tally_count = 0
        self._round_capacity = 1
        self._packed: Optional[np.ndarray] = None
        self._network: Optional[Proposal] = None

    # -- lifecycle ----------------------------------------------------------------------

    def prepare(self, scene: object) -> None:
        """Pack and upload the scene tables; reserve per-round tally banks on the device."""

        self.close()

        extra, preparation = None, None

        if isinstance(scene, ReferenceCase):
            # The reference prepares fixtures with its Fortran worker (metal.py:98-104); the
            # same tables come from the host restatement here.
            scene, prepared, preparation = prepare_case(scene, self.upstream, save_path=self.save_path,
                                                        save_scats=self.save_scats)  # fmt: skip
            extra = prepared.product_variables()

        if not isinstance(scene, SceneCase):
            raise TypeError("HIP.prepare expects a ReferenceCase or SceneCase")

        tables = replace(scene.scene, save_path=self.save_path, save_scats=self.save_scats)

        if (tables.nmat == 4 or scene.backward or self.save_scats) and (
            self.proposal is not None or self.directed is not None
        ):
            raise NotImplementedError("the guided and detector-directed tracers take forward nmat 1 or 3 scenes")

        if self.save_scats and tables.detector != "toa":
            raise NotImplementedError("the scattering decomposition runs with the detector at TOA")

        if scene.backward and (tables.detector != "toa" or tables.register not in (None, "toa")):
            # main.f90:765-781 allows backward runs only with source, detector and grid at TOA.
            raise NotImplementedError("backward transport runs with the detector and the result grid at TOA")

        packed, self._shape = pack_tables(tables, shards=self.tally_shards, backward=scene.backward)
        nmat, nx, ny, nz, views = self._shape
        per_round = tally_count(nmat, nx, ny, nz, views, self.save_path, self.save_scats)
        self._tally_count = per_round
        bank_bytes = per_round * self.tally_shards * (4 if self.precision == 32 else 8)
        self._round_capacity = max(
            1,
            min(
                32,
                self.photons_per_dispatch // (scene.photons_per_column * nx * ny),
                (256 * 1024 * 1024) // bank_bytes,
            ),
        )
        native = load_library(self.library)
        error = ctypes.create_string_buffer(8192)
        handle = native.monki_hip_create(
            self.device,
            packed.ctypes.data_as(FLOAT_POINTER),
            packed.size,
            per_round * self.tally_shards,
            self._round_capacity,
            error,
            len(error),
        )

        if not handle:
            raise RuntimeError(error.value.decode() or "HIP session creation failed")

        self._native = native
        self._handle = handle
        self._packed = packed
        self.metadata = scene_metadata(tables, scene.photons_per_column, extra=extra)
        self._configuration = self._settings()
        self._round = 0
        self.provenance = {
            "backend": "HIP",
            "device": native.monki_hip_device_name(handle).decode(),
            "case": scene.name,
            "scene": tables.description,
            "rng": "Philox4x32-10; seed/round/photon/draw-block",
            "seed": self.seed,
            "transport_precision": f"float{self.precision}",
            "round_precision": "float64",
            "transport": "backward" if scene.backward else "forward",
            "photons_per_dispatch": self.photons_per_dispatch,
            "tally_shards": self.tally_shards,
            "round_capacity": self._round_capacity,
            "max_steps": self.max_steps,
            "library_sha256": file_hash(self.library),
            "tables_sha256": hashlib.sha256(packed.tobytes()).hexdigest(),
            "python_version": platform.python_version(),
            "gpu_dispatches": 0,
        }

        if preparation is not None:
            self.provenance["preparation"] = preparation

        try:
            if self.proposal is not None:
                self._train_proposal(nx * ny)
                lobes = sharpest_species(tables)
                apply_steering(native, handle, replace(self.proposal, surface=self.proposal.surface and lobes >= 0), lobes,
                               self.proposal.lobe_fraction if lobes >= 0 else 0.0)  # fmt: skip
                self.provenance["proposal"]["lobe_species"] = lobes

            if self.directed is not None:
                settings = self.directed
                error = ctypes.create_string_buffer(4096)

                if native.monki_hip_set_directed(
                    handle, settings.alpha, settings.kappa, settings.species_mask() & 0xFFFFFFFF,
                    min(settings.max_order, 2**32 - 1), error, len(error),
                ):  # fmt: skip
                    raise RuntimeError(error.value.decode())

                lobes = sharpest_species(tables) if settings.lobes == "phase" else -1
                apply_steering(native, handle, settings, lobes, 1.0)
                self.provenance["directed"] = {key: getattr(settings, key) for key in settings.__dataclass_fields__}
                self.provenance["directed"]["lobe_species"] = lobes
                shape = "phase-function" if lobes >= 0 else "vMF"
                self.provenance["method"] = f"detector-directed ({shape} lobes toward the views, defensive physical sampler)"
        except BaseException:
            self.close()
            raise

    def release_network(self) -> Optional[Proposal]:
        """Hand the trained proposal to the caller (who then closes it); the session keeps it
        attached until it is closed or prepared again."""

        network, self._network = self._network, None
        return network

    def run_batch(self, controls: object) -> RawTallies:
        """Execute exactly one complete statistical round."""

        return self._run_group([controls])[0]

    def run_batches(self, controls: list) -> list:
        """Rounds in request order. Consecutive rounds with one photon budget share a dispatch
        group of at most the round capacity (and photons_per_dispatch photons); every round
        keeps its own tally banks and random streams, so grouping changes no result."""

        self._check_requests(controls)
        return [batch for group in self._groups(controls) for batch in self._run_group(group)]

    def _groups(self, controls: list):
        _, nx, ny, _, _ = self._shape

        for _, same_budget in itertools.groupby(controls, key=lambda request: request.photons_per_column):
            run = list(same_budget)
            fit = self.photons_per_dispatch // (run[0].photons_per_column * nx * ny)
            size = max(1, min(self._round_capacity, fit))

            for start in range(0, len(run), size):
                yield run[start : start + size]

    def close(self) -> None:
        """Release device buffers; returned arrays stay valid."""

        if self._handle is not None and self._native is not None:
            self._native.monki_hip_close(self._handle)

        if self._network is not None:
            self._network.close()

        self._network = None
        self._handle = None
        self._native = None
        self.metadata = None
        self._configuration = None

    # -- diagnostics --------------------------------------------------------------------

    def counters(self) -> dict:
        """Event counters accumulated since prepare (monotonic)."""

        self._require_session()
        values = (ctypes.c_uint64 * len(COUNTER_NAMES))()

        if self._native.monki_hip_counters(self._handle, values, len(values)):
            raise RuntimeError("HIP counter readback failed")

        return dict(zip(COUNTER_NAMES, (int(v) for v in values)))

    def kernel_seconds(self) -> float:
        """Transport-kernel seconds of the last dispatch (HIP events on its stream)."""

        self._require_session()
        return float(self._native.monki_hip_kernel_seconds(self._handle))

    def trace(self, controls: BatchControls, records: int, precision: int = 32):
        """Run one round and return (RawTallies-like totals, per-photon records).

        Records hold, per photon: atmospheric events, surface events, end status
        (1 escaped, 2 absorbed, 3 failed), final weight, sum of Stokes-I contributions
        times mu0/mu, sum of detector weights, sum of weighted path contributions, last layer.
        Does not advance the session's round counter.
        """

        self._require_session()
        _, nx, ny, _, _ = self._shape
        per_round = controls.photons_per_column * nx * ny
        count = per_round
        command = (ctypes.c_uint32 * 8)(
            count, 0, controls.round_index, self.seed, self.max_steps, 0, per_round, 1
        )
        out = np.zeros(self._tally_count, dtype=np.float64)
        buffer = np.zeros((records, TRACE_WORDS), dtype=np.float32)
        error = ctypes.create_string_buffer(8192)
        status = self._native.monki_hip_trace(
            self._handle,
            command,
            precision,
            out.ctypes.data_as(DOUBLE_POINTER),
            buffer.ctypes.data_as(FLOAT_POINTER),
            records,
            error,
            len(error),
        )

        if status:
            raise RuntimeError(error.value.decode())

        return out, buffer[: min(records, count)]

    @property
    def network(self) -> Optional[Proposal]:
        """The trained proposal of a guided session (None for the plain tracer)."""

        return self._network

    def trace_guided(self, controls: BatchControls, records: int):
        """Guided counterpart of ``trace`` (records of the same layout)."""

        self._require_session()

        if self._network is None:
            raise RuntimeError("trace_guided needs a guided session")

        _, nx, ny, _, _ = self._shape
        per_round = controls.photons_per_column * nx * ny
        command = (ctypes.c_uint32 * 8)(per_round, 0, controls.round_index, self.seed, self.max_steps, 0, per_round, 1)
        out = np.zeros(self._tally_count, dtype=np.float64)
        buffer = np.zeros((records, TRACE_WORDS), dtype=np.float32)
        error = ctypes.create_string_buffer(8192)
        status = self._native.monki_hip_trace_guided(
            self._handle, command, out.ctypes.data_as(DOUBLE_POINTER), buffer.ctypes.data_as(FLOAT_POINTER),
            records, error, len(error),
        )  # fmt: skip

        if status:
            raise RuntimeError(error.value.decode())

        return out, buffer[: min(records, per_round)]

    # -- internals ----------------------------------------------------------------------

    def _train_proposal(self, columns: int) -> None:
        settings = self.proposal
        start = time.perf_counter()
        network = Proposal(self._native, self.device, settings.seed)
        self._network = network
        if settings.path_target > 0.0 and not self.save_path:
            raise ValueError("a path target for the proposal needs the layer paths (save_path=True)")

        logged = log_rounds(
            self._native, self._handle, network, settings, columns,
            (self.seed + TRAINING_SEED_OFFSET + settings.seed) % 2**32, self.max_steps,
            self.photons_per_dispatch, self._tally_count, self._shape,
        )  # fmt: skip
        logging_seconds = time.perf_counter() - start
        losses = network.train(settings.steps, settings.batch, settings.learning_rate, settings.seed)
        error = ctypes.create_string_buffer(4096)

        if self._native.monki_hip_attach_proposal(
            self._handle, network.handle, settings.alpha, settings.species_mask(), min(settings.max_order, 2**32 - 1),
            error, len(error),
        ):  # fmt: skip
            raise RuntimeError(error.value.decode())

        self.provenance["proposal"] = {
            "settings": {key: getattr(settings, key) for key in settings.__dataclass_fields__},
            "logging_rounds": logged["rounds"],
            "training_events": logged["events"],
            **({"layer_weights": logged["layer_weights"]} if "layer_weights" in logged else {}),
            "logging_seconds": logging_seconds,
            "training_seconds": time.perf_counter() - start - logging_seconds,
            "loss_first": float(losses[0]) if len(losses) else None,
            "loss_last": float(np.mean(losses[-50:])) if len(losses) else None,
        }
        self.provenance["method"] = "guided (view lobes + learned vMF mixture, defensive physical sampler)"

    def _require_session(self):
        if self._handle is None or self._native is None:
            raise RuntimeError("HIP has no prepared session: call prepare first")

    def _check_requests(self, controls: list) -> None:
        """Refuse a request list before any device work."""

        self._require_session()

        if self._settings() != self._configuration:
            raise ValueError("HIP settings differ from the prepared session's: call prepare again")

        _, nx, ny, _, _ = self._shape
        expected = self._round

        for request in controls:
            if not isinstance(request, BatchControls):
                raise TypeError(f"rounds are requested with BatchControls, got {type(request).__name__}")

            expected += 1

            if request.round_index != expected:
                raise ValueError(f"round {request.round_index} requested where round {expected} comes next")

            # Photon ids of a dispatch group (rounds x photons per round) are 32-bit.
            if request.photons_per_column * nx * ny * len(controls) >= 2**32:
                raise ValueError("photon ids of this request overflow 32 bits: request fewer rounds at once")

    @contextlib.contextmanager
    def _closed_on_failure(self):
        """Any failure inside ends the session: its device state is no longer trusted."""

        try:
            yield
        except BaseException:
            self.close()
            raise

    def _kernel_entry(self):
        native = self._native

        if self.directed is not None:
            return native.monki_hip_run_directed

        if self._network is not None:
            return native.monki_hip_run_guided

        return native.monki_hip_run if self.precision == 32 else native.monki_hip_run_fp64

    def _unpack(self, request: BatchControls, values: np.ndarray) -> RawTallies:
        """One round's float64 tallies [Stokes | paths | weights] as owned Fortran-order arrays."""

        nmat, nx, ny, nz, views = self._shape
        stokes_end = nmat * nx * ny * views
        stokes = np.array(values[:stokes_end].reshape((nmat, nx, ny, views), order="F"))

        if self.save_scats:
            return RawTallies(request, stokes, scattering=split_scattering(values, self._shape))

        if not self.save_path:
            return RawTallies(request, stokes)

        paths_end = stokes_end + nx * ny * nz * views
        paths = np.array(values[stokes_end:paths_end].reshape((nx, ny, nz, views), order="F"))
        weights = np.array(values[paths_end:].reshape((nx, ny, views), order="F"))
        return RawTallies(request, stokes, paths, weights)

    def _run_group(self, controls: list) -> List[RawTallies]:
        """One dispatch group: its photons in slices of photons_per_dispatch, the float64
        tallies of each slice added per round."""

        self._check_requests(controls)
        _, nx, ny, _, _ = self._shape
        per_round = controls[0].photons_per_column * nx * ny
        photons = per_round * len(controls)
        sums = np.zeros((len(controls), self._tally_count))
        received = np.empty_like(sums)
        error = ctypes.create_string_buffer(8192)
        entry = self._kernel_entry()

        with self._closed_on_failure():
            for offset in range(0, photons, self.photons_per_dispatch):
                words = (ctypes.c_uint32 * 8)(
                    min(self.photons_per_dispatch, photons - offset), offset, controls[0].round_index, self.seed,
                    self.max_steps, int(self.profile), per_round, len(controls),
                )  # fmt: skip

                if entry(self._handle, words, received.ctypes.data_as(DOUBLE_POINTER), error, len(error)):
                    raise RuntimeError(error.value.decode() or "HIP dispatch failed")

                if not all_finite(received.ravel()):
                    raise RuntimeError("HIP returned tallies that are not finite")

                sums += received
                self.provenance["gpu_dispatches"] += 1

            if self.profile:
                self.provenance["gpu_seconds"] = self._native.monki_hip_gpu_seconds(self._handle)

            rounds = [self._unpack(request, values) for request, values in zip(controls, sums)]
            self._round = controls[-1].round_index
            return rounds

    def _settings(self) -> tuple:
        return (
            self.device,
            self.save_path,
            self.save_scats,
            self.seed,
            self.photons_per_dispatch,
            self.max_steps,
            self.profile,
            self.tally_shards,
            self.precision,
            self.proposal,
            self.directed,
        )
