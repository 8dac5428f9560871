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
from dataclasses import replace
from pathlib import Path
from typing import List, Optional

import numpy as np

from ..input import ReferenceCase, SceneCase
from ..metadata import file_hash, scene_metadata
from ..preparation import UPSTREAM, prepare_case
from .hip import DOUBLE_POINTER, FLOAT_POINTER, LIBRARY, UINT_POINTER, load_library
from .hip_tables import pack_tables, tally_count
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
            raise NotImplementedError("HIP backward transport is not implemented")

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

            if status or not np.all(np.isfinite(out)):
                raise RuntimeError(error.value.decode() or "Nonfinite HIP batch tallies")
        except BaseException:
            self.close()
            raise

        self._round = round_index
        self.provenance["gpu_rounds"] += 1
        results, start = [], 0

        for case, shape, count in zip(self._cases, self._shapes, self._counts):
            values = out[start : start + count]
            start += count
            nmat, nx, ny, nz, views = shape
            stokes_count = nmat * nx * ny * views
            stokes = values[:stokes_count].reshape((nmat, nx, ny, views), order="F").copy()
            paths = weights = None

            if self.save_path:
                path_count = nx * ny * nz * views
                paths = values[stokes_count : stokes_count + path_count].reshape((nx, ny, nz, views), order="F").copy()
                weights = values[stokes_count + path_count :].reshape((nx, ny, views), order="F").copy()

            results.append(RawTallies(BatchControls(round_index, case.photons_per_column), stokes, paths, weights))

        return results

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

        self._handle = None
        self._native = None
