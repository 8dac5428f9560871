"""Autotuner of the batched launch: lanes per block and tally shards.

No reference counterpart. Times ``HIPBatch`` rounds of the actual batch for every candidate
(block, tally shards) -- one warm-up round, which also captures the graph, then ``rounds`` timed
replays, device time from HIP events -- and returns the fastest. Results are cached in a JSON
file keyed by the device, the scene shapes and the photon counts, so a batch is tuned once.
"""

import hashlib
import json
import time
from dataclasses import replace
from pathlib import Path
from typing import Optional

from .batch import HIPBatch

CACHE = Path(__file__).resolve().parents[2] / "lib" / "autotune.json"


def signature(cases, device_name: str) -> str:
    shapes = [(c.scene.nx, c.scene.ny, c.scene.nz, c.scene.nmat, c.scene.angles, len(c.scene.species),
               c.photons_per_column) for c in cases]  # fmt: skip
    return hashlib.sha256(json.dumps([device_name, shapes]).encode()).hexdigest()[:16]


def autotune(cases, *, blocks=(64, 128, 256), shards=(1, 8, 16, 32, 64), rounds: int = 3,
             photons_per_column: Optional[int] = None, device: int = 0, cache: Optional[Path] = CACHE,
             seed: int = 23274) -> dict:
    """Return {"block", "tally_shards", "seconds_per_round", "table", "cached"} for the batch.

    photons_per_column, when given, replaces the cases' budget for the timing rounds.
    """

    cases = list(cases)
    timing_cases = [replace(c, photons_per_column=photons_per_column, rounds=rounds + 1) if photons_per_column
                    else replace(c, rounds=rounds + 1) for c in cases]  # fmt: skip
    key = None
    stored = {}

    if cache is not None and Path(cache).is_file():
        stored = json.loads(Path(cache).read_text())

    table = []

    for block in blocks:
        for shard_count in shards:
            with HIPBatch(device, tally_shards=shard_count, block=block, seed=seed) as batch:
                batch.prepare(timing_cases)

                if key is None:
                    key = signature(timing_cases, batch.provenance["device"])

                    if key in stored:
                        return {**stored[key], "cached": True}

                batch.run_round(1)
                start = time.perf_counter()
                device_seconds = 0.0

                for index in range(2, rounds + 2):
                    batch.run_round(index)
                    device_seconds += batch.kernel_seconds()

                wall = (time.perf_counter() - start) / rounds
                table.append({"block": block, "tally_shards": shard_count,
                              "device_seconds": device_seconds / rounds, "wall_seconds": wall})  # fmt: skip

    best = min(table, key=lambda row: row["device_seconds"])
    result = {"block": best["block"], "tally_shards": best["tally_shards"],
              "seconds_per_round": best["device_seconds"], "table": table}  # fmt: skip

    if cache is not None and key is not None:
        stored[key] = result
        Path(cache).parent.mkdir(parents=True, exist_ok=True)
        Path(cache).write_text(json.dumps(stored, indent=1) + "\n")

    return {**result, "cached": False}
