"""Benchmark of the MONKI HIP transport on one MI355X: photon-events/s and time-to-1%-AMF-SE.

Headline line (``python bench.py``): photon-events/s of the plain tracer on the largest workload
that fits one GPU, the spectral stochastic cloud field: 64 wavelengths (300-500 nm) of one
128 x 128 x 60 field (Kolmogorov spectrum, cover 0.6, mean cloudy optical depth 12, cloud
1-4 km, Rayleigh lambda^-4.08, synthetic ozone), I/Q/U at 3 views, layer AMF tallies, traced as
one batch: one graph-replayed launch per round, rounds accumulated on the device. A step is one
round of all 64 scenes (16 photons per column, 16.8e6 photons); tables are resident in HBM
before timing starts.

value = photon events (atmospheric collisions + surface hits, counted in-kernel) per second of
the timed rounds. roofline: algorithmic bytes per event (4 B per free-path cell crossing, per
explicit-equivalent detector cell and per phase-table entry read, 8 B read-modify-write per tally
element per view) over the average device time of a round (HIP events on the batch stream),
against the 8 TB/s HBM peak; ``traffic`` is the measured HBM bytes of the transport kernel
(rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size), and ``binding`` names what
actually limits the kernel (the tally's record atomics). cpu_baseline: the reference's Fortran
transport (mc.f90, amdflang -O2 -fopenmp) on a bounded sample of one wavelength of the same
field, all host threads this job has, one warm-up and the median of 3 runs, with 1 -> N thread
scaling.

time_to_1pct_amf_se: (a) the reference's own published setting (fixtures clear and
layered-cloud, 8,000 photons per round, AMFConvergence(), 8,192-round cap, median of 3 seeds
after a warm-up; validation/benchmark_amf_convergence.py:27-36) against its published Apple M4
times; (b) the 3-D box cloud (30 x 20 x 35, tau 20, I/Q/U) with the plain, detector-directed
and guided tracers, each to a fixed round cap with the worst relative AMF SE and the projection
to 1 % (SE ~ rounds^-1/2); (c) the same race on four wavelengths of the spectral field in one
batch per tracer.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-amf]
       python bench.py --layered-cloud          (the 30x20x35 layered-cloud round instead)
       python bench.py --race [--amf-rounds R]  (only the box-cloud race)
       python bench.py --spectral-race          (only the spectral-field race)
       python bench.py --batch [--steps K]      (only the 64-scene box-cloud batch)
       python bench.py --train [--steps K]      (only the proposal's Adam steps, 16,384 samples each)
       python bench.py --copy-calibration | --summarize-traffic DIR | --summarize-profile DB
"""

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent

for _path in (ROOT / "monki-gpu_amd", ROOT):
    if str(_path) not in sys.path:
        sys.path.insert(0, str(_path))

METRIC = "photon-events/s and time-to-1%-AMF-SE, guided vs plain MC, 1×MI355X"
HBM_PEAK_GBS = 8000.0
SPECTRAL_PPC = 16
LAYERED_PPC = 10_000
# Published Apple M4 medians of the reference (validation/README.md:80-84).
PUBLISHED_TIME_TO_1PCT = {"clear": {"metal": 0.466, "cpu_4_threads": 2.074},
                          "layered-cloud": {"metal": 26.160, "cpu_4_threads": 232.839}}  # fmt: skip


def algorithmic_bytes(counts: dict, views: int, nmat: int, nz: int) -> float:
    events = counts["atmospheric_events"] + counts["surface_events"]
    reads = 4.0 * (counts["walk_cells"] + counts["detector_cells"] + counts["table_reads"])
    tallies = 8.0 * events * views * (nmat + nz + 1)
    return reads + tallies


def delta(after: dict, before: dict) -> dict:
    return {key: after[key] - before[key] for key in after}


def host_threads() -> int:
    """Threads this job may use on the host: OMP_NUM_THREADS when the environment sets it (the
    GPU pool sets it to the job's CPU share), else the CPUs of the process's affinity mask."""

    configured = os.environ.get("OMP_NUM_THREADS", "")
    return int(configured) if configured.isdigit() and int(configured) > 0 else len(os.sched_getaffinity(0))


def cpu_baseline(scene, events_per_photon: float) -> dict:
    """The reference Fortran transport on one wavelength of the spectral field. Thread scaling
    1, 2, 4, ... up to every host thread this job has (one run each, one photon per thread and
    column); then, at the fastest thread count, one warm-up and the median of 3 runs. The
    compiler's RANDOM_NUMBER serializes across OpenMP threads, so the fastest count can be
    below the number of cores; the value is reported at the fastest count, the all-threads rate
    beside it."""

    from oracle.oracle import REFERENCE, run_reference

    threads = host_threads()
    columns = scene.nx * scene.ny
    base = {"unit": "photon-events/s", "kind": "reference", "host_threads": threads,
            "affinity_cpus": len(os.sched_getaffinity(0))}  # fmt: skip

    if not REFERENCE.is_file():
        return {**base, "cores": threads, "value": None, "sample": "reference build missing (make -C oracle ref)"}

    scaling = {}
    count = 1

    while True:
        scaling[count] = count * columns / run_reference(scene, count, 1, threads=count)[1][0]

        if count >= threads:
            break

        count = min(2 * count, threads)

    best = max(scaling, key=scaling.get)
    run_reference(scene, best, 1, threads=best)  # warm-up
    runs = [run_reference(scene, best, 1, seed=23274 + k, threads=best)[1][0] for k in range(3)]
    seconds = statistics.median(runs)
    photons_per_second = best * columns / seconds
    return {
        **base,
        "cores": best,
        "value": photons_per_second * events_per_photon,
        "sample": (
            f"{scene.description or 'spectral 400 nm'}: {best} photons/column x {columns} columns "
            f"({best * columns} photons) per run at {best} threads (the fastest of the scaling sweep), transport "
            f"seconds of 3 runs {[round(r, 2) for r in runs]}, median {seconds:.2f} s after a warm-up; events/s = "
            f"GPU-counted events per photon of the same scene ({events_per_photon:.2f}) x CPU photons/s "
            f"({photons_per_second:.0f})"
        ),
        "photons_per_second": photons_per_second,
        "all_threads_photon_events_per_s": scaling[threads] * events_per_photon,
        "thread_scaling_photons_per_s": {str(k): v for k, v in scaling.items()},
    }


def spectral_cases(rounds: int):
    from monki import SceneCase, configs

    return [SceneCase(f"spectral-{i:02d}", scene, SPECTRAL_PPC, rounds)
            for i, scene in enumerate(configs.spectral_scenes(64))]  # fmt: skip


def spectral_line(steps: int, warmup: int) -> dict:
    """Timed rounds of the 64-wavelength batch with the autotuned launch shape."""

    return batch_line(spectral_cases(steps + warmup + 1), steps, warmup, shards=(1, 4))


BOX_BATCH_PPC = 1000


def box_batch_line(steps: int = 3, warmup: int = 1) -> dict:
    """The 64-scene box-cloud batch (one seed: SZA U(0, 70), albedo U(0.02, 0.3), cloud optical
    depth lognormal with median 15, grids 16-30 x 12-20 x 35), 1,000 photons per column."""

    from monki import SceneCase, configs

    cases = [SceneCase(f"box-{i:02d}", scene, BOX_BATCH_PPC, steps + warmup + 1)
             for i, scene in enumerate(configs.box_cloud_batch(64))]  # fmt: skip
    run = batch_line(cases, steps, warmup, shards=(1, 8))
    return {
        "workload": "64 box-cloud scenes from one seed (grids 16-30 x 12-20 x 35, SZA U(0, 70), albedo U(0.02, 0.3), "
        f"cloud tau lognormal median 15), I/Q/U, 3 views, layer paths, {BOX_BATCH_PPC} photons/column, one "
        "graph-replayed launch per round",
        "photon_events_per_s": run["value"],
        "ms_per_round": 1e3 * run["elapsed"] / steps,
        "device_ms_per_round": 1e3 * run["device_round"],
        "photons_per_round": run["photons"] // steps,
        "autotune": run["autotune"],
    }


def batch_line(cases, steps: int, warmup: int, shards=(1, 4)) -> dict:
    """Timed rounds of a batch (rounds accumulated on the device) with the autotuned shape."""

    from monki.execution import HIPBatch
    from monki.execution.autotune import autotune

    tuned = autotune(cases, blocks=(128, 256), shards=shards, rounds=1)
    scene = cases[0].scene

    with HIPBatch(block=tuned["block"], tally_shards=tuned["tally_shards"]) as batch:
        batch.prepare(cases)

        for index in range(1, warmup + 1):
            batch.accumulate_round(index)

        before = batch.counters()
        device = 0.0
        start = time.perf_counter()

        for index in range(warmup + 1, warmup + steps + 1):
            batch.accumulate_round(index)
            device += batch.kernel_seconds()

        elapsed = time.perf_counter() - start
        counts = delta(batch.counters(), before)
        name = batch.provenance["device"]

    events = counts["atmospheric_events"] + counts["surface_events"]
    bytes_per_round = algorithmic_bytes(counts, scene.angles, scene.nmat, scene.nz) / steps
    device_round = device / steps
    return {
        "value": events / elapsed,
        "elapsed": elapsed,
        "device_round": device_round,
        "bytes_per_round": bytes_per_round,
        "counts": counts,
        "photons": counts["photons"],
        "events": events,
        "device": name,
        "autotune": {key: tuned[key] for key in ("block", "tally_shards", "seconds_per_round", "cached")},
    }


def single_scene_events_per_photon(scene) -> float:
    """Events per photon of one scene on the GPU (one round at 4 photons per column)."""

    from monki import SceneCase
    from monki.execution import HIP, BatchControls

    with HIP() as backend:
        backend.prepare(SceneCase("events", scene, 4, 1))
        backend.run_batch(BatchControls(1, 4))
        counts = backend.counters()

    return (counts["atmospheric_events"] + counts["surface_events"]) / counts["photons"]


def layered_cloud_line(steps: int = 5, warmup: int = 1) -> dict:
    """The 30 x 20 x 35 layered-cloud round (10,000 photons per column), one scene."""

    from monki import SceneCase, configs
    from monki.execution import HIP, BatchControls

    scene = configs.layered_cloud()
    backend = HIP()
    backend.prepare(SceneCase("layered-cloud", scene, LAYERED_PPC, warmup + steps))

    for index in range(1, warmup + 1):
        backend.run_batch(BatchControls(index, LAYERED_PPC))

    before = backend.counters()
    kernel = 0.0
    start = time.perf_counter()

    for index in range(warmup + 1, warmup + steps + 1):
        backend.run_batch(BatchControls(index, LAYERED_PPC))
        kernel += backend.kernel_seconds()

    elapsed = time.perf_counter() - start
    counts = delta(backend.counters(), before)
    backend.close()
    events = counts["atmospheric_events"] + counts["surface_events"]
    achieved = algorithmic_bytes(counts, scene.angles, scene.nmat, scene.nz) / kernel / 1e9
    return {
        "workload": "layered cloud 30x20x35, water cloud tau 10 + HG aerosol + Rayleigh/O3, I/Q/U, 3 views, "
        f"layer paths, {LAYERED_PPC} photons/column (6e6 photons/round), plain tracer",
        "photon_events_per_s": events / elapsed,
        "ms_per_round": 1e3 * elapsed / steps,
        "kernel_ms_per_round": 1e3 * kernel / steps,
        "events_per_photon": events / counts["photons"],
        "algorithmic_gb_per_s": achieved,
        "tally_record_instructions_per_s": counts["tally_atomics"] / (scene.nz + scene.nmat) / kernel,
    }


def reference_fixture_amf(seeds=(23274, 714103, 19088743)) -> dict:
    """Time to 1 % AMF SE at the reference's published setting, HIP backend."""

    import tempfile
    from dataclasses import replace

    from monki import AMFConvergence, ReferenceCase, simulate
    from monki.execution import HIP

    results = {}

    for name, fixture in (("clear", "cloudfree"), ("layered-cloud", "1layercloud")):
        case = ReferenceCase(name, f"testrun/MONKI_MLS_{fixture}_340_340_test.in", photons_per_column=8000,
                             rounds=8192)  # fmt: skip
        samples = []

        with tempfile.TemporaryDirectory() as directory:
            simulate(replace(case, rounds=2), HIP(), Path(directory) / "warm.nc")

            for seed in seeds:
                path = Path(directory) / f"{seed}.nc"
                start = time.perf_counter()
                simulate(case, HIP(seed=seed), path, convergence=AMFConvergence())
                seconds = time.perf_counter() - start
                receipt = json.loads(path.with_suffix(".run.json").read_text())
                samples.append({"seed": seed, "seconds": seconds, "status": receipt["status"],
                                "rounds": receipt["completed_rounds"],
                                "max_relative_se": receipt["amf_convergence"]["checks"][-1]["max_relative_se"]})  # fmt: skip

        median = statistics.median(sample["seconds"] for sample in samples)
        published = PUBLISHED_TIME_TO_1PCT[name]
        results[name] = {
            "median_seconds": median,
            "all_converged": all(sample["status"] == "converged" for sample in samples),
            "samples": samples,
            "published_apple_m4": published,
            "speedup_vs_published_metal": published["metal"] / median,
            "speedup_vs_published_cpu": published["cpu_4_threads"] / median,
        }

    results["setting"] = ("reference fixtures (1 x 1 column, SZA 45), 8000 photons/round, AMFConvergence(0.01, 128, "
                          "32, 3), cap 8192 rounds, end to end (prepare, transport, statistics, NetCDF, receipt), "
                          "median of 3 seeds after a warm-up")  # fmt: skip
    return results


def capped_amf_run(scene, name: str, rounds: int, photons_per_column: int, **backend) -> dict:
    """simulate(..., AMFConvergence(0.01, rounds, 32, 3)) to a cap of ``rounds``: wall time, the
    relative AMF SE of every cell at the cap, and the projected rounds and seconds to 1 %."""

    import tempfile

    import numpy as np

    from monki import AMFConvergence, SceneCase, simulate
    from monki.execution import HIP

    policy = AMFConvergence(0.01, rounds, 32, 3)
    case = SceneCase(name, scene, photons_per_column, rounds)

    with tempfile.TemporaryDirectory() as directory:
        start = time.perf_counter()
        result = simulate(case, HIP(**backend), Path(directory) / "amf.nc", convergence=policy)
        elapsed = time.perf_counter() - start
        receipt = json.loads((Path(directory) / "amf.run.json").read_text())

    relative = np.abs(result.values["std_AMF_layer"]) / np.abs(result.values["AMF_layer"])
    finite = np.sort(relative[np.isfinite(relative)].ravel())
    worst = float(finite[-1])
    record = {
        "seconds": elapsed,
        "status": receipt["status"],
        "rounds": receipt["completed_rounds"],
        "method": receipt.get("method", "plain"),
        "max_relative_se": worst,
        "p999_relative_se": float(finite[int(0.999 * (finite.size - 1))]),
        "worst_decile_mean_relative_se": float(finite[-max(1, finite.size // 10):].mean()),
        "median_relative_se": float(np.median(finite)),
        "cells_above_5pct": int((finite > 0.05).sum()),
        "cells": int(relative.size),
    }

    if "proposal" in receipt:
        record["training"] = {key: receipt["proposal"][key] for key in
                              ("training_events", "logging_seconds", "training_seconds", "loss_first", "loss_last")}  # fmt: skip

    # SE falls as rounds^-1/2: rounds and seconds the same run needs for 1 % in every cell (and in
    # 99.9 % of the cells), the fixed set-up time (training) counted once.
    fixed = record.get("training", {}).get("logging_seconds", 0.0) + record.get("training", {}).get("training_seconds", 0.0)
    per_round = (elapsed - fixed) / record["rounds"]
    for key, se in (("", worst), ("_p999", record["p999_relative_se"])):
        needed = record["rounds"] * max(se / 0.01, 1.0) ** 2
        record[f"projected_rounds{key}"] = needed
        record[f"projected_seconds{key}"] = fixed + per_round * needed

    return record


def amf_race(rounds: int = 128, photons_per_column: int = 10_000) -> dict:
    """Plain, detector-directed and guided tracers on the 3-D box cloud (30 x 20 x 35, tau 20 over
    10 x 6 columns, layers 2-6 km, Mie, I/Q/U, 3 views) to the same round cap."""

    from monki import configs
    from monki.execution.proposal import DirectedSettings, ProposalSettings

    scene = configs.box_cloud()
    legs = {
        "plain": capped_amf_run(scene, "box-cloud", rounds, photons_per_column),
        "directed": capped_amf_run(scene, "box-cloud", rounds, photons_per_column, directed=DirectedSettings()),
        "guided": capped_amf_run(scene, "box-cloud", rounds, photons_per_column, proposal=ProposalSettings()),
    }
    plain = legs["plain"]
    result = {"workload": "box cloud 30x20x35, tau 20 over 10x6 columns x layers 2-6 km, Mie r_eff 10 um, I/Q/U, "
                          f"3 views, {photons_per_column} photons/column, cap {rounds} rounds",  # fmt: skip
              **legs}

    for name in ("directed", "guided"):
        leg = legs[name]
        result[f"{name}_speedup_to_1pct"] = plain["projected_seconds"] / leg["projected_seconds"]
        result[f"{name}_speedup_to_1pct_p999"] = plain["projected_seconds_p999"] / leg["projected_seconds_p999"]
        # Per-photon variance at equal rounds, worst decile of cells.
        result[f"{name}_worst_decile_variance_ratio"] = (plain["worst_decile_mean_relative_se"] /
                                                         leg["worst_decile_mean_relative_se"]) ** 2  # fmt: skip

    result["speedup_to_1pct"] = result["guided_speedup_to_1pct"]
    return result


SPECTRAL_RACE_ROUNDS = 32


def spectral_race(rounds: int = SPECTRAL_RACE_ROUNDS) -> dict:
    """Plain, detector-directed and guided tracers on every 16th wavelength of the spectral field
    (300, 351, 402 and 452 nm of the 64; 128 x 128 x 60, 16 photons per column), each a batch of
    the four scenes accumulated on the device, to the same round cap: the relative layer-AMF SE of
    every cell at the cap and the projected seconds to 1 % (training counted once)."""

    import numpy as np

    from monki import SceneCase, configs
    from monki.execution import HIPBatch
    from monki.execution.proposal import DirectedSettings, ProposalSettings

    scenes = configs.spectral_scenes(64)[::16]
    cases = [SceneCase(f"spectral-{i:02d}", scene, SPECTRAL_PPC, rounds) for i, scene in enumerate(scenes)]
    legs = {}

    for name, backend in (("plain", {}), ("directed", {"directed": DirectedSettings()}),
                          ("guided", {"proposal": ProposalSettings()})):  # fmt: skip
        with HIPBatch(block=128, tally_shards=1, **backend) as batch:
            batch.prepare(cases)
            proposal = batch.provenance.get("proposal", {})
            fixed = proposal.get("logging_seconds", 0.0) + proposal.get("training_seconds", 0.0)
            device = 0.0
            start = time.perf_counter()

            for index in range(1, rounds + 1):
                batch.accumulate_round(index)
                device += batch.kernel_seconds()

            elapsed = time.perf_counter() - start
            relative = []

            for number, case in enumerate(cases):
                cells = batch.moments(number).amf_moments().relative_error(case.scene.zunit[:-1])
                relative.append(cells[np.isfinite(cells)].ravel())

        finite = np.sort(np.concatenate(relative))
        worst, p999 = float(finite[-1]), float(finite[int(0.999 * (finite.size - 1))])
        leg = {"seconds": elapsed, "device_seconds": device, "rounds": rounds, "max_relative_se": worst,
               "p999_relative_se": p999, "worst_decile_mean_relative_se": float(finite[-max(1, finite.size // 10):].mean()),
               "median_relative_se": float(np.median(finite)), "cells": int(finite.size)}  # fmt: skip

        if proposal:
            leg["training"] = {key: proposal[key] for key in ("training_events", "logging_seconds", "training_seconds")}

        for key, se in (("", worst), ("_p999", p999)):
            needed = rounds * max(se / 0.01, 1.0) ** 2
            leg[f"projected_rounds{key}"] = needed
            leg[f"projected_seconds{key}"] = fixed + elapsed / rounds * needed

        legs[name] = leg

    plain = legs["plain"]
    result = {"workload": "spectral field 128x128x60 at 300, 351, 402 and 452 nm in one batch, I/Q/U, 3 views, "
                          f"layer AMF tallies, {SPECTRAL_PPC} photons/column, cap {rounds} rounds",  # fmt: skip
              **legs}

    for name in ("directed", "guided"):
        leg = legs[name]
        result[f"{name}_speedup_to_1pct"] = plain["projected_seconds"] / leg["projected_seconds"]
        result[f"{name}_speedup_to_1pct_p999"] = plain["projected_seconds_p999"] / leg["projected_seconds_p999"]
        result[f"{name}_worst_decile_variance_ratio"] = (plain["worst_decile_mean_relative_se"] /
                                                         leg["worst_decile_mean_relative_se"]) ** 2  # fmt: skip

    return result


# Network arithmetic per training sample: forward (2 x 56,360 multiply-adds less the biases),
# input gradients of layers 1-4 and weight gradients of all five layers, about 332 kFLOP.
TRAIN_FLOP_PER_SAMPLE = 2 * (12 * 128 + 3 * 128 * 128 + 128 * 40) * 3 - 2 * 12 * 128
TRAIN_BATCH = 16_384


def train_line(steps: int, warmup: int) -> dict:
    """Adam steps of the learned proposal on 1e6 synthetic samples (uniform features, isotropic
    directions, weights in [0.1, 2)): wall time per step of one `train` call."""

    import numpy as np

    from monki.execution.hip import load_library
    from monki.execution.proposal import Proposal

    native = load_library()
    proposal = Proposal(native, 0, seed=3)
    rng = np.random.default_rng(1)
    count = 1_000_000
    features = rng.uniform(-1.0, 1.0, (count, 12)).astype(np.float32)
    directions = rng.normal(size=(count, 3))
    directions /= np.linalg.norm(directions, axis=1, keepdims=True)
    weights = rng.uniform(0.1, 2.0, count).astype(np.float32)

    try:
        proposal.set_samples(features, directions.astype(np.float32), weights)
        proposal.train(max(warmup, 1), TRAIN_BATCH, 1e-3)
        start = time.perf_counter()
        losses = proposal.train(steps, TRAIN_BATCH, 1e-3)
        seconds = (time.perf_counter() - start) / steps
    finally:
        proposal.close()

    return {"workload": "proposal training", "batch": TRAIN_BATCH, "steps": steps,
            "ms_per_step": 1e3 * seconds, "samples_per_s": TRAIN_BATCH / seconds,
            "gflop_per_step": TRAIN_FLOP_PER_SAMPLE * TRAIN_BATCH / 1e9,
            "tflops_wall": TRAIN_FLOP_PER_SAMPLE * TRAIN_BATCH / seconds / 1e12,
            "loss_first": float(losses[0]), "loss_last": float(losses[-1])}  # fmt: skip


COPY_BYTES = 1 << 30
TRAFFIC_FILES = {"spectral": ROOT / "profiles" / "traffic_spectral_field.json",
                 "layered-cloud": ROOT / "profiles" / "traffic_layered_cloud.json"}  # fmt: skip


def copy_calibration() -> None:
    """A plain device-to-device copy of COPY_BYTES, three times, for counter calibration."""

    import torch

    source = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    target = torch.empty_like(source)

    for _ in range(3):
        target.copy_(source)

    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": COPY_BYTES, "copies": 3}), flush=True)


def _counter_values(directory: Path, name: str, match) -> list:
    """Per-dispatch counter values of matching kernels from one rocprofv3 --pmc run (CSV or
    the SQLite database that rocprofv3 writes by default)."""

    import csv
    import sqlite3

    table = directory / f"{name}_counter_collection.csv"

    if table.is_file():
        with table.open() as handle:
            rows = [row for row in csv.DictReader(handle) if match(row["Kernel_Name"])]

        return [float(row["Counter_Value"]) for row in rows]

    with sqlite3.connect(directory / f"{name}_results.db") as db:
        rows = db.execute("SELECT kernel_name, value FROM counters_collection").fetchall()

    return [float(value) for kernel, value in rows if match(kernel)]


def summarize_traffic(directory: Path, workload: str) -> dict:
    """HBM bytes per transport launch from the four --pmc runs written under ``directory``.

    Expected runs (rocprofv3 -d DIR -o NAME): transport_fetch, transport_write, copy_fetch and
    copy_write. Counter values are KiB per dispatch.
    """

    def read(name: str, match) -> list:
        return _counter_values(directory, name, match)

    transport = lambda name: "transport_kernel" in name  # noqa: E731
    copy = lambda name: "copy" in name.lower() or "elementwise" in name.lower()  # noqa: E731
    fetch = read("transport_fetch", transport)
    write = read("transport_write", transport)
    copy_fetch = max(read("copy_fetch", copy))
    copy_write = max(read("copy_write", copy))
    fetch_factor = COPY_BYTES / (copy_fetch * 1024.0)
    write_factor = COPY_BYTES / (copy_write * 1024.0)
    fetch_bytes = sum(fetch) / len(fetch) * 1024.0 * fetch_factor
    write_bytes = sum(write) / len(write) * 1024.0 * write_factor
    flag = "" if workload == "spectral" else " --layered-cloud"
    summary = {
        "workload": workload,
        "hbm_bytes_per_launch": fetch_bytes + write_bytes,
        "fetch_bytes_per_launch": fetch_bytes,
        "write_bytes_per_launch": write_bytes,
        "fetch_factor": fetch_factor,
        "write_factor": write_factor,
        "transport_dispatches": [len(fetch), len(write)],
        "copy_kib": [copy_fetch, copy_write],
        "commands": [
            f"rocprofv3 --pmc FETCH_SIZE -d DIR -o transport_fetch -- python3 bench.py --steps 2 --warmup 0 --no-cpu --no-amf --no-layered{flag}",
            f"rocprofv3 --pmc WRITE_SIZE -d DIR -o transport_write -- python3 bench.py --steps 2 --warmup 0 --no-cpu --no-amf --no-layered{flag}",
            "rocprofv3 --pmc FETCH_SIZE -d DIR -o copy_fetch -- python3 bench.py --copy-calibration",
            "rocprofv3 --pmc WRITE_SIZE -d DIR -o copy_write -- python3 bench.py --copy-calibration",
            f"python3 bench.py --summarize-traffic DIR{flag}",
        ],
    }
    TRAFFIC_FILES[workload].write_text(json.dumps(summary, indent=1) + "\n")
    return summary


def summarize_profile(database: Path) -> str:
    """Markdown summary of a rocprofv3 --kernel-trace --stats run (its SQLite output)."""

    import sqlite3

    with sqlite3.connect(database) as db:
        top = db.execute("SELECT name, total_calls, total_duration, average, percentage FROM top_kernels").fetchall()
        dispatches = db.execute(
            "SELECT name, duration, grid_x, workgroup_x, lds_size, scratch_size, vgpr_count, "
            "accum_vgpr_count, sgpr_count FROM kernels ORDER BY start"
        ).fetchall()

    def short(name: str) -> str:
        name = name.replace("(anonymous namespace)::", "").removeprefix("void ")
        return name.split("(")[0]

    lines = ["| kernel | calls | total ms | average ms | % |", "|---|---|---|---|---|"]
    lines += [f"| `{short(n)}` | {c} | {t / 1e3:.3f} | {a / 1e3:.3f} | {p:.3f} |" for n, c, t, a, p in top]
    lines += ["", "| kernel | duration ms | grid | block | LDS B | scratch B | VGPR (as reported) | AGPR | SGPR |",
              "|---|---|---|---|---|---|---|---|---|"]  # fmt: skip
    lines += [
        f"| `{short(n)}` | {d / 1e6:.3f} | {g} | {w} | {lds} | {sc} | {v} | {acc} | {sg} |"
        for n, d, g, w, lds, sc, v, acc, sg in dispatches
        if "transport" in n or "reduce" in n or "assemble" in n
    ]
    # The proposal's kernels run thousands of times: the launch shape of the first of each.
    first = {}

    for n, d, g, w, lds, sc, v, acc, sg in dispatches:
        if any(key in n for key in ("train_kernel", "weight_gradients", "sum_gradients", "pack_kernel", "adam_kernel")):
            first.setdefault(short(n), f"| `{short(n)}` (first) | {d / 1e6:.3f} | {g} | {w} | {lds} | {sc} | {v} | {acc} | {sg} |")

    lines += list(first.values())
    return "\n".join(lines) + "\n"


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    parser.add_argument("--no-amf", action="store_true", help="skip the time-to-1%% AMF SE legs")
    parser.add_argument("--amf-rounds", type=int, default=128, help="round cap of the box-cloud race")
    parser.add_argument("--layered-cloud", action="store_true", help="time the layered-cloud round instead")
    parser.add_argument("--race", action="store_true", help="run only the box-cloud race (plain, directed, guided)")
    parser.add_argument("--spectral-race", action="store_true",
                        help="run only the spectral-field race (plain, directed, guided; four wavelengths)")
    parser.add_argument("--batch", action="store_true", help="time only the 64-scene box-cloud batch")
    parser.add_argument("--train", action="store_true", help="time only the proposal's Adam steps")
    parser.add_argument("--no-layered", action="store_true", help="skip the secondary layered-cloud and batch lines")
    parser.add_argument("--copy-calibration", action="store_true", help="run the copy kernel only")
    parser.add_argument("--summarize-traffic", type=Path, metavar="DIR")
    parser.add_argument("--summarize-profile", type=Path, metavar="DB")
    args = parser.parse_args()

    if args.summarize_profile:
        print(summarize_profile(args.summarize_profile), end="")
        return 0

    if args.copy_calibration:
        copy_calibration()
        return 0

    workload = "layered-cloud" if args.layered_cloud else "spectral"

    if args.summarize_traffic:
        print(json.dumps(summarize_traffic(args.summarize_traffic, workload)), flush=True)
        return 0

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU: there is no multi-GPU path (asked for {args.gpus}). "
            "Photon histories would shard by photon range, round or scene, with one reduction of "
            "each round's tallies.",
            file=sys.stderr,
        )
        return 2

    traffic = None

    if TRAFFIC_FILES[workload].is_file():
        traffic = json.loads(TRAFFIC_FILES[workload].read_text()).get("hbm_bytes_per_launch")

    if args.layered_cloud:
        leg = layered_cloud_line(args.steps, args.warmup)
        print(json.dumps(leg), flush=True)
        return 0

    if args.race:
        print(json.dumps(amf_race(args.amf_rounds)), flush=True)
        return 0

    if args.spectral_race:
        print(json.dumps(spectral_race()), flush=True)
        return 0

    if args.batch:
        print(json.dumps(box_batch_line(args.steps, args.warmup)), flush=True)
        return 0

    if args.train:
        print(json.dumps(train_line(args.steps, args.warmup)), flush=True)
        return 0

    print(f"bench: spectral batch, {args.warmup} warm-up + {args.steps} timed rounds", file=sys.stderr, flush=True)
    run = spectral_line(args.steps, args.warmup)
    achieved = run["bytes_per_round"] / run["device_round"] / 1e9
    # The in-kernel counter counts lane atomics; one record (nz layer paths + nmat Stokes values of
    # a pixel and view) is one wave-wide atomic instruction.
    record = 60 + 3
    lane_atomics = run["counts"]["tally_atomics"]
    records_rate = lane_atomics / record / (run["device_round"] * args.steps)
    line = {
        "metric": METRIC,
        "value": run["value"],
        "unit": "photon-events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * run["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "spectral stochastic cloud field, plain tracer: 64 wavelengths (300-500 nm) x 128x128 "
            "columns x 60 layers, lognormal Kolmogorov cloud 1-4 km (cover 0.6, mean tau 12), Rayleigh "
            f"lambda^-4.08 + synthetic O3, I/Q/U, 3 views, layer AMF tallies, {SPECTRAL_PPC} photons/column "
            "(16.8e6 photons per round), one graph-replayed batch launch per round",
            "photons_per_step": run["photons"] // args.steps,
            "events_per_photon": run["events"] / run["photons"],
            "device_ms_per_step": 1e3 * run["device_round"],
            "device": run["device"],
            "autotune": run["autotune"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "basis": "algorithmic bytes per round / device time per round (HIP events around the replayed round)",
            "measured_hbm_frac": traffic / run["device_round"] / 1e9 / HBM_PEAK_GBS if traffic else None,
            "binding": {"kind": "latency and lane divergence of the walks (SQ counters in the committed profile: VALU "
                                "busy 0.51 at lane utilization 0.34), then the tally's record atomics: one wave-wide "
                                "float atomic instruction per record write-out (a pixel's 63-value record of layer "
                                "paths and Stokes components)",
                        "record_instructions_per_s": records_rate,
                        "lane_atomics_per_s": lane_atomics / (run["device_round"] * args.steps),
                        "records_per_event": lane_atomics / record / run["events"]},  # fmt: skip
        },
    }

    if not args.no_amf:
        print("bench: time to 1% AMF SE, reference fixtures", file=sys.stderr, flush=True)
        fixtures = reference_fixture_amf()
        print("bench: box-cloud race, plain / directed / guided", file=sys.stderr, flush=True)
        race = amf_race(args.amf_rounds)
        print("bench: spectral-field race, plain / directed / guided", file=sys.stderr, flush=True)
        spectral = spectral_race()
        line["time_to_1pct_amf_se"] = {"reference_fixtures": fixtures, "box_cloud_guided_vs_plain": race,
                                       "spectral_guided_vs_plain": spectral}  # fmt: skip
        line["guided_vs_plain_speedup_to_1pct"] = race["speedup_to_1pct"]

    if not args.no_layered:
        print("bench: layered-cloud round", file=sys.stderr, flush=True)
        line["layered_cloud"] = layered_cloud_line()
        print("bench: 64-scene box-cloud batch", file=sys.stderr, flush=True)
        line["box_cloud_batch"] = box_batch_line()

    if not args.no_cpu:
        from monki import configs

        print("bench: CPU baseline (reference Fortran)", file=sys.stderr, flush=True)
        scene = configs.spectral_cloud_field(400.0, configs.stochastic_cloud_field(128, seed=7))
        line["cpu_baseline"] = cpu_baseline(scene, single_scene_events_per_photon(scene))

    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
