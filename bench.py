"""Benchmark: photon-events/s of the plain HIP tracer on the layered-cloud workload, 1 GPU.

A step is one complete statistical round of the layered-cloud scene with polarization:
30 x 20 columns x 35 layers (incl. vacuum), water cloud tau 10 at 2-3 km (Mie, r_eff 10 um),
HG aerosol tau 0.2, Rayleigh + ozone, I/Q/U at 3 views with per-layer path (AMF) tallies,
10,000 photons per column (6,000,000 photons per round). Tables are resident in HBM before
timing starts; the timed region covers K rounds: transport kernel, float64 shard reduction
and the copy of the round's tallies to the host.

value = photon events (atmospheric collisions + surface hits, counted in-kernel) per second.
roofline: algorithmic bytes per event (4 B per free-path cell crossing, per explicit-equivalent
detector cell and per phase-table entry read, plus 8 B read-modify-write per tally element
per view) over the transport kernel's average duration (HIP events on its stream), against
the 8 TB/s HBM peak. cpu_baseline: the reference's Fortran transport (mc.f90, amdflang -O2
-fopenmp) on a bounded sample of the same scene on the host cores.

traffic: HBM bytes per transport launch from two rocprofv3 --pmc runs (FETCH_SIZE, then
WRITE_SIZE, KiB per dispatch), each scaled by what the same counter reads for a plain 1 GiB
device copy (--copy-calibration); --summarize-traffic DIR turns the four runs into
profiles/traffic_layered_cloud.json, which later bench runs report.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--amf] [--no-cpu]
       python bench.py --copy-calibration
       python bench.py --summarize-traffic DIR
"""

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent

for _path in (ROOT / "monki-gpu_amd", ROOT):
    if str(_path) not in sys.path:
        sys.path.insert(0, str(_path))

METRIC = "photon-events/s and time-to-1%-AMF-SE, guided vs plain MC, 1×MI355X"
HBM_PEAK_GBS = 8000.0
PHOTONS_PER_COLUMN = 10_000


def algorithmic_bytes(counts: dict, views: int, nmat: int, nz: int) -> float:
    events = counts["atmospheric_events"] + counts["surface_events"]
    reads = 4.0 * (counts["walk_cells"] + counts["detector_cells"] + counts["table_reads"])
    tallies = 8.0 * events * views * (nmat + nz + 1)
    return reads + tallies


def delta(after: dict, before: dict) -> dict:
    return {key: after[key] - before[key] for key in after}


def cpu_baseline(scene, events_per_photon: float, target_seconds: float = 15.0) -> dict:
    """Time the reference Fortran transport on a bounded sample of the same scene."""

    from oracle.oracle import REFERENCE, run_reference

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    columns = scene.nx * scene.ny

    if not REFERENCE.is_file():
        return {"value": None, "unit": "photon-events/s", "cores": threads, "kind": "reference",
                "sample": "reference build missing (make -C oracle ref)"}  # fmt: skip

    # Calibrate on a short round, then size one round to about target_seconds.
    _, seconds, _ = run_reference(scene, threads, 1, threads=threads)
    rate = threads * columns / max(seconds[0], 1e-3)
    ppc = max(threads, int(rate * target_seconds / columns) // threads * threads)
    _, seconds, _ = run_reference(scene, ppc, 1, threads=threads)
    photons_per_second = ppc * columns / seconds[0]
    return {
        "value": photons_per_second * events_per_photon,
        "unit": "photon-events/s",
        "cores": threads,
        "kind": "reference",
        "sample": (
            f"one round of {ppc} photons/column ({ppc * columns} photons) of the same scene; "
            f"{seconds[0]:.2f} s transport; events/s = GPU-counted events per photon "
            f"({events_per_photon:.2f}) x CPU photons/s ({photons_per_second:.0f})"
        ),
        "photons_per_second": photons_per_second,
    }


def time_to_amf_se(scene, max_rounds: int = 1024, photons_per_column: int = PHOTONS_PER_COLUMN,
                   name: str = "layered-cloud", proposal=None, directed=None) -> dict:
    """Wall time of simulate(..., convergence=AMFConvergence(0.01, 128, 32, 3)) to converged."""

    import tempfile

    from monki import AMFConvergence, SceneCase, simulate
    from monki.execution import HIP
    from monki.execution.backend import Backend

    class Progress(Backend):
        """Delegates to HIP and prints one line per group of rounds (the run takes minutes)."""

        def __init__(self):
            self.inner = HIP(proposal=proposal, directed=directed)
            self.rounds = 0
            self.start = time.perf_counter()

        def prepare(self, case):
            self.inner.prepare(case)

        @property
        def metadata(self):
            return self.inner.metadata

        @property
        def provenance(self):
            return self.inner.provenance

        def run_batch(self, controls):
            return self.run_batches([controls])[0]

        def run_batches(self, controls):
            batches = self.inner.run_batches(controls)
            self.rounds += len(batches)
            print(f"amf: {self.rounds} rounds, {time.perf_counter() - self.start:.1f} s", file=sys.stderr, flush=True)
            return batches

        def close(self):
            self.inner.close()

    policy = AMFConvergence(0.01, 128, 32, 3)
    case = SceneCase(name, scene, photons_per_column, max_rounds)

    with tempfile.TemporaryDirectory() as directory:
        start = time.perf_counter()
        result = simulate(case, Progress(), Path(directory) / "amf.nc", convergence=policy)
        elapsed = time.perf_counter() - start
        receipt = json.loads((Path(directory) / "amf.run.json").read_text())

    import numpy as np

    amf = result.values["AMF_layer"]
    relative = np.abs(result.values["std_AMF_layer"]) / np.abs(amf)
    finite = relative[np.isfinite(relative)]
    worst = float(finite.max()) if finite.size else None
    quantiles = {"median": float(np.median(finite)), "p99": float(np.quantile(finite, 0.99)),
                 "max": worst, "nonfinite": int(relative.size - finite.size)} if finite.size else None  # fmt: skip
    result = {"seconds": elapsed, "status": receipt["status"], "rounds": receipt["completed_rounds"],
              "policy": "AMFConvergence(0.01, minimum 128, every 32, 3 passes)", "max_relative_se": worst,
              "relative_se": quantiles, "method": receipt.get("method", "plain")}  # fmt: skip

    if "proposal" in receipt:
        result["proposal"] = {key: receipt["proposal"][key] for key in
                              ("training_events", "logging_seconds", "training_seconds", "loss_first", "loss_last")}  # fmt: skip

    if worst and receipt["status"] != "converged":
        # SE falls as rounds^-1/2: rounds and seconds the same run needs to reach 1%.
        needed = receipt["completed_rounds"] * (worst / 0.01) ** 2
        result["projected_rounds"] = needed
        result["projected_seconds"] = elapsed * needed / receipt["completed_rounds"]

    return result


def amf_race(max_rounds: int = 256, photons_per_column: int = PHOTONS_PER_COLUMN) -> dict:
    """Plain vs guided tracer on the full-size 3-D box cloud (tau 20 over 10 x 6 columns,
    layers 2-6 km, Mie, I/Q/U, 3 views): wall time of simulate(..., AMFConvergence(0.01, 128,
    32, 3)) up to max_rounds, the final worst relative AMF SE, and the projection of both to 1%
    (SE ~ rounds^-1/2). The guided time includes logging and training the proposal. A third leg
    runs detector-directed sampling (vMF lobes toward the views, no network)."""

    from monki import configs
    from monki.execution.proposal import DirectedSettings, ProposalSettings

    scene = configs.box_cloud()
    plain = time_to_amf_se(scene, max_rounds, photons_per_column=photons_per_column, name="box-cloud")
    guided = time_to_amf_se(scene, max_rounds, photons_per_column=photons_per_column, name="box-cloud",
                            proposal=ProposalSettings())  # fmt: skip
    directed = time_to_amf_se(scene, max_rounds, photons_per_column=photons_per_column, name="box-cloud",
                              directed=DirectedSettings())  # fmt: skip
    result = {"plain": plain, "guided": guided, "directed": directed,
              "workload": "box cloud 30x20x35, tau 20, 10x6 columns x layers 2-6 km"}  # fmt: skip

    if plain.get("projected_seconds") and guided.get("projected_seconds"):
        result["speedup_to_1pct"] = plain["projected_seconds"] / guided["projected_seconds"]

    if plain.get("relative_se") and guided.get("relative_se"):
        # Per-photon variance ratio at equal photon counts, from the median relative AMF SE.
        result["variance_ratio_median"] = (guided["relative_se"]["median"] / plain["relative_se"]["median"]) ** 2

    if plain.get("projected_seconds") and directed.get("projected_seconds"):
        result["directed_speedup_to_1pct"] = plain["projected_seconds"] / directed["projected_seconds"]

    return result


def batch_rounds(rounds: int = 5, photons_per_column: int = 1_000) -> dict:
    """The 64-scene box-cloud batch (seed 2026: SZA U(0, 70), albedo U(0.02, 0.3), cloud optical
    depth lognormal(15, 0.5), grids 16-30 x 12-20 columns) traced one round per graph replay,
    with block and tally shards from the autotuner."""

    from monki import SceneCase, configs
    from monki.execution import HIPBatch
    from monki.execution.autotune import autotune

    cases = [SceneCase(f"box{i}", scene, photons_per_column, rounds + 1) for i, scene in
             enumerate(configs.box_cloud_batch(64, 2026))]  # fmt: skip
    tuned = autotune(cases, rounds=2)

    with HIPBatch(block=tuned["block"], tally_shards=tuned["tally_shards"]) as batch:
        batch.prepare(cases)
        batch.run_round(1)
        before = batch.counters()
        device = 0.0
        start = time.perf_counter()

        for index in range(2, rounds + 2):
            batch.run_round(index)
            device += batch.kernel_seconds()

        elapsed = time.perf_counter() - start
        counts = delta(batch.counters(), before)

    events = counts["atmospheric_events"] + counts["surface_events"]
    return {
        "workload": "64 box-cloud scenes (seed 2026), I/Q/U, 3 views, layer paths, "
        f"{photons_per_column} photons/column, one graph replay per round",
        "photons_per_round": counts["photons"] // rounds,
        "photon_events_per_s": events / elapsed,
        "ms_per_round": 1e3 * elapsed / rounds,
        "device_ms_per_round": 1e3 * device / rounds,
        "autotune": {key: tuned[key] for key in ("block", "tally_shards", "seconds_per_round", "cached")},
        "autotune_table": tuned["table"],
    }


def spectral_rounds(rounds: int = 2, photons_per_column: int = 16) -> dict:
    """The spectral case: 64 wavelengths (300-500 nm) of one 128 x 128 x 60 stochastic cloud
    field (Kolmogorov spectrum, cover 0.6, mean cloudy optical depth 12, cloud 1-4 km), I/Q/U,
    3 views, layer paths; one graph replay per round, launch shape from the autotuner."""

    from monki import SceneCase, configs
    from monki.execution import HIPBatch
    from monki.execution.autotune import autotune

    cases = [SceneCase(f"spectral-{i:02d}", scene, photons_per_column, 2 * rounds + 2)
             for i, scene in enumerate(configs.spectral_scenes(64))]  # fmt: skip
    tuned = autotune(cases, blocks=(128, 256), shards=(4, 16), rounds=1)

    with HIPBatch(block=tuned["block"], tally_shards=tuned["tally_shards"]) as batch:
        batch.prepare(cases)
        # Rounds copied whole to the host (run_round), then rounds whose layer paths stay on the
        # device (accumulate_round, what simulate_batch uses); one warm-up round each.
        timings = {}

        for label, run, first in (("host", batch.run_round, 1), ("device", batch.accumulate_round, rounds + 2)):
            run(first)
            before = batch.counters()
            device = 0.0
            start = time.perf_counter()

            for index in range(first + 1, first + rounds + 1):
                run(index)
                device += batch.kernel_seconds()

            timings[label] = (time.perf_counter() - start, device, delta(batch.counters(), before))

    elapsed, device, counts = timings["device"]
    events = counts["atmospheric_events"] + counts["surface_events"]
    return {
        "workload": f"spectral 64 x (128 x 128 x 60), {photons_per_column} photons/column, one graph replay per round",
        "photons_per_round": counts["photons"] // rounds,
        "events_per_photon": events / max(counts["photons"], 1),
        "photon_events_per_s": events / elapsed,
        "ms_per_round": 1e3 * elapsed / rounds,
        "device_ms_per_round": 1e3 * device / rounds,
        "host_copy_ms_per_round": 1e3 * timings["host"][0] / rounds,
        "round_path": "layer paths, weights and AMF moments accumulated on the device; Stokes copied per round",
        "autotune": {key: tuned[key] for key in ("block", "tally_shards", "seconds_per_round", "cached")},
    }


COPY_BYTES = 1 << 30
TRAFFIC_FILE = ROOT / "profiles" / "traffic_layered_cloud.json"


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


def summarize_traffic(directory: Path) -> dict:
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
    summary = {
        "hbm_bytes_per_launch": fetch_bytes + write_bytes,
        "fetch_bytes_per_launch": fetch_bytes,
        "write_bytes_per_launch": write_bytes,
        "fetch_factor": fetch_factor,
        "write_factor": write_factor,
        "transport_dispatches": [len(fetch), len(write)],
        "copy_kib": [copy_fetch, copy_write],
        "commands": [
            "rocprofv3 --pmc FETCH_SIZE -d DIR -o transport_fetch -- python3 bench.py --steps 2 --warmup 0 --no-cpu",
            "rocprofv3 --pmc WRITE_SIZE -d DIR -o transport_write -- python3 bench.py --steps 2 --warmup 0 --no-cpu",
            "rocprofv3 --pmc FETCH_SIZE -d DIR -o copy_fetch -- python3 bench.py --copy-calibration",
            "rocprofv3 --pmc WRITE_SIZE -d DIR -o copy_write -- python3 bench.py --copy-calibration",
            "python3 bench.py --summarize-traffic DIR",
        ],
    }
    TRAFFIC_FILE.write_text(json.dumps(summary, indent=1) + "\n")
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
    return "\n".join(lines) + "\n"


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--amf", action="store_true", help="also time AMF convergence at 1%%")
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    parser.add_argument("--copy-calibration", action="store_true", help="run the copy kernel only")
    parser.add_argument("--summarize-traffic", type=Path, metavar="DIR")
    parser.add_argument("--summarize-profile", type=Path, metavar="DB")
    parser.add_argument("--batch", action="store_true", help="also time the 64-scene box-cloud batch")
    parser.add_argument("--spectral", action="store_true", help="also time the 64-wavelength 128x128x60 field")
    parser.add_argument("--guided", type=int, metavar="ROUNDS", default=0,
                        help="also race plain vs guided to 1%% AMF SE on the box cloud (round cap)")
    args = parser.parse_args()

    if args.summarize_profile:
        print(summarize_profile(args.summarize_profile), end="")
        return 0

    if args.copy_calibration:
        copy_calibration()
        return 0

    if args.summarize_traffic:
        print(json.dumps(summarize_traffic(args.summarize_traffic)), flush=True)
        return 0

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU: there is no multi-GPU path (asked for {args.gpus}). "
            "Photon histories would shard by photon range, round or scene, with one reduction of "
            "each round's tallies.",
            file=sys.stderr,
        )
        return 2

    from monki import SceneCase, configs
    from monki.execution import HIP, BatchControls

    scene = configs.layered_cloud()
    backend = HIP()
    backend.prepare(SceneCase("layered-cloud", scene, PHOTONS_PER_COLUMN, args.warmup + args.steps))
    requests = [BatchControls(i, PHOTONS_PER_COLUMN) for i in range(1, args.warmup + args.steps + 1)]

    for request in requests[: args.warmup]:
        backend.run_batch(request)

    before = backend.counters()
    kernel_seconds = 0.0
    start = time.perf_counter()

    for request in requests[args.warmup :]:
        backend.run_batch(request)
        kernel_seconds += backend.kernel_seconds()

    elapsed = time.perf_counter() - start
    counts = delta(backend.counters(), before)
    device = backend.provenance["device"]
    backend.close()

    events = counts["atmospheric_events"] + counts["surface_events"]
    photons = counts["photons"]
    bytes_per_launch = algorithmic_bytes(counts, scene.angles, scene.nmat, scene.nz) / args.steps
    kernel_average = kernel_seconds / args.steps
    achieved = bytes_per_launch / kernel_average / 1e9
    traffic = None

    if TRAFFIC_FILE.is_file():
        traffic = json.loads(TRAFFIC_FILE.read_text()).get("hbm_bytes_per_launch")

    line = {
        "metric": METRIC,
        "value": events / elapsed,
        "unit": "photon-events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "layered-cloud plain tracer: 30x20 columns x 35 layers, water cloud tau 10 "
            "(Mie r_eff 10 um) + HG aerosol + Rayleigh/O3, I/Q/U, 3 views, layer AMF tallies, "
            "10000 photons/column (6e6 photons/round)",
            "photons_per_step": photons // args.steps,
            "events_per_photon": events / photons,
            "kernel_ms_per_step": 1e3 * kernel_average,
            "device": device,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
    }

    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(scene, events / photons)

    if args.amf:
        line["time_to_1pct_amf_se"] = time_to_amf_se(scene)

    if args.guided:
        line["guided_vs_plain"] = amf_race(args.guided)

    if args.batch:
        line["batch_box_clouds"] = batch_rounds()

    if args.spectral:
        line["spectral_field"] = spectral_rounds()

    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
