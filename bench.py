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

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--amf]
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


def time_to_amf_se(scene, max_rounds: int = 1024) -> dict:
    import tempfile

    from monki import AMFConvergence, SceneCase, simulate
    from monki.execution import HIP

    policy = AMFConvergence(0.01, 128, 32, 3)
    case = SceneCase("layered-cloud", scene, PHOTONS_PER_COLUMN, max_rounds)

    with tempfile.TemporaryDirectory() as directory:
        start = time.perf_counter()
        simulate(case, HIP(), Path(directory) / "amf.nc", convergence=policy)
        elapsed = time.perf_counter() - start
        receipt = json.loads((Path(directory) / "amf.run.json").read_text())

    return {"seconds": elapsed, "status": receipt["status"], "rounds": receipt["completed_rounds"]}


def main() -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--steps", type=int, default=20)
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--amf", action="store_true", help="also time AMF convergence at 1%%")
    parser.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = parser.parse_args()

    if args.gpus != 1:
        print(
            f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}). "
            "Photon histories shard by photon range, round or scene; see DESIGN.md.",
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
    traffic_file = ROOT / "profiles" / "traffic_layered_cloud.json"

    if traffic_file.is_file():
        traffic = json.loads(traffic_file.read_text()).get("hbm_bytes_per_launch")

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

    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
