"""Build libmonki_hip.so for gfx950 in-tree (monki-gpu_amd/lib/).

Run with ``python monki-gpu_amd/build.py``. hipcc cross-compiles without a GPU.
"""

import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent
SOURCES = [HERE / "csrc" / "transport.hip", HERE / "csrc" / "proposal.hip"]
OUTPUT = HERE / "lib" / "libmonki_hip.so"
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Keep SimplifyCFG from sinking stores of different fields of the photon state into one
    # store through a selected pointer: that pins the whole state in scratch memory.
    "-mllvm", "-simplifycfg-sink-common=false",
    "-Wall",
]  # fmt: skip


def build(verbose: bool = False) -> Path:
    OUTPUT.parent.mkdir(parents=True, exist_ok=True)
    temporary = OUTPUT.with_suffix(".so.tmp")
    command = ["hipcc", *FLAGS, "-I", str(ROOT / "include"), *map(str, SOURCES), "-o", str(temporary)]

    if verbose:
        print(" ".join(command), flush=True)

    subprocess.run(command, check=True)
    temporary.replace(OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
