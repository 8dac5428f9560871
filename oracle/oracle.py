"""ctypes binding of the float64 C oracle and a runner for the reference Fortran build.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` as the checker, never by the product package.

* ``Oracle`` wraps ``oracle/libmonki_oracle.so`` (``monki_oracle.c``): the Fortran transport
  restated in float64 C on the product's Philox streams and packed tables.
* ``run_reference`` runs ``oracle/_ref/monki_ref``: the reference's unchanged ``mc.f90`` and
  helper modules built with amdflang, driven by ``ref_driver.f90`` over a transport.bin v1
  export, with its own compiler random numbers.
"""

import ctypes
import os
import subprocess
import tempfile
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "libmonki_oracle.so"
REFERENCE = HERE / "_ref" / "monki_ref"
PROBE = HERE / "_ref" / "helper_probe"
TRACE_WORDS = 8

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_U = ctypes.POINTER(ctypes.c_uint32)
_U64 = ctypes.POINTER(ctypes.c_uint64)


def build() -> None:
    """Compile the C oracle (and the reference build when its sources are present)."""

    subprocess.run(["make", "-s", "-C", str(HERE), "oracle"], check=True)

    if Path("/root/reference/fortran").is_dir():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", "probe"], check=True)


class Oracle:
    def __init__(self, library: Path = LIBRARY):
        if not Path(library).is_file():
            raise FileNotFoundError(f"{library} is missing: run make -C oracle")

        lib = ctypes.CDLL(str(library))
        lib.oracle_run.argtypes = [_F, ctypes.c_size_t, _U, _D, _D, ctypes.c_size_t, _U64]
        lib.oracle_run.restype = ctypes.c_uint
        lib.oracle_tally_count.argtypes = [_F, ctypes.c_size_t]
        lib.oracle_sample.argtypes = [_F, ctypes.c_size_t, ctypes.c_int] + [ctypes.c_double] * 3 + [ctypes.c_uint32, ctypes.c_uint32, _D]
        lib.oracle_sample.restype = ctypes.c_int
        lib.oracle_tally_count.restype = ctypes.c_int
        lib.oracle_philox.argtypes = [_U, _U, _U]
        lib.oracle_rayleigh.argtypes = [ctypes.c_double, ctypes.c_double, _D]
        lib.oracle_rotation.argtypes = [ctypes.c_double] * 5 + [_D]
        lib.oracle_scatter.argtypes = [_D, _D, _D, _D]
        lib.oracle_tabulated.argtypes = [_F, _F, _F, _F, _F, ctypes.c_int, ctypes.c_double, _D]
        lib.oracle_new_direction.argtypes = [ctypes.c_double] * 3 + [_D, _D]
        lib.oracle_rayleigh_f44.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.oracle_rayleigh_f44.restype = ctypes.c_double
        lib.oracle_scatter4.argtypes = [_D, _D, _D, _D]
        lib.oracle_tabulated_v.argtypes = [_F, _F, _F, ctypes.c_int, ctypes.c_double, _D]
        self.lib = lib

    # -- helpers ------------------------------------------------------------------------
    def philox(self, counter, key):
        c = (ctypes.c_uint32 * 4)(*counter)
        k = (ctypes.c_uint32 * 2)(*key)
        out = (ctypes.c_uint32 * 4)()
        self.lib.oracle_philox(c, k, out)
        return list(out)

    def rayleigh(self, depol, cosine):
        out = (ctypes.c_double * 4)()
        self.lib.oracle_rayleigh(depol, cosine, out)
        return np.array(out)

    def rotation(self, cosine, uo, u1, phio, phi1):
        out = (ctypes.c_double * 4)()
        self.lib.oracle_rotation(cosine, uo, u1, phio, phi1, out)
        return np.array(out)

    def scatter(self, elements, rotation, stokes):
        f = (ctypes.c_double * 4)(*elements)
        r = (ctypes.c_double * 4)(*rotation)
        s = (ctypes.c_double * 3)(*stokes)
        out = (ctypes.c_double * 3)()
        self.lib.oracle_scatter(f, r, s, out)
        return np.array(out)

    def rayleigh_f44(self, depol, cosine):
        return self.lib.oracle_rayleigh_f44(depol, cosine)

    def scatter4(self, elements, rotation, stokes):
        """P * S of the full Stokes vector; elements F11, F22, F33, F12, F34, F44 (/4)."""

        f = (ctypes.c_double * 6)(*elements)
        r = (ctypes.c_double * 4)(*rotation)
        s = (ctypes.c_double * 4)(*stokes)
        out = (ctypes.c_double * 4)()
        self.lib.oracle_scatter4(f, r, s, out)
        return np.array(out)

    def tabulated_v(self, species, cosine):
        """F34 / 4 and F44 / 4 of a tabulated species at a cosine (float32 table)."""

        arrays = [np.ascontiguousarray(getattr(species, n), dtype=np.float32) for n in ("cosines", "f34", "f44")]
        out = (ctypes.c_double * 2)()
        self.lib.oracle_tabulated_v(*[a.ctypes.data_as(_F) for a in arrays], arrays[0].size, cosine, out)
        return np.array(out)

    def tabulated(self, species, cosine):
        arrays = [np.ascontiguousarray(getattr(species, n), dtype=np.float32) for n in
                  ("cosines", "f11", "f22", "f33", "f12")]  # fmt: skip
        out = (ctypes.c_double * 4)()
        self.lib.oracle_tabulated(*[a.ctypes.data_as(_F) for a in arrays], arrays[0].size, cosine, out)
        return np.array(out)

    def new_direction(self, cos_in, cos_sca, sigma):
        co, dphi = ctypes.c_double(), ctypes.c_double()
        self.lib.oracle_new_direction(cos_in, cos_sca, sigma, ctypes.byref(co), ctypes.byref(dphi))
        return co.value, dphi.value

    # -- transport ----------------------------------------------------------------------
    def sample(self, packed, species: int, q: float, u: float, cos_in: float, count: int, seed: int = 1):
        """Sampled (cos_sca, sigma, cos_out, dphi) rows of the oracle's rejection sampler."""

        packed = np.ascontiguousarray(packed, dtype=np.float32)
        out = np.zeros((count, 4), dtype=np.float64)
        status = self.lib.oracle_sample(
            packed.ctypes.data_as(_F), packed.size, species, q, u, cos_in, seed, count,
            out.ctypes.data_as(_D),
        )  # fmt: skip

        if status:
            raise ValueError("oracle sampling failed")

        return out

    def run(self, packed, controls, records: int = 0):
        """Run photons with the product's control words; return (tallies, records, counters, flag)."""

        packed = np.ascontiguousarray(packed, dtype=np.float32)
        count = self.lib.oracle_tally_count(packed.ctypes.data_as(_F), packed.size)

        if count < 0:
            raise ValueError("oracle rejected the packed tables")

        rounds = int(controls[7])
        tallies = np.zeros((rounds, count), dtype=np.float64)
        rec = np.zeros((max(records, 1), TRACE_WORDS), dtype=np.float64)
        counters = np.zeros(8, dtype=np.uint64)
        words = (ctypes.c_uint32 * 8)(*[int(c) for c in controls])
        flag = self.lib.oracle_run(
            packed.ctypes.data_as(_F), packed.size, words, tallies.ctypes.data_as(_D),
            rec.ctypes.data_as(_D), records, counters.ctypes.data_as(_U64),
        )  # fmt: skip
        return tallies, rec[:records], counters, int(flag)


def run_reference(scene, photons_per_column: int, rounds: int, seed: int = 23274, threads: int = 1,
                  executable: Path = REFERENCE, backward: bool = False, save_scats: bool = False):
    """Run the reference Fortran transport (backward_mode = ``backward``); return (per-round
    tallies, seconds per round, wall seconds). With ``save_scats`` (one thread) each round's
    row is [Stokes | Stokes_sing | Stokes_mult | contrib | contrib_sing | contrib_mult], the
    3-D arrays (nmat, nx, ny, 0:nz, nangles) in Fortran order."""

    if not Path(executable).is_file():
        raise FileNotFoundError(f"{executable} is missing: run make -C oracle ref")

    with tempfile.TemporaryDirectory(prefix="monki-ref-") as directory:
        tables = scene.write_transport_bin(Path(directory) / "transport.bin", backward)
        output = Path(directory) / "rounds.bin"
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_DYNAMIC="FALSE",
                   OMP_STACKSIZE="128M", GOMP_STACKSIZE="128M")  # fmt: skip
        start = time.perf_counter()
        result = subprocess.run(
            [str(executable), str(tables), str(photons_per_column), str(rounds), str(seed), str(output),
             "1" if save_scats else "0"],
            env=env, capture_output=True, text=True,
        )  # fmt: skip
        elapsed = time.perf_counter() - start

        if result.returncode:
            raise RuntimeError(f"reference transport failed: {result.stdout[-2000:]}{result.stderr[-2000:]}")

        seconds = [float(line.split()[-1]) for line in result.stdout.splitlines()
                   if line.startswith("MONKI_REF_ROUND")]  # fmt: skip
        data = np.fromfile(output, dtype=np.float64).reshape(rounds, -1)

    return data, seconds, elapsed


def fortran_scats_row(row, shape):
    """The reference driver's decomposition row [Stokes | sing | mult | contrib | contrib_sing |
    contrib_mult] (3-D arrays over levels 0..nz) in the product's order [Stokes | sing | mult |
    contrib_sing | contrib_mult] (levels 0..nz-1; level nz, the vacuum layer, holds no events)."""

    nmat, nx, ny, nz, views = shape
    stokes = nmat * nx * ny * views
    full = nmat * nx * ny * (nz + 1) * views
    cut = lambda a: a.reshape(nmat, nx, ny, nz + 1, views, order="F")[:, :, :, :nz].ravel(order="F")
    head = row[: 3 * stokes]
    single = row[3 * stokes + full : 3 * stokes + 2 * full]
    multiple = row[3 * stokes + 2 * full :]
    if np.any(row[3 * stokes : 3 * stokes + full].reshape(nmat, nx, ny, nz + 1, views, order="F")[:, :, :, nz] != 0):
        raise ValueError("contribution recorded in the vacuum layer")
    return np.concatenate([head, cut(single), cut(multiple)])
