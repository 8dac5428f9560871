"""Validate and pack prepared transport tables into the float32 layout of ``libmonki_hip``.

Counterpart of the reference's ``python/monki/execution/metal_tables.py:6-133``: a
versioned float32 buffer with a fixed header, then the tables. Differences: several
tabulated species (cumulative extinction fractions per species, gas last and implicit),
explicit offsets in the header instead of a fixed order, and scene constants (sizes,
geometry, shard count) as header words rather than compiled-in defines.

Header (72 float32 words; integers are stored exactly, all < 2**24):

====  =====================================================================
0     layout version (2)
1-9   nx, ny, nz (incl. vacuum), nmat, views, species, paths, uniform_top, shards
10-15 xunit, yunit (km), depol, propagation zenith, solar azimuth, view azimuth (rad)
16-18 view zeniths (rad)
19-25 offsets: zunit, albedo, beta, cumulative fractions, gas albedo,
      height above, optical depth above
26    total buffer length
27    offset of the per-layer uniformity flags (1: extinction and species mix are the same in
      every column of the layer, so the free path crosses it without cell-by-cell steps)
28    transport direction: 0 forward, 1 backward (mc.f90 backward_mode)
29    scattering decomposition: 1 runs calc_radiances_and_scats (no layer paths)
30    detector and result grid: 0 TOA, 1 surface (views point down)
31    result grid on the other level than the detector: 1 (0: at the detector)
32+8s species s: LUT size, albedo, offsets of cosine block (cos, 1-cos, 1+cos),
      CDF, F11, F22, F33, F12; with nmat 4, F34 and F44 follow F12 directly (same length)
64-65 factored extinction (one tabulated species; 0 when absent): offsets of the species'
      extinction per cell and of the gas extinction per layer, with beta = species + gas
      and the species fraction = species / beta exactly in float32
====  =====================================================================

Grids are x-fastest (Fortran order). ``height_above[k]``/``tau_above[k]`` are the thickness
and vertical optical depth of the physical layers above layer k, integrated from the same
float32 values the kernel reads (``metal_tables.py:93-107``); ``uniform_top`` is one above the
highest layer whose extinction or species mix varies across columns.
"""

import numpy as np

HEADER_WORDS = 72
LAYOUT_VERSION = 2
MAX_LAYERS = 64
MAX_SPECIES = 4
MAX_VIEWS = 3
MAX_OFFSET = 2**24


def pack_tables(scene, shards: int = 32, backward: bool = False):
    """Return (float32 buffer, (nmat, nx, ny, nz, views)) for ``monki_hip_create``."""

    scene.validate()
    nx, ny, nz, nmat = scene.nx, scene.ny, scene.nz, scene.nmat
    views = scene.angles
    nspecies = len(scene.species)

    if nz > MAX_LAYERS:
        raise ValueError(f"at most {MAX_LAYERS} layers including vacuum are supported")

    if scene.save_scats and scene.save_path:
        raise ValueError("the scattering decomposition replaces the layer paths: save_path must be False")

    if scene.save_path and nz + nmat > 64:
        raise ValueError("layers (incl. vacuum) plus Stokes components must not exceed 64 with paths")

    if nspecies > MAX_SPECIES:
        raise NotImplementedError(f"at most {MAX_SPECIES} tabulated species are supported")

    if views > MAX_VIEWS:
        raise ValueError(f"at most {MAX_VIEWS} view zenith angles are supported")

    if type(shards) is not int or not 1 <= shards <= 64:
        raise ValueError("shards must be an integer in [1, 64]")

    beta32 = scene.beta.astype(np.float32)
    z32 = scene.zunit.astype(np.float32)
    by_layer = beta32.reshape(nx * ny, nz, order="F")
    varying = scene.varying_layers()
    uniform_top = int(varying[-1] + 1) if varying.size else 0
    height_above = np.zeros(nz)
    tau_above = np.zeros(nz)

    for layer in range(nz - 3, -1, -1):
        height_above[layer] = height_above[layer + 1] + float(z32[layer + 1])
        tau_above[layer] = tau_above[layer + 1] + float(z32[layer + 1]) * float(by_layer[0, layer + 1])

    cumulative = np.cumsum(scene.fractions[:nspecies], axis=0)
    parts = []
    offset = HEADER_WORDS

    def place(array):
        nonlocal offset
        array = np.asarray(array, dtype=np.float64).ravel(order="F")
        start = offset
        parts.append(array)
        offset += array.size
        return start

    header = np.zeros(HEADER_WORDS)
    header[0] = LAYOUT_VERSION
    header[1:10] = [nx, ny, nz, nmat, views, nspecies, int(scene.save_path), uniform_top, shards]
    header[10:16] = [
        scene.xunit,
        scene.yunit,
        scene.depol,
        np.deg2rad(scene.propagation_zenith),
        np.deg2rad(scene.solar_azimuth),
        np.deg2rad(scene.view_azimuth),
    ]
    header[16 : 16 + views] = np.deg2rad(scene.view_zeniths)
    header[19] = place(scene.zunit)
    header[20] = place(scene.albedo)
    header[21] = place(scene.beta)
    header[22] = place(np.moveaxis(cumulative, 0, -1)) if nspecies else offset
    header[23] = place(scene.ssalb_gas)
    header[24] = place(height_above)
    header[25] = place(tau_above)
    uniform = np.ones(nz)
    uniform[varying] = 0.0
    header[27] = place(uniform)
    header[28] = 1.0 if backward else 0.0
    header[29] = 1.0 if scene.save_scats else 0.0
    header[30] = 1.0 if scene.detector == "surface" else 0.0
    header[31] = 1.0 if (scene.register or scene.detector) != scene.detector else 0.0

    if scene.extinction_parts is not None:
        header[64] = place(scene.extinction_parts[0])
        header[65] = place(scene.extinction_parts[1])

    for index, sp in enumerate(scene.species):
        cdf = sp.bound_cdf(nmat) if sp.cdf is None else sp.cdf
        base = 32 + 8 * index
        header[base : base + 2] = [sp.cosines.size, sp.ssalb]
        # Cosine block: cos, then 1 - cos and 1 + cos, which keep float32 precision near the
        # forward and backward peaks where 1 - cos^2 would cancel.
        header[base + 2] = place(np.concatenate([sp.cosines, 1.0 - sp.cosines, 1.0 + sp.cosines]))
        header[base + 3] = place(cdf)
        header[base + 4] = place(sp.f11)
        header[base + 5] = place(sp.f22)
        header[base + 6] = place(sp.f33)
        header[base + 7] = place(sp.f12)

        if nmat == 4:
            place(sp.f34)  # F12 offset + LUT size
            place(sp.f44)  # F12 offset + 2 LUT size

    header[26] = offset

    if offset >= MAX_OFFSET:
        raise ValueError("scene tables exceed the 2**24-word float32 header range")

    packed = np.concatenate([header, *parts]).astype(np.float32)

    if not np.all(np.isfinite(packed)):
        raise ValueError("Transport tables overflow float32 precision")

    for sp in scene.species:
        if np.any(np.diff(sp.cosines.astype(np.float32)) <= 0):
            raise ValueError("Scattering angles are not distinct at float32 precision")

    view_cosines = np.cos(np.deg2rad(np.asarray(scene.view_zeniths)).astype(np.float32))

    if np.any(view_cosines <= 0 if scene.detector == "toa" else view_cosines >= 0):
        raise ValueError("Viewing angles must stay off the horizon in float32")

    assert cumulative.shape == (nspecies, nx, ny, nz) and packed.size == offset
    return packed, (nmat, nx, ny, nz, views)


def tally_count(nmat: int, nx: int, ny: int, nz: int, views: int, paths: bool, scats: bool = False) -> int:
    """Float count of one round's tallies: Stokes, then paths and weights when saved; with the
    scattering decomposition Stokes, Stokes_sing, Stokes_mult, contrib_sing, contrib_mult."""

    count = nmat * nx * ny * views

    if scats:
        return 3 * count + 2 * nmat * nx * ny * nz * views

    if paths:
        count += nx * ny * (nz + 1) * views

    return count


def split_scattering(values: np.ndarray, shape) -> dict:
    """The decomposition arrays of one round's tallies (``tally_count(..., scats=True)``), in
    Fortran order: Stokes_sing and Stokes_mult (nmat, nx, ny, views) and the single- and
    multiple-scattering contributions by event cell (nmat, nx, ny, nz, views), level 0 for
    surface reflections and k for physical layer k (calc_radiances_and_scats,
    mc_routines_cartesian_mod.f90:405-836)."""

    nmat, nx, ny, nz, views = shape
    stokes = nmat * nx * ny * views
    contrib = nmat * nx * ny * nz * views
    part = lambda start, count, dims: values[start : start + count].reshape(dims, order="F").copy()
    return {
        "stokes_sing": part(stokes, stokes, (nmat, nx, ny, views)),
        "stokes_mult": part(2 * stokes, stokes, (nmat, nx, ny, views)),
        "contrib_sing": part(3 * stokes, contrib, (nmat, nx, ny, nz, views)),
        "contrib_mult": part(3 * stokes + contrib, contrib, (nmat, nx, ny, nz, views)),
    }
