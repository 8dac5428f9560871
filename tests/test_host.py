"""Host logic without a GPU: the reference's unit-test semantics plus the C-ABI boundary.

Re-expresses tests/unit/{test_convergence,test_execution,test_inputs,test_validation}.py of the
reference for this package, and checks table packing, the transport.bin v1 export, NetCDF
products and that libmonki_hip.so exports every symbol include/monki_hip.h declares.
"""

import ctypes
import re
from dataclasses import replace
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


# -- convergence (tests/unit/test_convergence.py) ----------------------------------------


def test_online_amf_error_matches_explicit_correlated_ratio():
    from monki.convergence import RatioMoments
    from monki.execution.tallies import BatchControls, RawTallies

    rng = np.random.default_rng(123)
    weights = rng.uniform(1, 4, (100, 1, 1, 2))
    paths = weights[:, :, :, None, :] * rng.uniform(1, 5, (100, 1, 1, 4, 2))
    moments = RatioMoments()

    for index in range(100):
        moments.add(RawTallies(BatchControls(index + 1, 100), np.zeros((3, 1, 1, 2)), paths[index], weights[index]))

    thickness = np.array([0.5, 2.0, 5.0])
    actual, error = moments.estimate(thickness)
    ratio = paths[:, :, :, :-1, :].mean(axis=0) / weights.mean(axis=0)[:, :, None, :]
    residual = paths[:, :, :, :-1, :] - ratio * weights[:, :, :, None, :]
    expected = residual.std(axis=0, ddof=1) / 10 / weights.mean(axis=0)[:, :, None, :]
    np.testing.assert_allclose(actual, ratio / thickness[None, None, :, None])
    np.testing.assert_allclose(error, expected / thickness[None, None, :, None], rtol=1e-12)


@pytest.mark.parametrize(
    "settings",
    [{"relative_error": 0}, {"relative_error": np.nan}, {"minimum_rounds": 31},
     {"check_every": 0}, {"consecutive_checks": True}],
)  # fmt: skip
def test_invalid_stopping_policy(settings):
    from monki import AMFConvergence

    with pytest.raises(ValueError):
        AMFConvergence(**settings)


class _Case:
    photons_per_column = 100

    def __init__(self, rounds):
        self.rounds = rounds


def test_late_outlier_resets_precision_streak():
    from monki import AMFConvergence
    from monki.convergence import PrecisionStopping
    from monki.execution.tallies import RawTallies

    class Synthetic:
        def run_batches(self, requests):
            return [
                RawTallies(r, np.ones((3, 1, 1, 1)), np.full((1, 1, 2, 1), 100.0 if r.round_index == 48 else 1.0), np.ones((1, 1, 1)))
                for r in requests
            ]

    stopping = PrecisionStopping(AMFConvergence(minimum_rounds=32, check_every=16), np.array([1.0]))
    assert len(list(stopping.schedule(Synthetic(), _Case(64)))) == 64
    assert not stopping.converged
    assert stopping.history[0]["streak"] == 1
    assert stopping.history[1]["streak"] == 0


def test_zero_amf_cannot_satisfy_relative_precision():
    from monki import AMFConvergence
    from monki.convergence import PrecisionStopping
    from monki.execution.tallies import RawTallies

    class Synthetic:
        def run_batches(self, requests):
            return [RawTallies(r, np.ones((3, 1, 1, 1)), np.zeros((1, 1, 2, 1)), np.ones((1, 1, 1))) for r in requests]

    stopping = PrecisionStopping(AMFConvergence(minimum_rounds=32, consecutive_checks=1), np.array([1.0]))
    list(stopping.schedule(Synthetic(), _Case(32)))
    assert not stopping.converged
    assert stopping.history[-1]["max_relative_se"] is None


# -- execution (tests/unit/test_execution.py) --------------------------------------------


def test_backend_is_abstract():
    from monki.execution import Backend

    with pytest.raises(TypeError, match="abstract"):
        Backend()


@pytest.mark.parametrize("missing", ["prepare", "run_batch", "close"])
def test_every_lifecycle_method_is_required(missing):
    from monki.execution import Backend

    methods = {"prepare": lambda self, s: None, "run_batch": lambda self, c: object(), "close": lambda self: None}
    del methods[missing]

    with pytest.raises(TypeError, match=missing):
        type("Incomplete", (Backend,), methods)()


def test_context_manager_closes_on_success_and_failure():
    from monki.execution import Backend

    class Recording(Backend):
        def __init__(self):
            self.events = []

        def prepare(self, scene):
            self.events.append(("prepare", scene))

        def run_batch(self, controls):
            self.events.append(("batch", controls))
            return controls

        def close(self):
            self.events.append(("close", None))

    backend = Recording()

    with backend as strategy:
        strategy.prepare(1)
        assert strategy.run_batch(2) == 2

    assert backend.events == [("prepare", 1), ("batch", 2), ("close", None)]

    with pytest.raises(RuntimeError, match="failed"), backend:
        raise RuntimeError("simulation failed")

    assert backend.events.count(("close", None)) == 2


@pytest.mark.parametrize(
    "settings",
    [{"device": -1}, {"device": True}, {"seed": -1}, {"seed": True}, {"seed": 2**32},
     {"photons_per_dispatch": 0}, {"tally_shards": 0}, {"tally_shards": 65},
     {"tally_shards": True}, {"profile": "yes"}, {"max_steps": 0}, {"max_steps": 100_001},
     {"save_path": "yes"}, {"precision": 16}],
)  # fmt: skip
def test_invalid_hip_settings_fail_before_gpu_access(settings):
    from monki.execution import HIP

    with pytest.raises(ValueError):
        HIP(**settings)


def test_hip_does_not_accept_cloud_settings():
    from monki.execution import HIP

    with pytest.raises(TypeError):
        HIP(provider="aws")


def test_missing_library_fails_loudly(tmp_path):
    from monki import SceneCase, configs
    from monki.execution import HIP

    backend = HIP(library=tmp_path / "missing.so")

    with pytest.raises(FileNotFoundError, match="build"):
        backend.prepare(SceneCase("x", configs.clear_sky(nx=2, ny=2), 10, 1))

    assert backend._handle is None


def test_stale_library_is_rejected_before_native_call(monkeypatch):
    from monki import SceneCase, configs
    from monki.execution import HIP

    class Stale:
        monki_hip_abi_version = staticmethod(lambda: 0)

    monkeypatch.setattr(ctypes, "CDLL", lambda path: Stale())

    with HIP() as backend, pytest.raises(RuntimeError, match="Rebuild"):
        backend.prepare(SceneCase("x", configs.clear_sky(nx=2, ny=2), 10, 1))


def test_reference_case_errors_precede_device_access():
    """Fixture checks of the reference's CPU preparer (cpu.py:82-101) raise before the HIP
    library or a device is touched (this host has neither a GPU nor needs one here)."""

    from monki import ReferenceCase
    from monki.execution import HIP

    clear = "testrun/MONKI_MLS_cloudfree_340_340_test.in"
    box = "testrun/MONKI_MLS_2boxclouds_340_340_test.in"

    with HIP() as backend:
        with pytest.raises(ValueError, match="layer AMF setting"):
            backend.prepare(ReferenceCase("box", box))

        with pytest.raises(ValueError, match="3 viewing angles"):
            backend.prepare(ReferenceCase("clear", clear, view_zeniths=(0.0, 10.0, 20.0, 30.0)))

        with pytest.raises(FileNotFoundError):
            backend.prepare(ReferenceCase("clear", "testrun/missing.in"))

        with pytest.raises(TypeError):
            backend.prepare("clear")

        assert backend.metadata is None

    # Backward fixtures reach the tables with the flag set (the session then needs a device).
    from monki.preparation import prepare_case

    scene_case, _, _ = prepare_case(ReferenceCase("clear", clear, backward=True))
    assert scene_case.backward is True
    with pytest.raises(ValueError, match="backward"):
        ReferenceCase("clear", clear, backward=1)


# -- inputs (tests/unit/test_inputs.py) ---------------------------------------------------


@pytest.mark.parametrize(
    "changes",
    [{"name": "bad name"}, {"input_file": "/abs.in"}, {"input_file": "../x.in"}, {"photons_per_column": 0},
     {"rounds": 2**31}, {"surface_albedo": 1.5}, {"view_zeniths": ()}, {"solar_zenith": 90},
     {"view_azimuth": 360}, {"backward": 1}],
)  # fmt: skip
def test_reference_case_validation(changes):
    from monki import ReferenceCase

    with pytest.raises(ValueError):
        replace(ReferenceCase("clear", "testrun/clear.in"), **changes)


def test_batch_controls_are_positive_32_bit():
    from monki.execution import BatchControls

    for bad in ((0, 1), (1, 0), (1, 2**31), (True, 1)):
        with pytest.raises(ValueError):
            BatchControls(*bad)


# -- scene tables ---------------------------------------------------------------------------


def test_scene_validation_rejects_inconsistent_tables():
    from monki import configs

    scene = configs.clear_sky(nx=2, ny=2)
    scene.validate()
    broken = replace(scene, fractions=scene.fractions * 0.5)

    with pytest.raises(ValueError, match="fractions"):
        broken.validate()

    replace(scene, nmat=4).validate()  # the full Stokes vector is supported

    with pytest.raises(ValueError, match="nmat"):
        replace(scene, nmat=2).validate()

    with pytest.raises(ValueError, match="zenith"):
        replace(scene, view_zeniths=(0.0, 90.0)).validate()


def test_transport_bin_round_trip_and_layout():
    from monki import configs
    from monki.scene import TransportScene

    scene = configs.layered_cloud(nx=3, ny=2)
    raw = scene.transport_bin()
    # worker.f90:1102-1120 header: version, nx, ny, nz, nmat, angles, species, source,
    # detector, register, save_path, backward.
    assert list(raw[:12]) == [1, 3, 2, 35, 3, 3, 2, 35, 35, 35, 1, 0]
    assert raw[15] == 140.0  # propagation zenith = 180 - SZA (worker.f90:789)
    back = TransportScene.from_transport_bin(raw)
    np.testing.assert_array_equal(back.beta, scene.beta)
    np.testing.assert_array_equal(back.fractions, scene.fractions)
    np.testing.assert_array_equal(back.species[0].f11, scene.species[0].f11)
    np.testing.assert_allclose(back.species[1].cdf, scene.species[1].bound_cdf(3))


def test_full_stokes_tables_carry_f34_and_f44():
    """nmat 4: the export appends F34 and F44 per species (read back exactly), and the packed
    float32 tables place them right after F12, where the device reads them."""

    from monki import configs
    from monki.execution.hip_tables import pack_tables, tally_count
    from monki.scene import TransportScene

    scene = configs.box_cloud(nx=4, ny=3, nmat=4, box=(2, 1), offset=(1, 1))
    cloud = scene.species[0]
    assert np.any(cloud.f34 != 0) and not np.array_equal(cloud.f44, cloud.f33 * 0)
    back = TransportScene.from_transport_bin(scene.transport_bin())
    for a, b in zip(back.species, scene.species):
        np.testing.assert_array_equal(a.f34, b.f34)
        np.testing.assert_array_equal(a.f44, b.f44)
    assert back.nmat == 4
    three = replace(scene, nmat=3)
    assert scene.transport_bin().size == three.transport_bin().size + sum(2 * sp.cosines.size for sp in scene.species)

    packed, shape = pack_tables(scene, shards=2)
    assert shape[0] == 4 and packed[4] == 4
    for index, sp in enumerate(scene.species):
        n, f12 = int(packed[32 + 8 * index]), int(packed[32 + 8 * index + 7])
        np.testing.assert_array_equal(packed[f12 + n : f12 + 2 * n], sp.f34.astype(np.float32))
        np.testing.assert_array_equal(packed[f12 + 2 * n : f12 + 3 * n], sp.f44.astype(np.float32))
    assert tally_count(4, 4, 3, scene.nz, 3, True) == 4 * 4 * 3 * 3 + 4 * 3 * scene.nz * 3 + 4 * 3 * 3


def test_steered_tracers_and_batches_reject_full_stokes_before_device():
    from monki import SceneCase, configs
    from monki.execution import HIP, DirectedSettings, HIPBatch

    case = SceneCase("iquv", configs.clear_sky(nx=2, ny=2, nmat=4), 10, 2)

    with pytest.raises(NotImplementedError, match="nmat"):
        HIP(directed=DirectedSettings()).prepare(case)

    with pytest.raises(NotImplementedError, match="nmat"):
        HIPBatch().prepare([case])


def test_packed_tables_uniform_top_and_prefixes():
    from monki import configs
    from monki.execution.hip_tables import pack_tables

    box = configs.box_cloud()
    packed, (nmat, nx, ny, nz, views) = pack_tables(box)
    # The box cloud occupies layers 3-6: the detector walks explicitly below layer 7.
    assert packed[8] == 7
    clear, _ = pack_tables(configs.clear_sky())
    assert clear[8] == 0
    height = clear[int(clear[24]) : int(clear[24]) + nz]
    tau = clear[int(clear[25]) : int(clear[25]) + nz]
    # Above the last physical layer nothing is left; at layer 0 the whole column above it.
    assert height[nz - 2] == 0 and tau[nz - 2] == 0
    zunit = configs.clear_sky().zunit
    assert height[0] == pytest.approx(zunit[1 : nz - 1].sum(), rel=1e-6)
    # Moving a varying layer to the last physical layer disables the shortcut everywhere.
    scene = configs.box_cloud()
    scene.beta[0, 0, nz - 2] += 0.01
    changed, _ = pack_tables(scene)
    assert changed[8] == nz - 1


def test_packed_tables_reject_oversized_scenes():
    from monki import configs
    from monki.execution.hip_tables import pack_tables

    with pytest.raises(ValueError, match="shards"):
        pack_tables(configs.clear_sky(nx=2, ny=2), shards=65)

    # Layer paths and Stokes components share one 64-lane tally record.
    grid = lambda physical, **options: configs._assemble(
        "layers", 2, 2, np.linspace(0.0, 60.0, physical + 1), [], [], 3, 40.0, 0.05, **options
    )
    assert pack_tables(grid(60))[1] == (3, 2, 2, 61, 3)
    with pytest.raises(ValueError, match="exceed 64"):
        pack_tables(grid(61))
    assert pack_tables(grid(63, save_path=False))[1] == (3, 2, 2, 64, 3)
    with pytest.raises(ValueError, match="at most 64 layers"):
        pack_tables(grid(64, save_path=False))

    aerosol = configs.aerosol()
    species = lambda count: configs._assemble(
        "species", 2, 2, configs.LEVELS_KM, [np.full(len(configs.LEVELS_KM) - 1, 0.01)] * count,
        [replace(aerosol, name=f"aerosol{index}") for index in range(count)], 3, 40.0, 0.05,
    )  # fmt: skip
    assert pack_tables(species(4))[0][6] == 4
    with pytest.raises(NotImplementedError, match="at most 4 tabulated species"):
        pack_tables(species(5))


# -- validation and the NetCDF product (tests/unit/test_validation.py) ---------------------


@pytest.fixture
def result():
    from monki import SimulationResult

    values = {
        "radiance": np.array([[[[[0.5], [0.1], [-0.05]]]]]),
        "std_radiance": np.full((1, 1, 1, 3, 1), 0.001),
        "DOP": np.array([[[[np.hypot(0.1, 0.05) / 0.5]]]]),
        "wav": np.array([340.0]),
        "theta": np.array([30.0]),
        "n_rounds": np.array([32]),
        "n_photons_per_column": np.array([2000]),
    }
    dims = {
        "radiance": ("nvza", "ny", "nx", "nmat", "nwavs"),
        "std_radiance": ("nvza", "ny", "nx", "nmat", "nwavs"),
        "DOP": ("nvza", "ny", "nx", "nwavs"),
        "wav": ("nwavs",),
        "theta": ("nvza",),
        "n_rounds": ("single",),
        "n_photons_per_column": ("single",),
    }
    return SimulationResult(Path("synthetic.nc"), values, dims, {})


def changed(result, name, value):
    return replace(result, values={**result.values, name: value})


def test_identical_products_pass_both_modes(result):
    from agreement import agree, invariant_failures

    assert invariant_failures(result) == []
    assert agree(result, result, statistical=False).passed
    assert agree(result, result).passed


def test_wrong_but_physical_radiance_fails(result):
    from agreement import agree, invariant_failures

    candidate = changed(result, "radiance", result.radiance * 1.5)
    assert invariant_failures(candidate) == []
    assert not agree(result, candidate).passed


def test_dop_and_units_mismatches_fail(result):
    from agreement import agree, invariant_failures

    assert "DOP does not follow from the Stokes vector" in invariant_failures(changed(result, "DOP", result.values["DOP"] * 2))
    other = replace(result, attributes={"radiance": {"units": "W"}})
    assert not agree(result, other).passed


def test_combined_se_statistics_excludes_zero_uncertainty():
    from agreement import z_statistics

    stats = z_statistics([1.0, 2.0, 0.0], [0.1, 0.1, 0.0], [1.1, 2.0, 0.0], [0.1, 0.1, 0.0])
    assert stats["count"] == 2
    assert stats["max_abs_z"] == pytest.approx(0.1 / np.hypot(0.1, 0.1))


def test_netcdf_product_round_trip_and_exclusive_create(tmp_path, result):
    from monki import SimulationResult

    product = replace(result, path=tmp_path / "out.nc", attributes={"radiance": {"units": "-"}})
    product.save()
    loaded = SimulationResult.load(tmp_path / "out.nc")
    np.testing.assert_array_equal(loaded.radiance, product.radiance)
    assert loaded.dimensions["radiance"] == product.dimensions["radiance"]
    assert loaded.attributes["radiance"]["units"] == "-"

    with pytest.raises(FileExistsError):
        product.save()


def test_simulate_with_a_recording_backend(tmp_path):
    """simulate() statistics on synthetic tallies: mean, SE with 999 through round 30, AMF."""

    from monki import SceneCase, configs, simulate
    from monki.execution import Backend
    from monki.execution.tallies import RawTallies
    from monki.metadata import scene_metadata

    scene = configs.clear_sky(nx=2, ny=1, nmat=3)
    rng = np.random.default_rng(5)

    class Synthetic(Backend):
        def prepare(self, case):
            self.metadata = scene_metadata(case.scene, case.photons_per_column)
            self.provenance = {"backend": "synthetic"}

        def run_batch(self, controls):
            stokes = np.abs(rng.normal(10, 1, (3, 2, 1, 3))) * [[[[1]]], [[[-0.1]]], [[[0.01]]]]
            paths = np.abs(rng.normal(5, 0.5, (2, 1, scene.nz, 3)))
            return RawTallies(controls, stokes, paths, np.full((2, 1, 3), 2.0))

        def close(self):
            pass

    case = SceneCase("synthetic", scene, 100, 40)
    out = simulate(case, Synthetic(), tmp_path / "s.nc")
    assert out.radiance.shape == (3, 1, 2, 3, 1)
    assert np.all(out.standard_error < 999) and np.all(out.standard_error > 0)
    assert (tmp_path / "s.run.json").is_file()

    with pytest.raises(FileExistsError):
        simulate(case, Synthetic(), tmp_path / "s.nc")

    short = simulate(replace(case, rounds=2), Synthetic(), tmp_path / "short.nc")
    assert np.all(short.standard_error == 999)


# -- the C-ABI ------------------------------------------------------------------------------


def declared_symbols():
    header = (ROOT / "include" / "monki_hip.h").read_text()
    return sorted(set(re.findall(r"\b(monki_hip_\w+)\s*\(", header)))


def test_library_exports_every_declared_symbol():
    from monki.execution.hip import LIBRARY

    if not LIBRARY.is_file():
        pytest.fail("libmonki_hip.so is not built (python monki-gpu_amd/build.py)")

    native = ctypes.CDLL(str(LIBRARY))
    symbols = declared_symbols()
    assert len(symbols) >= 12

    for name in symbols:
        assert hasattr(native, name), name

    from monki.execution.hip import ABI_VERSION

    assert native.monki_hip_abi_version() == ABI_VERSION == 2


def test_library_binding_declares_every_symbol():
    from monki.execution.hip import load_library

    native = load_library()

    for name in declared_symbols():
        assert getattr(native, name) is not None


def test_spectral_cloud_field_statistics():
    """The stochastic field: Kolmogorov 1-D spectrum slope -5/3, cloud cover 0.6, mean cloudy
    optical depth 12; the spectral scenes scale Rayleigh as lambda^-4.08."""

    from monki import configs

    field = configs.kolmogorov_field(256, 3)
    power = np.abs(np.fft.fft2(field)) ** 2
    k = np.hypot(*np.meshgrid(np.fft.fftfreq(256) * 256, np.fft.fftfreq(256) * 256, indexing="ij"))
    shells = np.arange(4, 64)
    spectrum = np.array([power[(k >= a) & (k < a + 1)].sum() for a in shells])
    slope = np.polyfit(np.log(shells + 0.5), np.log(spectrum), 1)[0]
    assert slope == pytest.approx(-5 / 3, abs=0.15)
    tau = configs.stochastic_cloud_field(128, seed=7)
    assert (tau > 0).mean() == pytest.approx(0.6, abs=2e-3)
    assert tau[tau > 0].mean() == pytest.approx(12.0)
    short, long = (configs.spectral_cloud_field(w, tau[:8, :8]) for w in (300.0, 500.0))
    assert short.nz == 60 and len(configs.SPECTRAL_LEVELS_KM) == 60
    # The tables hold float32 values (factored extinction): the gas column to float32 precision.
    gas = lambda scene: (scene.beta[0, 0, :-1] * scene.zunit[:-1] * scene.fractions[-1, 0, 0, :-1]).sum()  # noqa: E731
    rayleigh = lambda scene: gas(scene) - 0.01 * np.exp(-(scene.wavelength_nm - 340) / 12)  # noqa: E731
    assert rayleigh(short) / rayleigh(long) == pytest.approx((300 / 500) ** -4.08, rel=1e-6)
    # Every wavelength carries the same cloud extinction; tables and parts agree in float32.
    assert np.array_equal(short.extinction_parts[0], long.extinction_parts[0])
    for scene in (short, long):
        scene.check_extinction_parts()


def test_factored_extinction_must_reproduce_the_tables():
    """The factored extinction is refused unless it gives back the float32 tables bit for bit."""

    from monki import configs

    scene = configs.spectral_cloud_field(400.0, configs.stochastic_cloud_field(16, seed=3))
    species, gas = scene.extinction_parts
    scene.validate()
    nudged = gas.copy()
    nudged[5] = np.nextafter(nudged[5], np.float32(1))
    scene.extinction_parts = (species, nudged)

    with pytest.raises(ValueError, match="reproduce the extinction table"):
        scene.validate()

    scene.extinction_parts = (species.astype(np.float64), gas)

    with pytest.raises(ValueError, match="float32"):
        scene.validate()


def test_directed_settings_validate_and_build_species_masks():
    from monki.execution import DirectedSettings

    assert DirectedSettings(guided_kinds=("tabulated",)).species_mask() == 0xFFFFFFFE
    assert DirectedSettings(guided_kinds=("gas", 1)).species_mask() == 0b101
    assert DirectedSettings().species_mask() == 0xFFFFFFFF
    assert DirectedSettings().max_order == 2**32 - 1 and DirectedSettings().split_above > 1

    for bad in ({"alpha": 0.0}, {"alpha": 1.5}, {"kappa": 0.5}, {"kappa": 2e6}, {"max_order": 0}, {"max_order": 1.0},
                {"lobes": "cone"}, {"surface": 1}, {"split_above": 0.5}, {"roulette_weight": 3.0},
                {"roulette_below": 0.2, "roulette_weight": 0.1}):  # fmt: skip
        with pytest.raises(ValueError):
            DirectedSettings(**bad)


@pytest.mark.parametrize("settings", ["DirectedSettings", "ProposalSettings"])
def test_unbounded_steering_of_tabulated_species_needs_the_weight_window(settings):
    """Steering Mie events of every order without the weight window is refused: the product of
    the steering factors over a thick cloud's tens of events leaves the standard error
    meaningless. Bounded orders, gas only, or the window make it acceptable."""

    from monki.execution import DirectedSettings, ProposalSettings

    kind = {"DirectedSettings": DirectedSettings, "ProposalSettings": ProposalSettings}[settings]

    with pytest.raises(ValueError, match="weight window"):
        kind(split_above=0.0)

    kind(split_above=0.0, max_order=4)
    kind(split_above=0.0, guided_kinds=("gas",))
    kind()


def test_network_orders_setting():
    """network_orders: a positive integer, 2 by default (the orders at which the network's
    mixture is part of the guided proposal)."""

    from monki.execution import ProposalSettings

    assert ProposalSettings().network_orders == 2
    ProposalSettings(network_orders=1)
    ProposalSettings(network_orders=2**32 - 1)

    for bad in (0, -1, 2.0, 2**32):
        with pytest.raises(ValueError, match="network_orders"):
            ProposalSettings(network_orders=bad)


def test_path_target_settings_and_layer_weights():
    """path_target is a share in [0, 1]; the layer weights are 1 / (the layer's path tally summed
    over pixels and views), read from the documented output order, 0 for layers without path."""

    from monki.execution import ProposalSettings
    from monki.execution.hip_tables import tally_count
    from monki.execution.proposal import path_layer_weights

    assert ProposalSettings().path_target == 0.0
    ProposalSettings(path_target=1.0)

    for bad in (-0.1, 1.5):
        with pytest.raises(ValueError, match="path_target"):
            ProposalSettings(path_target=bad)

    nmat, nx, ny, nz, na = 3, 4, 2, 5, 3
    count = tally_count(nmat, nx, ny, nz, na, True)
    tallies = np.zeros(count)
    paths = np.zeros((nx, ny, nz, na))
    paths[:, :, 0, :] = 0.25  # layer 0: total 0.25 x 24
    paths[1, 0, 2, 1] = 3.0  # layer 2: one cell
    start = nmat * nx * ny * na
    tallies[start:start + paths.size] = paths.ravel(order="F")
    tallies[start + paths.size:] = 7.0  # weights, not read
    weights = path_layer_weights(tallies, (nmat, nx, ny, nz, na))
    np.testing.assert_allclose(weights, [1 / 6.0, 0.0, 1 / 3.0, 0.0, 0.0])


def test_decomposition_settings_layout_and_product_variables():
    """save_scats: exclusive with layer paths, matched against the fixture line, tallies split
    in the documented order, and the Report_ncdf variables in the metadata."""

    from monki import ReferenceCase, configs
    from monki.execution import HIP
    from monki.execution.hip_tables import pack_tables, split_scattering, tally_count
    from monki.metadata import scene_metadata
    from monki.preparation import prepare_case

    with pytest.raises(ValueError, match="exclusive"):
        HIP(save_scats=True)

    with pytest.raises(ValueError, match="scattering-contribution"):
        prepare_case(ReferenceCase("clear", "testrun/MONKI_MLS_cloudfree_340_340_test.in"), save_scats=True)

    scene = replace(configs.box_cloud(nx=3, ny=2, nmat=3, box=(1, 1), offset=(1, 0)), save_path=False, save_scats=True)
    packed, shape = pack_tables(scene, shards=1)
    assert packed[29] == 1 and packed[7] == 0
    nmat, nx, ny, nz, views = shape
    count = tally_count(nmat, nx, ny, nz, views, False, True)
    assert count == 3 * nmat * nx * ny * views + 2 * nmat * nx * ny * nz * views
    values = np.arange(count, dtype=np.float64)
    parts = split_scattering(values, shape)
    stokes = nmat * nx * ny * views
    assert parts["stokes_sing"].ravel(order="F")[0] == stokes
    assert parts["contrib_mult"].shape == (nmat, nx, ny, nz, views)
    assert parts["contrib_mult"].ravel(order="F")[-1] == count - 1

    with pytest.raises(ValueError, match="save_path"):
        pack_tables(replace(scene, save_path=True))

    metadata = scene_metadata(scene, 10)
    assert metadata.dimensions["contrib_radiance_3D_sing"] == ("nvza", "nz_withsurf", "ny", "nx", "nmat", "nwavs")
    assert metadata.values["radiance_mult"].shape == (views, ny, nx, nmat, 1)
    assert metadata.attributes["radiance_sing"]["long_name"] == "TOA radiance single scattering"
