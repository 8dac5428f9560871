"""GPU tests of the batched launch: scenes traced together equal the same scenes traced alone
(same Philox streams), the captured graph replays equal direct launches, simulate_batch equals
simulate per case, and the autotuner returns a measured configuration."""

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def scenes():
    from monki import configs

    return [
        configs.box_cloud(nx=8, ny=6, nmat=3, box=(3, 2), offset=(2, 1)),
        configs.clear_sky(nx=6, ny=4, nmat=1),
        configs.layered_cloud(nx=3, ny=2, nmat=3),
        configs.box_cloud(nx=5, ny=7, nmat=3, solar_zenith=65.0, albedo=0.2, optical_depth=9.0, box=(2, 3)),
    ]


def cases(rounds=3, ppc=150):
    from monki import SceneCase

    return [SceneCase(f"scene{i}", scene, ppc + 25 * i, rounds) for i, scene in enumerate(scenes())]


def assert_tallies_equal(a, b):
    # Identical histories; float32 bank sums in another order (tolerance of the reference's
    # dispatch and shard invariance tests, test_metal.py:133-140).
    np.testing.assert_allclose(a.stokes, b.stokes, rtol=2e-5, atol=2e-5)

    if a.weighted_paths is not None:
        np.testing.assert_allclose(a.weighted_paths, b.weighted_paths, rtol=2e-5, atol=2e-5)
        np.testing.assert_allclose(a.weights, b.weights, rtol=2e-5, atol=2e-5)


@pytest.mark.parametrize("block", [64, 256])
def test_batched_round_equals_single_scene_sessions(block):
    from monki.execution import HIP, HIPBatch, BatchControls

    batch_cases = cases()

    with HIPBatch(block=block, tally_shards=8) as batch:
        batch.prepare(batch_cases)
        batched = [batch.run_round(1), batch.run_round(2)]

    for number, case in enumerate(batch_cases):
        with HIP(tally_shards=32) as single:
            single.prepare(case)
            for index in (1, 2):
                alone = single.run_batch(BatchControls(index, case.photons_per_column))
                assert_tallies_equal(batched[index - 1][number], alone)


def test_round_arrays_survive_later_rounds_and_close():
    """Returned tallies alias no buffer of the session: later rounds and close leave them as
    they were (the reference's ownership rule, tallies.py:21-27, test_metal.py:126, :327-329)."""

    from monki.execution import HIPBatch

    batch_cases = cases(rounds=4)

    with HIPBatch(tally_shards=8) as batch:
        batch.prepare(batch_cases)
        first = batch.run_round(1)
        saved = [(r.stokes.copy(), r.weighted_paths.copy(), r.weights.copy()) for r in first]
        stokes = batch.accumulate_round(2)
        stokes_saved = [a.copy() for a in stokes]
        batch.run_round(3)
        batch.accumulate_round(4)

    for r, (a, b, c) in zip(first, saved):
        np.testing.assert_array_equal(r.stokes, a)
        np.testing.assert_array_equal(r.weighted_paths, b)
        np.testing.assert_array_equal(r.weights, c)

    for a, b in zip(stokes, stokes_saved):
        np.testing.assert_array_equal(a, b)
        assert np.abs(a).max() > 0


def test_graph_replays_equal_direct_launches():
    """The captured round, replayed several times with direct launches of another session and
    host work in between, equals the directly launched round every time."""

    from monki.execution import HIPBatch

    batch_cases = cases(rounds=6)

    with HIPBatch(use_graph=True) as graph, HIPBatch(use_graph=False) as direct:
        graph.prepare(batch_cases)
        direct.prepare(batch_cases)

        for index in range(1, 7):
            replayed = graph.run_round(index)
            launched = direct.run_round(index)
            np.asarray([b.stokes.sum() for b in launched]).sum()  # host work between replays

            for a, b in zip(replayed, launched):
                assert_tallies_equal(a, b)


def test_simulate_batch_equals_simulate(tmp_path):
    from monki import AMFConvergence, simulate, simulate_batch
    from monki.execution import HIP, HIPBatch

    batch_cases = cases(rounds=32, ppc=100)
    policy = AMFConvergence(0.5, 32, 16, 1)
    outputs = [tmp_path / f"batch{i}.nc" for i in range(len(batch_cases))]
    batched = simulate_batch(batch_cases, HIPBatch(), outputs, convergence=policy)

    for i, case in enumerate(batch_cases):
        alone = simulate(case, HIP(), tmp_path / f"alone{i}.nc", convergence=policy)
        # Sums of 32 rounds of float32 banks: absolute tolerance scaled to the largest value.
        expected = alone.values["radiance"]
        np.testing.assert_allclose(batched[i].values["radiance"], expected, rtol=2e-5,
                                   atol=2e-5 * np.abs(expected).max())  # fmt: skip
        np.testing.assert_allclose(batched[i].values["AMF_layer"], alone.values["AMF_layer"], rtol=2e-5)
        receipt = json.loads(outputs[i].with_suffix(".run.json").read_text())
        assert receipt["backend"] == "HIPBatch" and receipt["case"] == case.name


def test_autotuner_picks_a_measured_configuration(tmp_path):
    from monki.execution.autotune import autotune

    cache = tmp_path / "autotune.json"
    result = autotune(cases(), blocks=(128, 256), shards=(8, 32), rounds=2, cache=cache)
    assert not result["cached"] and len(result["table"]) == 4
    assert result["seconds_per_round"] == min(row["device_seconds"] for row in result["table"])
    again = autotune(cases(), blocks=(128, 256), shards=(8, 32), rounds=2, cache=cache)
    assert again["cached"] and again["block"] == result["block"]


def test_spectral_batch_equals_single_wavelength_sessions():
    """Wavelengths of the spectral stochastic field in one batch (59 layers + vacuum: the widest
    tally record, 60 + 3 lanes) equal their single-scene runs."""

    from monki import SceneCase, configs
    from monki.execution import HIP, HIPBatch, BatchControls

    field = configs.stochastic_cloud_field(24, seed=7)
    spectral = [SceneCase(f"w{w:.0f}", configs.spectral_cloud_field(w, field), 20, 2) for w in (300.0, 400.0, 500.0)]

    with HIPBatch(tally_shards=8) as batch:
        batch.prepare(spectral)
        batched = batch.run_round(1)

    for case, result in zip(spectral, batched):
        i = result.stokes[0]
        assert np.all(np.isfinite(result.stokes)) and i.mean() > 0
        with HIP(tally_shards=8) as single:
            single.prepare(case)
            assert_tallies_equal(result, single.run_batch(BatchControls(1, case.photons_per_column)))


class HostRounds:
    """A HIPBatch seen without its device accumulation: simulate_batch then copies every round's
    tallies to the host."""

    def __init__(self, batch):
        self.batch = batch

    def __getattr__(self, name):
        if name == "accumulate_round":
            raise AttributeError(name)

        return getattr(self.batch, name)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.batch.close()


def test_device_moments_equal_host_moments():
    """Sums and AMF moments accumulated on the device equal the host's over the same rounds
    (two sessions, same Philox streams: float32 bank sums differ only in atomic order)."""

    from monki.convergence import RatioMoments
    from monki.execution import HIPBatch

    batch_cases = cases(rounds=6)
    rounds = 6

    with HIPBatch(tally_shards=8) as host, HIPBatch(tally_shards=8) as device:
        host.prepare(batch_cases)
        device.prepare(batch_cases)
        sums = [None] * len(batch_cases)
        moments = [RatioMoments() for _ in batch_cases]

        for index in range(1, rounds + 1):
            stokes = device.accumulate_round(index)

            for number, (full, alone) in enumerate(zip(host.run_round(index), stokes)):
                np.testing.assert_allclose(alone, full.stokes, rtol=2e-5, atol=2e-5 * np.abs(full.stokes).max())
                moments[number].add(full)
                previous = sums[number]
                sums[number] = (full.weighted_paths, full.weights) if previous is None else (
                    previous[0] + full.weighted_paths, previous[1] + full.weights)  # fmt: skip

        for number in range(len(batch_cases)):
            got = device.moments(number)
            assert got.rounds == rounds
            reference = moments[number]
            reference.estimate(np.ones(got.path_mean.shape[2]))  # folds the buffered rounds in
            pairs = [
                (got.paths, sums[number][0]), (got.weights, sums[number][1]),
                (got.path_mean, reference.mean_path), (got.weight_mean, reference.mean_weight[:, :, 0, :]),
            ]  # fmt: skip
            for a, b in pairs:
                np.testing.assert_allclose(a, b, rtol=1e-4, atol=1e-4 * np.abs(b).max())
            # Second moments: differences of rounds, at the scale of the variance.
            for a, b in ((got.path_m2, reference.s_pp), (got.covariance, reference.s_pw),
                         (got.weight_m2, reference.s_ww[:, :, 0, :])):  # fmt: skip
                np.testing.assert_allclose(a, b, rtol=1e-2, atol=1e-3 * np.abs(b).max())


def test_inactive_scenes_are_neither_traced_nor_accumulated():
    from monki.execution import HIPBatch

    batch_cases = cases(rounds=3)

    with HIPBatch(tally_shards=8) as batch:
        batch.prepare(batch_cases)
        first = batch.accumulate_round(1)
        frozen = batch.moments(1)
        second = batch.accumulate_round(2, active=[True, False, True, True])
        assert second[1] is None and all(r is not None for r in first)
        again = batch.moments(1)
        assert again.rounds == 1 == frozen.rounds
        np.testing.assert_array_equal(again.paths, frozen.paths)
        np.testing.assert_array_equal(again.path_m2, frozen.path_m2)
        assert batch.moments(0).rounds == 2


def test_simulate_batch_on_device_equals_host_rounds(tmp_path):
    """simulate_batch with AMF stopping: device-accumulated rounds give the products of the
    host path (scenes stop at their own round counts)."""

    from monki import AMFConvergence, simulate_batch
    from monki.execution import HIPBatch

    batch_cases = cases(rounds=64, ppc=100)
    policy = AMFConvergence(0.2, 32, 16, 1)
    device = simulate_batch(batch_cases, HIPBatch(), [tmp_path / f"d{i}.nc" for i in range(len(batch_cases))],
                            convergence=policy)  # fmt: skip
    host = simulate_batch(batch_cases, HostRounds(HIPBatch()), [tmp_path / f"h{i}.nc" for i in range(len(batch_cases))],
                          convergence=policy)  # fmt: skip

    for i, (a, b) in enumerate(zip(device, host)):
        da = json.loads((tmp_path / f"d{i}.run.json").read_text())
        db = json.loads((tmp_path / f"h{i}.run.json").read_text())
        assert da["completed_rounds"] == db["completed_rounds"] and da["status"] == db["status"]
        for variable in ("radiance", "AMF_layer", "std_AMF_layer"):
            expected = b.values[variable]
            np.testing.assert_allclose(a.values[variable], expected, rtol=1e-4, atol=1e-4 * np.abs(expected).max(),
                                       err_msg=variable)  # fmt: skip


def test_spectral_field_full_size():
    """The spectral case at full size: 64 wavelengths of the 128 x 128 x 60 stochastic field in one
    batch round (1 photon per column). Every tally finite, |Q|, |U| <= I per pixel, and two
    wavelengths equal to their single-scene sessions. (The physics of these scenes is pinned
    history by history against the oracle on a reduced field in test_gpu_transport.py; at one
    photon per column the clear-column radiance is dominated by forward-peak contributions of
    neighbouring clouds, so it carries no usable wavelength trend.)"""

    from monki import SceneCase, configs
    from monki.execution import HIP, HIPBatch, BatchControls

    field = configs.stochastic_cloud_field(128, seed=7)
    wavelengths = configs.spectral_wavelengths(64)
    spectral = [SceneCase(f"w{i:02d}", configs.spectral_cloud_field(w, field), 1, 2) for i, w in enumerate(wavelengths)]
    assert spectral[0].scene.nx == 128 and spectral[0].scene.nz == 60
    print("spectral full size: 64 x 128 x 128 x 60, one batch round", flush=True)

    with HIPBatch(tally_shards=4) as batch:
        batch.prepare(spectral)
        rounds = batch.run_round(1)

    clear = field == 0
    means = []

    for result in rounds:
        i, q, u = result.stokes
        assert np.isfinite(result.stokes).all() and np.isfinite(result.weighted_paths).all()
        assert (i >= 0).all() and (result.weighted_paths >= 0).all() and (result.weights >= 0).all()
        assert (np.abs(q) <= i * (1 + 1e-5) + 1e-12).all() and (np.abs(u) <= i * (1 + 1e-5) + 1e-12).all()
        means.append(i[clear].mean())

    print("clear-column mean I, 300 to 500 nm:", np.round(means[::8], 4), flush=True)

    for number in (0, 63):
        print(f"single-scene session of wavelength {wavelengths[number]:.1f} nm", flush=True)
        with HIP(tally_shards=4) as single:
            single.prepare(spectral[number])
            assert_tallies_equal(rounds[number], single.run_batch(BatchControls(1, 1)))


@pytest.mark.parametrize("mode", ["directed", "guided"])
def test_steered_batch_equals_single_scene_sessions(mode):
    """Steered rounds (view lobes, surface steering, weight window; guided: one network for the
    batch) in a batch equal the same scenes steered alone: same streams, same copies."""

    from monki.execution import HIP, HIPBatch, BatchControls
    from monki.execution.proposal import DirectedSettings, ProposalSettings

    batch_cases = cases(rounds=2, ppc=100)
    box_cases = [batch_cases[0], batch_cases[3]]  # every scene needs the lobe species
    directed = DirectedSettings() if mode == "directed" else None
    proposal = ProposalSettings(events=50_000, photons_per_column=200, steps=20, batch=4096) if mode == "guided" else None

    with HIPBatch(tally_shards=8, directed=directed, proposal=proposal) as batch:
        batch.prepare(box_cases)
        batched = batch.run_round(1)
        parameters = batch._network.parameters if proposal else None

    for number, case in enumerate(box_cases):
        with HIP(tally_shards=8, directed=directed, proposal=proposal) as single:
            single.prepare(case)
            if proposal:
                # The batch trained on its first case; give the single session the same network.
                single.network.parameters = parameters
            assert_tallies_equal(batched[number], single.run_batch(BatchControls(1, case.photons_per_column)))
