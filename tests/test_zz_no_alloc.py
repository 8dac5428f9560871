"""monki_hip_run_static, the entry point that allocates nothing, against monki_hip_run on the
same tables and Philox streams: two calls in a row, the second on a smaller input."""

import ctypes

import numpy as np
import pytest


def _controls(per_round, seed=23274):
    return (ctypes.c_uint32 * 8)(per_round, 0, 1, seed, 100_000, 0, per_round, 1)


def _static(native, packed, per_round, tally_count):
    from monki.execution.hip import DOUBLE_POINTER, FLOAT_POINTER

    out = np.zeros(tally_count)
    error = ctypes.create_string_buffer(4096)
    status = native.monki_hip_run_static(0, packed.ctypes.data_as(FLOAT_POINTER), packed.size, _controls(per_round),
                                         out.ctypes.data_as(DOUBLE_POINTER), error, len(error))  # fmt: skip
    assert status == 0, error.value.decode()
    return out


def _session(native, packed, per_round, tally_count, shards):
    from monki.execution.hip import DOUBLE_POINTER, FLOAT_POINTER

    out = np.zeros(tally_count)
    error = ctypes.create_string_buffer(4096)
    handle = native.monki_hip_create(0, packed.ctypes.data_as(FLOAT_POINTER), packed.size, tally_count * shards, 1,
                                     error, len(error))  # fmt: skip
    assert handle, error.value.decode()

    try:
        status = native.monki_hip_run(handle, _controls(per_round), out.ctypes.data_as(DOUBLE_POINTER), error,
                                      len(error))  # fmt: skip
        assert status == 0, error.value.decode()
    finally:
        native.monki_hip_close(handle)

    return out


@pytest.mark.gpu
def test_static_round_equals_session_round_twice_second_smaller():
    from monki import configs
    from monki.execution.hip import load_library
    from monki.execution.hip_tables import pack_tables, tally_count

    native = load_library()

    for scene, ppc, shards in ((configs.box_cloud(nx=12, ny=8, box=(4, 3)), 400, 8),
                               (configs.layered_cloud(nx=3, ny=2), 150, 4)):  # fmt: skip
        packed, shape = pack_tables(scene, shards=shards)
        count = tally_count(*shape, True, False)
        per_round = ppc * scene.nx * scene.ny
        print(f"static vs session: {scene.nx}x{scene.ny}, {per_round} photons, {shards} shards", flush=True)
        static = _static(native, packed, per_round, count)
        session = _session(native, packed, per_round, count, shards)
        # Identical histories; float32 bank sums differ only in the order of the atomics.
        np.testing.assert_allclose(static, session, rtol=2e-5, atol=2e-6 * np.abs(session).max())
        assert np.abs(session).max() > 0


def test_static_round_rejects_a_group_beyond_the_scratch():
    """Rejected before any device call (runs without a GPU)."""

    from monki import configs
    from monki.execution.hip import DOUBLE_POINTER, FLOAT_POINTER, load_library
    from monki.execution.hip_tables import pack_tables, tally_count

    native = load_library()
    scene = configs.layered_cloud(nx=200, ny=200)
    packed, shape = pack_tables(scene, shards=64)
    out = np.zeros(8)
    error = ctypes.create_string_buffer(4096)
    status = native.monki_hip_run_static(0, packed.ctypes.data_as(FLOAT_POINTER), packed.size, _controls(40_000),
                                         out.ctypes.data_as(DOUBLE_POINTER), error, len(error))  # fmt: skip
    assert status != 0
    assert "static scratch" in error.value.decode()
    assert tally_count(*shape, True, False) > 0
