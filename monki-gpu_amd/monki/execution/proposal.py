"""Learned scattering-direction proposal for the guided HIP tracer.

No reference counterpart: the reference samples every scattering direction from the phase
function (``mc_routines_general_mod.f90:492-932``). Here a BF16 MLP (12 inputs: periodic
position, height, direction, log extinction, single-scattering albedo, scatterer kind and
scattering order; 4 hidden layers of 128 SiLU units) maps the state of a scattering event to
an 8-component von Mises-Fisher mixture q over the outgoing direction. The guided tracer
draws the direction from the phase function p with probability ``alpha`` and from q
otherwise, and multiplies the photon weight by p / (alpha p + (1 - alpha) q): the estimator
stays unbiased for any network, and the factor is at most 1 / alpha.

Training data are scattering events of the plain tracer on their own random streams (seed
distinct from the estimation rounds), weighted by the detector weight the photon contributed
after the event; the loss is the weighted negative log-likelihood of the observed direction.
Forward and backward passes run in the library's own matrix-core kernels.
"""

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

PARAMETERS = 56360
INPUTS = 12
OUTPUTS = 40
_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)
_U = ctypes.POINTER(ctypes.c_uint32)
_U64 = ctypes.POINTER(ctypes.c_uint64)


def _check_window(settings) -> None:
    """Shared validation of the steering and weight-window fields."""

    if not 0.0 < float(settings.alpha) <= 1.0:
        raise ValueError("alpha must be in (0, 1]")

    if type(settings.surface) is not bool:
        raise ValueError("surface must be a bool")

    split, below, weight = float(settings.split_above), float(settings.roulette_below), float(settings.roulette_weight)

    if split != 0.0 and not (split > 1.0 and 0.0 <= below < split and (below == 0.0 or below < weight <= split)):
        raise ValueError("weight window: split_above 0 (off) or > 1, 0 <= roulette_below < roulette_weight <= split_above")

    tabulated = settings.species_mask() & 0xFFFFFFFE
    # Without a window the steering factors of a history multiply unchecked: over the tens of
    # events of a thick cloud the estimator, though unbiased, becomes too heavy-tailed for its
    # standard error to mean anything (guided = plain failed at |z| 11.6 that way).
    if split == 0.0 and tabulated and settings.max_order > 4:
        raise ValueError("steering tabulated species beyond order 4 needs the weight window (split_above > 1)")


@dataclass(frozen=True)
class ProposalSettings:
    """How the guided tracer's proposal is trained and used.

    Training: ``events`` scattering events logged from the plain tracer (``photons_per_column``
    per logging round), then ``steps`` Adam steps over random minibatches of ``batch`` samples
    at ``learning_rate``; ``seed`` initializes the network and the logging streams. The
    defaults log 10^8 events (80 bytes each on the device, 8 GB) and draw about one epoch of
    them (6,000 x 16,384 samples).

    Use: a steered event draws its direction from the phase function with probability
    ``alpha``, else from the proposal: ``lobe_fraction`` x lobes around the view directions
    shaped by the sharpest tabulated phase function + (1 - ``lobe_fraction``) x the network's
    vMF mixture. Steered events: atmospheric events of ``guided_kinds`` ("gas" and/or tabulated
    species indices; None for every kind) up to ``max_order``, and Lambertian reflections when
    ``surface``. The network's mixture is part of the proposal at the first ``network_orders``
    scattering orders only; beyond them the lobes alone (no steering when ``lobe_fraction`` is
    0), and the network does not run: on the box cloud the mixture lowers the variance at the
    first two orders and raises it at higher ones. Weight window: photons heavier than ``split_above`` are split, lighter than
    ``roulette_below`` play Russian roulette for ``roulette_weight`` (``split_above`` 0 turns
    the window off, which is refused for tabulated species beyond order 4).

    Training target: a sample's weight is the detector weight G the photon contributed after the
    event; with ``path_target`` p > 0 it is (1 - p) G + p s P, P the layer path the photon still
    travelled before each later estimate, weighted by that estimate's detector weight and by
    1 / (the layer's path tally in a plain round), s = sum G / sum P: the network then also
    values directions that feed the layers whose paths are rare.
    """

    events: int = 100_000_000
    photons_per_column: int = 2_000
    steps: int = 6_000
    batch: int = 16_384
    learning_rate: float = 1e-3
    alpha: float = 0.9
    guided_kinds: Optional[tuple] = None
    max_order: int = 2**32 - 1
    seed: int = 97
    lobe_fraction: float = 0.5
    surface: bool = True
    split_above: float = 2.0
    roulette_below: float = 0.005
    roulette_weight: float = 0.01
    path_target: float = 0.0
    network_orders: int = 2

    def species_mask(self) -> int:
        if self.guided_kinds is None:
            return 0xFFFFFFFF

        mask = 0

        for kind in self.guided_kinds:
            mask |= 1 if kind == "gas" else 1 << (int(kind) + 1)

        return mask

    def __post_init__(self):
        for name, value, low, high in (
            ("events", self.events, 64, 1 << 31),
            ("photons_per_column", self.photons_per_column, 1, 1 << 24),
            ("steps", self.steps, 0, 1 << 24),
            ("batch", self.batch, 64, 1 << 24),
            ("seed", self.seed, 0, 2**32 - 1),
            ("max_order", self.max_order, 1, 2**32 - 1),
            ("network_orders", self.network_orders, 1, 2**32 - 1),
        ):
            if type(value) is not int or not low <= value <= high:
                raise ValueError(f"{name} must be an integer in [{low}, {high}]")

        if not 0.0 < float(self.learning_rate) < 1.0:
            raise ValueError("learning_rate must be in (0, 1)")

        if not 0.0 <= float(self.lobe_fraction) <= 1.0:
            raise ValueError("lobe_fraction must be in [0, 1]")

        if not 0.0 <= float(self.path_target) <= 1.0:
            raise ValueError("path_target must be in [0, 1]")

        if self.guided_kinds is not None:
            for kind in self.guided_kinds:
                if kind != "gas" and not (type(kind) is int and 0 <= kind < 4):
                    raise ValueError("guided_kinds holds 'gas' or tabulated species indices 0-3")

        _check_window(self)


def declare(native) -> None:
    """ctypes declarations of the proposal entry points of libmonki_hip.so."""

    native.monki_hip_proposal_parameter_count.restype = ctypes.c_size_t
    native.monki_hip_proposal_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_proposal_create.restype = ctypes.c_void_p
    native.monki_hip_proposal_set_parameters.argtypes = [ctypes.c_void_p, _F, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_proposal_get_parameters.argtypes = [ctypes.c_void_p, _F, ctypes.c_size_t]
    native.monki_hip_proposal_reserve.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_proposal_set_samples.argtypes = [ctypes.c_void_p, _F, _F, _F, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_proposal_samples.argtypes = [ctypes.c_void_p, _U64, _D]
    native.monki_hip_proposal_set_path_target.argtypes = [ctypes.c_void_p, _F, ctypes.c_size_t, ctypes.c_float, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_proposal_read_samples.argtypes = [ctypes.c_void_p, _F, _F, _F, ctypes.c_size_t]
    native.monki_hip_proposal_gradients.argtypes = [ctypes.c_void_p, ctypes.c_size_t, _F, _D, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_proposal_train.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint64, _D, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_proposal_forward.argtypes = [ctypes.c_void_p, _F, ctypes.c_size_t, _F, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_proposal_close.argtypes = [ctypes.c_void_p]
    native.monki_hip_proposal_close.restype = None
    native.monki_hip_attach_proposal.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_set_network_orders.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_run_guided.argtypes = [ctypes.c_void_p, _U, _D, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_trace_guided.argtypes = [ctypes.c_void_p, _U, _D, _F, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_run_logged.argtypes = [ctypes.c_void_p, _U, _D, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_set_directed.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_run_directed.argtypes = [ctypes.c_void_p, _U, _D, ctypes.c_char_p, ctypes.c_size_t]
    native.monki_hip_set_steering.argtypes = [
        ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
        ctypes.c_char_p, ctypes.c_size_t,
    ]  # fmt: skip

    for name in (
        "monki_hip_proposal_set_parameters", "monki_hip_proposal_get_parameters", "monki_hip_proposal_reserve",
        "monki_hip_proposal_set_samples", "monki_hip_proposal_samples", "monki_hip_proposal_read_samples",
        "monki_hip_proposal_set_path_target",
        "monki_hip_proposal_gradients", "monki_hip_proposal_train", "monki_hip_proposal_forward",
        "monki_hip_attach_proposal", "monki_hip_run_guided", "monki_hip_trace_guided", "monki_hip_run_logged",
        "monki_hip_set_directed", "monki_hip_run_directed", "monki_hip_set_steering", "monki_hip_set_network_orders",
    ):  # fmt: skip
        getattr(native, name).restype = ctypes.c_int


@dataclass(frozen=True)
class DirectedSettings:
    """Detector-directed sampling (no network). A steered event draws its direction from the
    physical sampler (phase function; Lambert at the surface) with probability ``alpha``, else
    from lobes around the view directions, and multiplies the weight by
    p / (alpha p + (1 - alpha) q). ``lobes="phase"`` shapes the lobes by the sharpest tabulated
    phase function of the scene (its forward peak measured from each view direction), so that
    a photon steered into the peak toward a view carries the inverse peak in its weight and the
    local estimate at its next event loses its forward-peak spike; ``lobes="vmf"`` uses vMF
    lobes of concentration ``kappa``. ``guided_kinds``, ``max_order``, ``surface`` and the
    weight window as in ProposalSettings. The defaults steer every event, the surface
    included, with the window on.
    """

    alpha: float = 0.9
    kappa: float = 500.0
    guided_kinds: Optional[tuple] = None
    max_order: int = 2**32 - 1
    lobes: str = "phase"
    surface: bool = True
    split_above: float = 2.0
    roulette_below: float = 0.005
    roulette_weight: float = 0.01

    def species_mask(self) -> int:
        if self.guided_kinds is None:
            return 0xFFFFFFFF
        mask = 0
        for kind in self.guided_kinds:
            if kind == "gas":
                mask |= 1
            elif kind == "tabulated":
                mask |= 0xFFFFFFFE
            else:
                mask |= 1 << (int(kind) + 1)
        return mask

    def __post_init__(self):
        if not (isinstance(self.kappa, (int, float)) and 1 <= self.kappa <= 1e6):
            raise ValueError("kappa must be in [1, 1e6]")

        if type(self.max_order) is not int or not 1 <= self.max_order <= 2**32 - 1:
            raise ValueError("max_order must be an integer in [1, 2**32 - 1]")

        if self.lobes not in ("phase", "vmf"):
            raise ValueError("lobes is 'phase' or 'vmf'")

        _check_window(self)


def sharpest_species(scene) -> int:
    """Index of the tabulated species with the highest forward peak (F11 at cos 1, F11 being
    normalized to a sphere mean of one), or -1 without tabulated species."""

    if not scene.species:
        return -1

    return int(np.argmax([float(species.f11[-1]) for species in scene.species]))


def apply_steering(native, handle, settings, lobe_species: int, lobe_fraction: float) -> None:
    """Lobes, surface steering and weight window of a session (monki_hip_set_steering)."""

    error = ctypes.create_string_buffer(4096)

    if native.monki_hip_set_steering(
        handle, lobe_species, lobe_fraction, int(settings.surface), settings.split_above, settings.roulette_below,
        settings.roulette_weight, error, len(error),
    ):  # fmt: skip
        raise RuntimeError(error.value.decode())

    # The guided tracer's network takes part in the proposal up to its orders.
    orders = getattr(settings, "network_orders", None)

    if orders is not None and native.monki_hip_set_network_orders(handle, min(orders, 2**32 - 1), error, len(error)):
        raise RuntimeError(error.value.decode())


class Proposal:
    """A proposal network resident on one device (parameters, Adam state, samples)."""

    def __init__(self, native, device: int = 0, seed: int = 97):
        declare(native)

        if native.monki_hip_proposal_parameter_count() != PARAMETERS:
            raise RuntimeError("Rebuild libmonki_hip.so: proposal parameter count differs")

        error = ctypes.create_string_buffer(4096)
        handle = native.monki_hip_proposal_create(device, seed, error, len(error))

        if not handle:
            raise RuntimeError(error.value.decode() or "proposal creation failed")

        self.native = native
        self.handle = handle
        self.device = device

    def _check(self, status: int, error) -> None:
        if status:
            raise RuntimeError(error.value.decode() or "proposal call failed")

    @property
    def parameters(self) -> np.ndarray:
        out = np.zeros(PARAMETERS, dtype=np.float32)
        self._check(self.native.monki_hip_proposal_get_parameters(self.handle, out.ctypes.data_as(_F), out.size), None)
        return out

    @parameters.setter
    def parameters(self, values) -> None:
        values = np.ascontiguousarray(values, dtype=np.float32)
        error = ctypes.create_string_buffer(4096)
        self._check(
            self.native.monki_hip_proposal_set_parameters(self.handle, values.ctypes.data_as(_F), values.size, error, len(error)),
            error,
        )  # fmt: skip

    def reserve(self, events: int, photons: int) -> None:
        error = ctypes.create_string_buffer(4096)
        self._check(self.native.monki_hip_proposal_reserve(self.handle, events, photons, error, len(error)), error)

    def set_samples(self, features, directions, weights) -> None:
        features = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, INPUTS)
        directions = np.ascontiguousarray(directions, dtype=np.float32).reshape(-1, 3)
        weights = np.ascontiguousarray(weights, dtype=np.float32).ravel()
        error = ctypes.create_string_buffer(4096)
        self._check(
            self.native.monki_hip_proposal_set_samples(
                self.handle, features.ctypes.data_as(_F), directions.ctypes.data_as(_F),
                weights.ctypes.data_as(_F), len(weights), error, len(error),
            ),
            error,
        )  # fmt: skip

    def set_path_target(self, layer_weights, mix: float) -> None:
        weights = np.ascontiguousarray(layer_weights, dtype=np.float32)
        error = ctypes.create_string_buffer(1024)
        self._check(self.native.monki_hip_proposal_set_path_target(self.handle, weights.ctypes.data_as(_F), weights.size,
                                                                   float(mix), error, len(error)), error)  # fmt: skip

    def samples(self):
        """(sample count, sum of weights)."""

        count = ctypes.c_uint64()
        total = ctypes.c_double()
        self._check(self.native.monki_hip_proposal_samples(self.handle, ctypes.byref(count), ctypes.byref(total)), None)
        return int(count.value), float(total.value)

    def read_samples(self, count: Optional[int] = None):
        available, _ = self.samples()
        count = available if count is None else count
        features = np.zeros((count, INPUTS), dtype=np.float32)
        directions = np.zeros((count, 3), dtype=np.float32)
        weights = np.zeros(count, dtype=np.float32)
        self._check(
            self.native.monki_hip_proposal_read_samples(
                self.handle, features.ctypes.data_as(_F), directions.ctypes.data_as(_F),
                weights.ctypes.data_as(_F), count,
            ),
            None,
        )  # fmt: skip
        return features, directions, weights

    def gradients(self, count: int):
        """Gradient of -sum w log q / sum w over the first ``count`` samples, and that loss."""

        grad = np.zeros(PARAMETERS, dtype=np.float32)
        loss = ctypes.c_double()
        error = ctypes.create_string_buffer(4096)
        self._check(
            self.native.monki_hip_proposal_gradients(self.handle, count, grad.ctypes.data_as(_F), ctypes.byref(loss), error, len(error)),
            error,
        )  # fmt: skip
        return grad, float(loss.value)

    def train(self, steps: int, batch: int, learning_rate: float, seed: int = 1) -> np.ndarray:
        losses = np.zeros(max(steps, 1), dtype=np.float64)
        error = ctypes.create_string_buffer(4096)
        self._check(
            self.native.monki_hip_proposal_train(
                self.handle, steps, batch, learning_rate, seed, losses.ctypes.data_as(_D), error, len(error)
            ),
            error,
        )  # fmt: skip
        return losses[:steps]

    def forward(self, features) -> np.ndarray:
        features = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, INPUTS)
        out = np.zeros((len(features), OUTPUTS), dtype=np.float32)
        error = ctypes.create_string_buffer(4096)
        self._check(
            self.native.monki_hip_proposal_forward(
                self.handle, features.ctypes.data_as(_F), len(features), out.ctypes.data_as(_F), error, len(error)
            ),
            error,
        )  # fmt: skip
        return out

    def close(self) -> None:
        if self.handle is not None:
            self.native.monki_hip_proposal_close(self.handle)

        self.handle = None


def path_layer_weights(tallies: np.ndarray, layout: tuple) -> np.ndarray:
    """1 / (path tally of each layer, summed over pixels and views) of one round's tallies in the
    output order [Stokes (nmat, nx, ny, na) | paths (nx, ny, nz, na) | weights (nx, ny, na)];
    layers without path get 0. layout = (nmat, nx, ny, nz, na)."""

    nmat, nx, ny, nz, na = layout
    start = nmat * nx * ny * na
    paths = tallies[start:start + nx * ny * nz * na].reshape((nx, ny, nz, na), order="F").sum(axis=(0, 1, 3))
    return np.where(paths > 0.0, 1.0 / np.where(paths > 0.0, paths, 1.0), 0.0)


def log_rounds(native, session, proposal: Proposal, settings: ProposalSettings, columns: int, seed: int,
               max_steps: int, photons_per_dispatch: int, tally_count: int, layout: tuple = None) -> dict:
    """Run plain logging rounds (their own seed) until ``settings.events`` samples exist.
    layout = (nmat, nx, ny, nz, na) of the scene, needed for a path target."""

    per_round = settings.photons_per_column * columns
    dispatch = min(per_round, photons_per_dispatch)
    proposal.reserve(settings.events, dispatch)
    out = np.zeros(tally_count, dtype=np.float64)
    error = ctypes.create_string_buffer(8192)
    round_index = 0
    events = 0
    layer_weights = None

    if settings.path_target > 0.0:
        # Layer weights from a plain round of the first dispatch's photons on streams of their own:
        # 1 / (the layer's path tally summed over pixels and views); its samples are dropped.
        command = (ctypes.c_uint32 * 8)(dispatch, 0, 1, (seed + 1) % 2**32, max_steps, 0, per_round, 1)
        if native.monki_hip_run_logged(session, command, out.ctypes.data_as(_D), proposal.handle, error, len(error)):
            raise RuntimeError(error.value.decode() or "logging round failed")
        layer_weights = path_layer_weights(out, layout)
        proposal.reserve(settings.events, dispatch)
        proposal.set_path_target(layer_weights, settings.path_target)

    while events < settings.events and round_index < 4096:
        round_index += 1

        for first in range(0, per_round, dispatch):
            count = min(dispatch, per_round - first)
            command = (ctypes.c_uint32 * 8)(count, first, round_index, seed, max_steps, 0, per_round, 1)
            status = native.monki_hip_run_logged(session, command, out.ctypes.data_as(_D), proposal.handle, error, len(error))

            if status:
                raise RuntimeError(error.value.decode() or "logging round failed")

            events, _ = proposal.samples()

            if events >= settings.events:
                break

    result = {"rounds": round_index, "events": events, "weight_sum": proposal.samples()[1]}

    if layer_weights is not None:
        result["layer_weights"] = [float(w) for w in layer_weights]

    return result

