"""Named batch requests and owned transport buffers shared by backend implementations.

Same contract as the reference's ``python/monki/execution/tallies.py:8-32``: one request is
one complete statistical round, numbered from one within a session, and returned arrays are
owned, unnormalized and in Fortran axis order.
"""

from dataclasses import dataclass
from typing import Optional

import numpy as np


@dataclass(frozen=True)
class BatchControls:
    """Request one complete statistical round, numbered from one within a session."""

    round_index: int
    photons_per_column: int

    def __post_init__(self):
        # tallies.py:15-19 — both counts are positive 32-bit integers.
        for value in (self.round_index, self.photons_per_column):
            if type(value) is not int or not 0 < value < 2**31:
                raise ValueError("batch counts must be positive 32-bit integers")


@dataclass(frozen=True)
class RawTallies:
    """Own unnormalized Fortran-order Stokes and optional weighted-path sums.

    Stokes axes are (component, x, y, angle); paths are (x, y, layer, angle), including
    the vacuum layer; weights are (x, y, angle). No buffer aliases device memory.
    ``scattering`` (no reference counterpart: the reference's backends refuse the option)
    holds the single/multiple-scattering decomposition when requested: ``stokes_sing``,
    ``stokes_mult`` (like ``stokes``) and ``contrib_sing``, ``contrib_mult``
    (component, x, y, level, angle), level 0 the surface.
    """

    controls: BatchControls
    stokes: np.ndarray
    weighted_paths: Optional[np.ndarray] = None
    weights: Optional[np.ndarray] = None
    scattering: Optional[dict] = None
