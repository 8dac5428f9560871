"""MONKI photon transport with an MI355X (gfx950) HIP backend.

Public surface mirrors the reference package ``python/monki/__init__.py``:
``simulate``, ``AMFConvergence``, ``ReferenceCase``, ``SimulationResult``; plus ``SceneCase``
for prepared in-memory transport tables.
"""

from .convergence import AMFConvergence
from .input import ReferenceCase, SceneCase
from .output import SimulationResult
from .simulation import simulate, simulate_batch

__all__ = ["AMFConvergence", "ReferenceCase", "SceneCase", "SimulationResult", "simulate", "simulate_batch"]
