"""Interchangeable backend strategies (python/monki/execution/__init__.py)."""

from .backend import Backend
from .batch import HIPBatch
from .hip import HIP
from .proposal import DirectedSettings, ProposalSettings
from .tallies import BatchControls, RawTallies

__all__ = ["Backend", "HIP", "HIPBatch", "DirectedSettings", "ProposalSettings", "BatchControls", "RawTallies"]
