"""Interchangeable backend strategies (python/monki/execution/__init__.py)."""

from .backend import Backend
from .hip import HIP
from .tallies import BatchControls, RawTallies

__all__ = ["Backend", "HIP", "BatchControls", "RawTallies"]
