"""Common lifecycle for interchangeable transport implementations.

Mirrors the reference strategy interface ``python/monki/execution/backend.py:10-72``: the
driver owns rounds, convergence and normalization; a backend owns prepared tables and
device memory and returns one owned ``RawTallies`` per requested round.
"""

from abc import ABC, abstractmethod
from types import TracebackType
from typing import Optional, Type


class Backend(ABC):
    """Strategy interface consumed by the shared simulation driver."""

    metadata = None
    provenance: dict = {}

    @abstractmethod
    def prepare(self, scene: object) -> None:
        """Validate and retain scene tables before running batches.

        Repreparing replaces prior scene state. Failure must leave resources safe to close.
        """

        raise NotImplementedError

    @abstractmethod
    def run_batch(self, controls: object) -> object:
        """Return owned raw tallies for one complete round, without normalization."""

        raise NotImplementedError

    def run_batches(self, controls: list) -> list:
        """Return one owned result per requested round, preserving order (backend.py:44-51)."""

        return [self.run_batch(request) for request in controls]

    @abstractmethod
    def close(self) -> None:
        """Release resources; safe before preparation and on repeated calls."""

        raise NotImplementedError

    def __enter__(self):
        return self

    def __exit__(
        self,
        exception_type: Optional[Type[BaseException]],
        exception: Optional[BaseException],
        traceback: Optional[TracebackType],
    ) -> None:
        self.close()
