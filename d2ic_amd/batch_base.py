"""Batch runner contract (reference ``d2ic/batch_base.py:9-52``):
``run`` = ``before`` -> ``sequence`` -> ``end``."""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import Sequence

from .dataclasses import BatchResult
from .types import Array


class BatchBase(ABC):
    def __init__(self) -> None:
        super().__init__()

    def run(self, images: Sequence[Array]) -> BatchResult:
        self.before(images)
        result = self.sequence(images)
        return self.end(result)

    @abstractmethod
    def before(self, images: Sequence[Array]) -> None:
        raise NotImplementedError

    @abstractmethod
    def sequence(self, images: Sequence[Array]) -> BatchResult:
        raise NotImplementedError

    @abstractmethod
    def end(self, result: BatchResult) -> BatchResult:
        raise NotImplementedError
