"""CudaStatus / CudaStatusError (the reference's cupano/status.py interface):
a code (0 = ok; a hipError_t code for failures inside the library) and a
message."""
from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class CudaStatus:
    code: int = 0
    message: str = ""

    def ok(self) -> bool:
        return self.code == 0


class CudaStatusError(RuntimeError):
    def __init__(self, status: CudaStatus):
        self.status = status
        super().__init__(status.message or f"CudaStatus(code={status.code})")


def raise_if_error(status: CudaStatus) -> None:
    if not status.ok():
        raise CudaStatusError(status)
