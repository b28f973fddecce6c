"""The refusal classes of the reference (python/qed/errors.py; include/ed/core/errors.h:24-50): each is
also the builtin it refines, so callers that catch ValueError / MemoryError / RuntimeError keep working."""

from __future__ import annotations


class QEDError(Exception):
    """Base of every error the backend raises."""


class InvalidRequest(QEDError, ValueError):
    """Arguments that cannot describe a calculation."""


class ResourceLimit(QEDError, MemoryError):
    """Does not fit the device memory or a budget."""


class DeviceUnavailable(QEDError, RuntimeError):
    """device='gpu' without a built library or a visible HIP device."""


class DeviceUnsupported(QEDError, RuntimeError):
    """The block cannot run on the device."""


class ConvergenceError(QEDError, RuntimeError):
    """A Krylov solve did not converge or met a non-finite entry."""


class EngineError(QEDError, RuntimeError):
    """Any other failure inside the engine."""


_BY_STATUS = {
    1: InvalidRequest,
    2: ResourceLimit,
    3: DeviceUnavailable,
    4: DeviceUnsupported,
    5: EngineError,
    6: ConvergenceError,
}


def from_status(status: int, message: str) -> QEDError:
    return _BY_STATUS.get(int(status), EngineError)(message or f"engine status {status}")
