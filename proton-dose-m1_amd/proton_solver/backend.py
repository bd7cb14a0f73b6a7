"""Backend resolution (mirrors ``proton_solver/backend.py:6-59``).

``Backend.GPU`` and ``Backend.AUTO`` resolve to the MI355X march
(libm1march.so); the host arrays exchanged with it are NumPy arrays, so the
resolved array module is NumPy.  ``Backend.CPU`` named the reference's Numba
march, which this package does not contain.
"""

from dataclasses import dataclass

from .enums import Backend


@dataclass(frozen=True)
class ResolvedBackend:
    backend: Backend
    xp: object


def resolve_backend(backend: Backend) -> ResolvedBackend:
    import numpy as np

    if backend is Backend.CPU:
        raise NotImplementedError(
            "The CPU Numba march is not part of the MI355X package; use Backend.GPU."
        )
    if backend in (Backend.GPU, Backend.AUTO):
        from . import _native

        try:
            _native.load_library()
        except RuntimeError as exc:
            raise RuntimeError("GPU backend requested, but libm1march.so is unavailable.") from exc
        return ResolvedBackend(backend=Backend.GPU, xp=np)
    raise ValueError(f"Unknown backend: {backend}")
