"""Exception types mirroring the Julia exceptions of the reference API."""
from __future__ import annotations


class ArgumentError(ValueError):
    """Julia `ArgumentError`."""


class DimensionMismatch(ValueError):
    """Julia `DimensionMismatch`."""


class MissingSchmidtCoefficientsException(RuntimeError):
    """`Qrochet.MissingSchmidtCoefficientsException` (src/Ansatz.jl:91-99)."""

    def __init__(self, bond, msg=None):
        self.bond = tuple(bond) if bond is not None else None
        super().__init__(msg or f"Can't access the spectrum on bond {self.bond}")


class QHError(RuntimeError):
    """Device-side failure (HIP error, out of memory, non-convergence)."""

    def __init__(self, code, msg):
        self.code = code
        super().__init__(f"[{code}] {msg}")


class AdmissionError(QHError):
    """A serving request whose projected footprint exceeds the memory budget."""
