"""Isotropic elastic material as (mu, lam), never nu.

Mirrors src/mhs/materials.py:33-84: mu > 0, lam > -2 mu / 3, and the one
(mu, nu) adapter that refuses nu >= 1/2.
"""
from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class Material:
    """Shear modulus ``mu`` and Lame's first parameter ``lam``."""

    mu: float
    lam: float

    def __post_init__(self) -> None:
        if not (self.mu > 0.0):
            raise ValueError(f"mu must be positive, got {self.mu!r}")
        if not (self.lam > -2.0 * self.mu / 3.0):
            raise ValueError(
                f"lam must exceed -2 mu / 3 = {-2.0 * self.mu / 3.0!r} for a "
                f"positive bulk modulus, got {self.lam!r}")

    @property
    def bulk(self) -> float:
        return self.lam + 2.0 * self.mu / 3.0

    @property
    def nu_if_you_must(self) -> float:
        return 0.5 * self.lam / (self.lam + self.mu)

    @classmethod
    def from_mu_nu(cls, mu: float, nu: float) -> "Material":
        if not (-1.0 < nu < 0.5):
            raise ValueError(
                f"nu must lie in (-1, 1/2) to convert to lam; got {nu!r}. An "
                f"incompressible material has no finite lam -- construct "
                f"Material(mu, lam) with a large lam directly.")
        return cls(mu=float(mu), lam=2.0 * float(mu) * nu / (1.0 - 2.0 * nu))
