"""Dispersive susceptibilities (Lorentz / Drude poles) and their ADE coefficients.

Mirrors src/Susceptibility.jl: the pole types (:16-45), the coefficients of
the auxiliary-differential-equation update (:47-85, evaluated in float64 in
the same operation order) and the closed-form susceptibility chi(f) (:87-101).
Frequencies follow the reference's convention (angular frequency / 2 pi).
"""
from __future__ import annotations

import math
from dataclasses import dataclass


@dataclass(frozen=True)
class LorentzianSusceptibility:
    """chi(f) = sigma w0^2 / (w0^2 - w^2 - i w gamma) with w = 2 pi f (src/Susceptibility.jl:16-33)."""

    omega_0: float
    gamma: float
    sigma: float


def DrudeSusceptibility(gamma: float, sigma: float) -> LorentzianSusceptibility:
    """A Lorentzian pole with omega_0 = 0 (src/Susceptibility.jl:35-45)."""
    return LorentzianSusceptibility(0.0, float(gamma), float(sigma))


@dataclass(frozen=True)
class ADECoefficients:
    """src/Susceptibility.jl:47-65; the per-voxel sigma is applied separately."""

    gamma1_inv: float
    gamma1: float
    omega0_dt_sq: float
    sigma_omega0_dt_sq: float
    drude_coeff: float
    is_drude: bool

    def as_array(self):
        return [self.gamma1_inv, self.gamma1, self.omega0_dt_sq, self.sigma_omega0_dt_sq, self.drude_coeff]


def compute_ade_coefficients(s: LorentzianSusceptibility, dt: float) -> ADECoefficients:
    """src/Susceptibility.jl:67-85, float64 with the reference's association."""
    gamma_pi_dt = (s.gamma * math.pi) * dt
    gamma1 = 1.0 - gamma_pi_dt
    gamma1_inv = 1.0 / (1.0 + gamma_pi_dt)
    omega0_dt_sq = ((2 * math.pi) * s.omega_0 * dt) ** 2
    drude_coeff = (s.gamma * (2 * math.pi)) * dt * dt
    return ADECoefficients(gamma1_inv, gamma1, omega0_dt_sq, omega0_dt_sq, drude_coeff, s.omega_0 == 0.0)


def eval_susceptibility(s: LorentzianSusceptibility, freq: float) -> complex:
    """chi at frequency ``freq`` (src/Susceptibility.jl:87-101)."""
    w = 2 * math.pi * freq
    w0 = 2 * math.pi * s.omega_0
    g = 2 * math.pi * s.gamma
    if s.omega_0 == 0:
        return -s.sigma * g / (w ** 2 + 1j * w * g)
    return s.sigma * w0 ** 2 / (w0 ** 2 - w ** 2 - 1j * w * g)


def has_susceptibilities(material) -> bool:
    sus = getattr(material, "susceptibilities", None)
    return bool(sus)


def chi1_coefficient(c: ADECoefficients) -> float:
    """Per-sigma semi-implicit correction gamma1^-1 C / 2 of one pole
    (src/Geometry.jl:1224-1236), before the cast to the backend type."""
    if c.is_drude:
        return c.gamma1_inv * c.drude_coeff / 2
    return c.gamma1_inv * c.sigma_omega0_dt_sq / 2
