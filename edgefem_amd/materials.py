"""Dispersive material models (include/edgefem/materials/dispersive.hpp:14-290; the
``materials`` submodule of the reference binding, python/pyedgefem.cpp:1100-1221).

Each model gives the relative permittivity at an angular frequency omega (rad/s), with the
reference's e^{j omega t} sign convention (losses as a negative imaginary part):

* Debye:   eps_inf + (eps_s - eps_inf) / (1 + j omega tau)
* Lorentz: eps_inf + sum_p d_p omega_p0^2 / (omega_p0^2 - omega^2 + j gamma_p omega)
* Drude:   1 - omega_p^2 / (omega^2 + j gamma omega)            (-1e30 at omega = 0)
* Drude-Lorentz: eps_inf - omega_p^2 / (omega^2 + j gamma_d omega) + Lorentz poles
  (-1e30 replaces the whole value at omega = 0, before the poles are added)

``eval_mu`` is 1 for every model.  MaxwellParams.set_eps_model / set_mu_model attach a
model to a physical region; it takes precedence over the region's static value and is
evaluated at the assembly's omega.  Host code: a model is evaluated once per region and
frequency, then scattered to the per-tet arrays the GPU assembly reads.
"""
from __future__ import annotations

from typing import List, NamedTuple


class DispersiveMaterial:
    """Base class (dispersive.hpp:19-35)."""

    def eval_eps(self, omega: float) -> complex:
        raise NotImplementedError

    def eval_mu(self, omega: float) -> complex:
        return 1.0 + 0j


class DebyeMaterial(DispersiveMaterial):
    """Single-relaxation polar dielectric (dispersive.hpp:37-77)."""

    def __init__(self, eps_static: float, eps_inf: float, tau: float):
        if tau <= 0.0:
            raise ValueError("DebyeMaterial: tau must be positive")
        self._eps_s, self._eps_inf, self._tau = float(eps_static), float(eps_inf), float(tau)

    def eval_eps(self, omega: float) -> complex:
        return self._eps_inf + (self._eps_s - self._eps_inf) / complex(1.0, float(omega) * self._tau)

    @property
    def eps_static(self) -> float:
        return self._eps_s

    @property
    def eps_inf(self) -> float:
        return self._eps_inf

    @property
    def tau(self) -> float:
        return self._tau


class LorentzPole(NamedTuple):
    delta_eps: float
    omega0: float
    gamma: float


def _lorentz_sum(poles: List[LorentzPole], omega: float) -> complex:
    acc = 0j
    w2 = omega * omega
    for p in poles:
        w02 = p.omega0 * p.omega0
        acc += p.delta_eps * w02 / complex(w02 - w2, p.gamma * omega)
    return acc


def _check_pole(owner: str, omega0: float, gamma: float) -> None:
    if omega0 <= 0.0:
        raise ValueError(f"{owner}: omega0 must be positive")
    if gamma < 0.0:
        raise ValueError(f"{owner}: gamma must be non-negative")


class LorentzMaterial(DispersiveMaterial):
    """Multi-pole Lorentz oscillator (dispersive.hpp:79-146)."""

    def __init__(self, eps_inf: float = 1.0):
        self._eps_inf = float(eps_inf)
        self._poles: List[LorentzPole] = []

    def add_pole(self, delta_eps: float, omega0: float, gamma: float) -> None:
        _check_pole("LorentzMaterial", omega0, gamma)
        self._poles.append(LorentzPole(float(delta_eps), float(omega0), float(gamma)))

    def eval_eps(self, omega: float) -> complex:
        return self._eps_inf + _lorentz_sum(self._poles, float(omega))

    @property
    def eps_inf(self) -> float:
        return self._eps_inf

    def set_eps_inf(self, eps_inf: float) -> None:
        self._eps_inf = float(eps_inf)

    @property
    def poles(self) -> List[LorentzPole]:
        return list(self._poles)

    @property
    def num_poles(self) -> int:
        return len(self._poles)


class DrudeMaterial(DispersiveMaterial):
    """Free-electron metal or plasma (dispersive.hpp:148-197)."""

    def __init__(self, omega_p: float, gamma: float):
        if omega_p <= 0.0:
            raise ValueError("DrudeMaterial: omega_p must be positive")
        if gamma < 0.0:
            raise ValueError("DrudeMaterial: gamma must be non-negative")
        self._omega_p, self._gamma = float(omega_p), float(gamma)

    def eval_eps(self, omega: float) -> complex:
        omega = float(omega)
        if omega == 0.0:
            return complex(-1e30, 0.0)  # the DC limit of a pure Drude metal
        return 1.0 - (self._omega_p * self._omega_p) / complex(omega * omega, self._gamma * omega)

    @property
    def omega_p(self) -> float:
        return self._omega_p

    @property
    def gamma(self) -> float:
        return self._gamma


class DrudeLorentzMaterial(DispersiveMaterial):
    """Drude term plus Lorentz poles (dispersive.hpp:199-287)."""

    def __init__(self, eps_inf: float, omega_p: float, gamma_d: float):
        if omega_p <= 0.0:
            raise ValueError("DrudeLorentzMaterial: omega_p must be positive")
        if gamma_d < 0.0:
            raise ValueError("DrudeLorentzMaterial: gamma_d must be non-negative")
        self._eps_inf, self._omega_p, self._gamma_d = float(eps_inf), float(omega_p), float(gamma_d)
        self._poles: List[LorentzPole] = []

    def add_lorentz_pole(self, delta_eps: float, omega0: float, gamma: float) -> None:
        _check_pole("DrudeLorentzMaterial", omega0, gamma)
        self._poles.append(LorentzPole(float(delta_eps), float(omega0), float(gamma)))

    def eval_eps(self, omega: float) -> complex:
        omega = float(omega)
        if omega != 0.0:
            eps = self._eps_inf - (self._omega_p * self._omega_p) / complex(omega * omega, self._gamma_d * omega)
        else:
            eps = complex(-1e30, 0.0)
        return eps + _lorentz_sum(self._poles, omega)

    @property
    def eps_inf(self) -> float:
        return self._eps_inf

    @property
    def omega_p(self) -> float:
        return self._omega_p

    @property
    def gamma_d(self) -> float:
        return self._gamma_d

    @property
    def lorentz_poles(self) -> List[LorentzPole]:
        return list(self._poles)
