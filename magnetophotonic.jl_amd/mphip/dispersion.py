"""Drude-Lorentz ADE poles (host side).

Restates src/fdtd/Dispersion.jl: `create_pole` (:12-19), the discrete susceptibility
probe `discrete_pole_chi` (:24-48), the two-pole strength fit `_fit_pole_strengths`
(:50-54) and the pump / probe / MO pole sets (:56-91).  A pole is the triple
(C1, C2, C3) of DLPole{T}.
"""
from __future__ import annotations

from fractions import Fraction

import numpy as np

PROBE_POLE_POSITIONS = ((0.0, 3.0e14), (4.0e15, 8.0e14))
PROBE_MO_POLE_POSITIONS = ((4.0e15, 8.0e14), (5.0e15, 1.2e15))


def create_pole(omega0: float, gamma: float, delta_eps_omega0_sq: float, dt: float, eps0: float):
    """create_pole (Dispersion.jl:12-19) -> (C1, C2, C3)."""
    alpha = float(dt) / 2.0
    ao = alpha * float(omega0)
    D = 1.0 + alpha * float(gamma) + ao * ao
    C1 = (1.0 + alpha * float(gamma) - ao * ao) / D
    C2 = (2.0 * alpha) / D
    C3 = (alpha * alpha * float(eps0) * float(delta_eps_omega0_sq)) / D
    return (C1, C2, C3)


def _fma(a: float, b: float, c: float) -> float:
    """Correctly rounded a*b + c (Julia's muladd fused on FMA hardware)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def chi_cont_pole(omega0, gamma, S, omega):
    """_chi_cont_pole (Dispersion.jl:21-22): S / (w0^2 - w^2 - i*gamma*w)."""
    return S / (omega0 * omega0 - omega * omega - 1j * gamma * omega)


def discrete_pole_chi(pole, dt: float, eps0: float, omega: float, warm: int = 4000,
                      meas: int = 2000) -> complex:
    """discrete_pole_chi (Dispersion.jl:24-48): drive the discrete ADE recursion with
    cos(w n dt) and fit P to (cos, sin, Nyquist, const)."""
    C1, C2, C3 = pole
    beta = C3
    P = J = Eprev = 0.0
    Pv = np.empty(meas + 1)
    k = 0
    for n in range(0, warm + meas + 1):
        E = np.cos(float(omega) * float(dt) * n)
        psi = _fma(C1, P, _fma(C2, J, beta * Eprev))
        Pnew = psi + beta * E
        Jnew = (Pnew - P) * (2.0 / float(dt)) - J
        P, J, Eprev = Pnew, Jnew, E
        if n >= warm:
            Pv[k] = Pnew
            k += 1
    ns = np.arange(warm, warm + meas + 1)
    t = ns * float(dt)
    nyq = np.where(ns % 2 == 1, -1.0, 1.0)
    M = np.column_stack([np.cos(omega * t), np.sin(omega * t), nyq, np.ones(len(ns))])
    coef, *_ = np.linalg.lstsq(M, Pv, rcond=None)
    return (coef[0] + 1j * coef[1]) / float(eps0)


def fit_pole_strengths(pos, omega: float, tgt: complex):
    """_fit_pole_strengths (Dispersion.jl:50-54): solve the 2xN real system for S."""
    U = np.array([chi_cont_pole(float(p[0]), float(p[1]), 1.0, float(omega)) for p in pos])
    A = np.vstack([U.real, U.imag])
    rhs = np.array([tgt.real, tgt.imag])
    if A.shape[0] == A.shape[1]:
        return np.linalg.solve(A, rhs)
    return np.linalg.lstsq(A, rhs, rcond=None)[0]


def build_pole_set(dt, eps0, omega, target: complex, pos=PROBE_POLE_POSITIONS,
                   require_passive=False, name="pole set"):
    """_build_probe_pole_set (Dispersion.jl:56-71)."""
    S = fit_pole_strengths(pos, float(omega), complex(target))
    poles = [create_pole(float(pos[i][0]), float(pos[i][1]), float(S[i]), float(dt), float(eps0))
             for i in range(len(pos))]
    if require_passive and float(np.min(S)) < 0.0:
        raise ValueError(f"{name} pole fit produced a negative oscillator strength (min S = {np.min(S)})")
    return poles


def build_pump_poles(dt, eps0, gd):
    """build_pump_poles (Dispersion.jl:73-77): fit eps_pump - 1 at omega_pump."""
    epsd = complex(gd["eps_real_pump"], gd["eps_imag_pump"])
    return build_pole_set(dt, eps0, gd["omega_pump"], epsd - 1.0, PROBE_POLE_POSITIONS, False,
                          "pump diagonal")


def build_probe_poles(dt, eps0, gd):
    """build_probe_poles (Dispersion.jl:79-83)."""
    epsd = complex(gd["eps_real_probe"], gd["eps_imag_probe"])
    return build_pole_set(dt, eps0, gd["omega_probe"], epsd - 1.0, PROBE_POLE_POSITIONS, True,
                          "probe diagonal")


def build_probe_mo_poles(dt, eps0, gd):
    """build_probe_mo_poles (Dispersion.jl:85-89): fit -i*eps_probe."""
    epsd = complex(gd["eps_real_probe"], gd["eps_imag_probe"])
    return build_pole_set(dt, eps0, gd["omega_probe"], -1j * epsd, PROBE_MO_POLE_POSITIONS, True,
                          "probe MO off-diagonal")


def poles_array(poles, dtype=np.float32) -> np.ndarray:
    """Pack DLPole{T} triples into the C-ABI layout: npoles x {C1, C2, C3}."""
    return np.asarray([[c for c in p] for p in poles], dtype=np.float64).astype(dtype).reshape(-1)
